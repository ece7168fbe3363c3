!------------------------------------------------------------------------------!
! The reference's host-array procedures called between resident time steps.  !
! Two RK3 steps of the static-body channel (as cylinders_fd); with mode 1,   !
! between the steps the program calls, on its own scratch arrays, what the  !
! reference calls on U_reg, the Heaviside fields and friends:                !
!   apply_boundary_conditions, divergence, compute_rhs_u/v/w,               !
!   compute_mean_mass_flow_U, compute_dPx_for_constant_mass_flow (arrays),  !
!   regT, regT_1n, regTc_1n, regu, regv, regw, remove_mean_per_body,        !
!   compute_heaviside                                                        !
! and checks that the resident velocity is bitwise what it was before.       !
!                                                                            !
! usage: dropin_fd nx ny nz nu dt dPdx mode markers in scratch out           !
!   markers: as cylinders_fd                                                 !
!   in     : raw FP64 U, V, W                                                !
!   scratch: raw FP64 Ua, Va, Wa (field shapes), Pa (rhs_p shape), f (3nb)   !
!   out    : U, V, W after the two steps, fb (4nb), the resident-unchanged   !
!            flag (FP64 1 or 0); mode 1 adds Ua, Va, Wa after                !
!            apply_boundary_conditions, div (rhs_p), Fu, Fv, Fw (interiors), !
!            Qflow, dPdx from the arrays, Eu, E1nu (3nb), E1np (nb),         !
!            regu/regv/regw of f(1:nb), f(nb+1:2nb), f(2nb+1:3nb),          !
!            remove_mean_per_body of f(1:nb)                                 !
!------------------------------------------------------------------------------!
Program dropin_fd
  Use iso_fortran_env, Only : Int32, Int64, error_unit
  Use c3d_global
  Use c3d_flow
  Use boundary_conditions
  Use differential_operators
  Use immersed_boundary_operators
  Use heaviside
  Use equations
  Use mass_flow
  Use time_integration
  Implicit None

  Integer(Int32) :: nxa, nya, nza, mode, iu, nb, nzb, nxb, nxb1
  Real(Int64)    :: nu_in, dt_in, qflow, dpx, same
  Character(256) :: arg, fmk, fin, fsc, fout
  Real(Int64), Allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), U0(:,:,:), V0(:,:,:), W0(:,:,:)
  Real(Int64), Allocatable :: Ua(:,:,:), Va(:,:,:), Wa(:,:,:), Pa(:,:,:), dv(:,:,:)
  Real(Int64), Allocatable :: Fu(:,:,:), Fv(:,:,:), Fw(:,:,:), Ur(:,:,:), Vr(:,:,:), Wr(:,:,:)
  Real(Int64), Allocatable :: xb(:), yb(:), zb(:), normals(:), sb(:), ub(:), fb(:), f(:)
  Real(Int64), Allocatable :: Eu(:), E1nu(:), E1np(:), fm(:)

  If (command_argument_count() /= 11) Then
     Write(error_unit,*) 'usage: dropin_fd nx ny nz nu dt dPdx mode markers in scratch out'
     Error Stop 2
  End If
  Call get_command_argument(1, arg);  Read(arg,*) nxa
  Call get_command_argument(2, arg);  Read(arg,*) nya
  Call get_command_argument(3, arg);  Read(arg,*) nza
  Call get_command_argument(4, arg);  Read(arg,*) nu_in
  Call get_command_argument(5, arg);  Read(arg,*) dt_in
  Call get_command_argument(6, arg);  Read(arg,*) dPdx
  Call get_command_argument(7, arg);  Read(arg,*) mode
  Call get_command_argument(8, fmk)
  Call get_command_argument(9, fin)
  Call get_command_argument(10, fsc)
  Call get_command_argument(11, fout)

  Call c3d_initialize(nxa, nya, nza, 0, 8d0*Atan(1d0), 4d0*Atan(1d0), 2d0, 0d0, nu_in)
  x_mass_cte = 0
  dt = dt_in

  Open(newunit=iu, file=fmk, access='stream', form='unformatted', status='old')
  Read(iu) nb, nzb, nxb, nxb1
  Allocate(xb(nb), yb(nb), zb(nb), normals(3*nb), sb(nb), ub(3*nb), fb(4*nb), f(3*nb))
  Read(iu) xb, yb, zb, normals, sb, ub
  Close(iu)

  Allocate(U(nx, nyg, nzg), V(nxg, ny, nzg), W(nxg, nyg, nz))
  Allocate(U0(nx, nyg, nzg), V0(nxg, ny, nzg), W0(nxg, nyg, nz))
  Allocate(Ua(nx, nyg, nzg), Va(nxg, ny, nzg), Wa(nxg, nyg, nz), Pa(nxg-2, nyg-2, nzg-1))
  Allocate(dv(nxg-2, nyg-2, nzg-1), Ur(nx, nyg, nzg), Vr(nxg, ny, nzg), Wr(nxg, nyg, nz))
  Allocate(Fu(nx-2, nyg-2, nzg-2), Fv(nxg-2, ny-2, nzg-2), Fw(nxg-2, nyg-2, nz-2))
  Allocate(Eu(3*nb), E1nu(3*nb), E1np(nb), fm(nb))
  Open(newunit=iu, file=fin, access='stream', form='unformatted', status='old')
  Read(iu) U, V, W
  Close(iu)
  Open(newunit=iu, file=fsc, access='stream', form='unformatted', status='old')
  Read(iu) Ua, Va, Wa, Pa, f
  Close(iu)

  Call c3d_upload(U, V, W)
  Call apply_boundary_conditions_resident
  Call setup_IB_operators(xb, yb, zb, normals, sb, ub, nzb, nxb, nxb1)

  Call compute_time_step_RK3
  same = 1d0
  If (mode == 1) Then
     Call c3d_download(U0, V0, W0)
     ! src/projection.f90:73,116 and src/heaviside.f90:33,43,49 call these on
     ! scratch fields; src/heaviside.f90:51-52 calls the interpolations
     Call apply_boundary_conditions(Ua, Va, Wa)
     Call divergence(dv, Ua, Va, Wa)
     Call compute_rhs_u(Ua, Va, Wa, Fu)
     Call compute_rhs_v(Ua, Va, Wa, Fv)
     Call compute_rhs_w(Ua, Va, Wa, Fw)
     Call compute_mean_mass_flow_U(Ua, qflow)
     Qflow_x_0 = 1d0
     Call compute_dPx_for_constant_mass_flow(Ua, dpx)
     Call regT(Eu, Ua, Va, Wa)
     Call regT_1n(E1nu, Ua, Va, Wa)
     Call regTc_1n(E1np, Pa)
     Call regu(Ur, f(1:nb))
     Call regv(Vr, f(nb+1:2*nb))
     Call regw(Wr, f(2*nb+1:3*nb))
     fm = f(1:nb)
     Call remove_mean_per_body(fm)
     Call compute_heaviside
     Call c3d_download(U, V, W)
     If (Any(U /= U0) .Or. Any(V /= V0) .Or. Any(W /= W0)) same = 0d0
  End If
  Call compute_time_step_RK3

  Call c3d_download(U, V, W)
  Call c3d_check(c3d_ib_get_force(ctx, fb), 'ib_get_force')
  Open(newunit=iu, file=fout, access='stream', form='unformatted', status='replace')
  Write(iu) U, V, W, fb, same
  If (mode == 1) Write(iu) Ua, Va, Wa, dv, Fu, Fv, Fw, qflow, dpx, Eu, E1nu, E1np, Ur, Vr, Wr, fm
  Close(iu)
  Call c3d_finalize
End Program dropin_fd
