!------------------------------------------------------------------------------!
! Channel with static immersed bodies on the GPU: the time loop of           !
! app/main.f90:78-119 with setup_IB_operators + compute_heaviside at start,  !
! a fixed time step and a constant driving pressure gradient applied         !
! everywhere (exterior_pressure_gradient off, no constant-flow correction).  !
!                                                                            !
! usage: cylinders_fd nx ny nz nu dt dPdx nsteps nmonitor cg_guess markers in out
!   markers: int32 nb nzb nxb nxb1, then FP64 xb yb zb (nb), normals (3nb),  !
!            sb (nb), ub (3nb)                                               !
!   in     : raw FP64 U, V, W (reference layout, uniform y grid)             !
!   out    : raw FP64 U, V, W, fb (4nb), then max|div|, max slip and the     !
!            summed BiCGSTAB iterations (FP64)                               !
!------------------------------------------------------------------------------!
Program cylinders_fd
  Use iso_fortran_env, Only : Int32, Int64, error_unit, output_unit
  Use c3d_global
  Use c3d_flow
  Use boundary_conditions
  Use immersed_boundary_operators
  Use time_integration
  Use monitor
  Implicit None

  Integer(Int32) :: nxa, nya, nza, nsteps, nmonitor, istep, iu, nb, nzb, nxb, nxb1
  Real(Int64)    :: nu_in, dt_in, max_div, max_slip
  Character(256) :: arg, fmk, fin, fout
  Real(Int64), Allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), xb(:), yb(:), zb(:), normals(:), &
                              sb(:), ub(:), fb(:)

  If (command_argument_count() /= 12) Then
     Write(error_unit,*) 'usage: cylinders_fd nx ny nz nu dt dPdx nsteps nmonitor cg_guess markers in out'
     Error Stop 2
  End If
  Call get_command_argument(1, arg);  Read(arg,*) nxa
  Call get_command_argument(2, arg);  Read(arg,*) nya
  Call get_command_argument(3, arg);  Read(arg,*) nza
  Call get_command_argument(4, arg);  Read(arg,*) nu_in
  Call get_command_argument(5, arg);  Read(arg,*) dt_in
  Call get_command_argument(6, arg);  Read(arg,*) dPdx
  Call get_command_argument(7, arg);  Read(arg,*) nsteps
  Call get_command_argument(8, arg);  Read(arg,*) nmonitor
  Call get_command_argument(9, arg);  Read(arg,*) cg_guess
  Call get_command_argument(10, fmk)
  Call get_command_argument(11, fin)
  Call get_command_argument(12, fout)

  Call c3d_initialize(nxa, nya, nza, 0, 8d0*Atan(1d0), 4d0*Atan(1d0), 2d0, 0d0, nu_in)
  x_mass_cte = 0
  dt = dt_in

  Open(newunit=iu, file=fmk, access='stream', form='unformatted', status='old')
  Read(iu) nb, nzb, nxb, nxb1
  Allocate(xb(nb), yb(nb), zb(nb), normals(3*nb), sb(nb), ub(3*nb), fb(4*nb))
  Read(iu) xb, yb, zb, normals, sb, ub
  Close(iu)

  Allocate(U(nx, nyg, nzg), V(nxg, ny, nzg), W(nxg, nyg, nz))
  Open(newunit=iu, file=fin, access='stream', form='unformatted', status='old')
  Read(iu) U, V, W
  Close(iu)
  Call c3d_upload(U, V, W)
  Call apply_boundary_conditions_resident

  ! app/main.f90:78-87
  Call setup_IB_operators(xb, yb, zb, normals, sb, ub, nzb, nxb, nxb1)

  Do istep = 1, nsteps
     Call compute_time_step_RK3
     If (nmonitor > 0) Then
        If (Mod(istep, nmonitor) == 0) Then
           Call check_divergence(max_div)
           Call check_slip(max_slip)
           Write(output_unit,'(A,I6,A,ES12.5,A,ES10.3,A,ES10.3,A,I5,A,I9)') 'step ', istep, &
                '  t ', t, '  max|div| ', max_div, '  max slip ', max_slip, &
                '  cg iters ', cg_last_iter, '  total ', cg_accum_iter
        End If
     End If
  End Do

  Call check_divergence(max_div)
  Call check_slip(max_slip)
  Call c3d_download(U, V, W)
  Call c3d_check(c3d_ib_get_force(ctx, fb), 'ib_get_force')
  Open(newunit=iu, file=fout, access='stream', form='unformatted', status='replace')
  Write(iu) U, V, W, fb, max_div, max_slip, Real(cg_accum_iter, Int64)
  Close(iu)
  Call c3d_finalize
End Program cylinders_fd
