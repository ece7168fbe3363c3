!------------------------------------------------------------------------------!
! Fortran host driver: one projection step on the GPU through the modules    !
! that keep the reference's names, in the call order of                      !
! src/time_integration.f90:117-123 (BC, projection, BC).                     !
!                                                                            !
! usage: projection_driver nx ny nz grid_type alpha dt rk_step in out        !
!                          [precision nref]                                  !
!   in : raw FP64 U, V, W in the reference's Fortran layout                  !
!   out: raw FP64 U, V, W, P, P_interim, then max|div|                       !
!   precision: 0 FP64 (default), 1 mixed; nref >= 0: refinement sweeps       !
!------------------------------------------------------------------------------!
Program projection_driver
  Use iso_fortran_env, Only : Int32, Int64, error_unit
  Use c3d_global
  Use boundary_conditions
  Use projection
  Use monitor
  Implicit None

  Integer(Int32) :: nxa, nya, nza, gtype, rk, iu, prec, nref
  Real(Int64)    :: alpha, dt_in, max_div
  Character(256) :: arg, fin, fout
  Real(Int64), Allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), P(:,:,:), Pint(:,:,:)

  If (command_argument_count() /= 9 .And. command_argument_count() /= 11) Then
     Write(error_unit,*) 'usage: projection_driver nx ny nz grid_type alpha dt rk_step in out ', &
          '[precision nref]'
     Error Stop 2
  End If
  prec = C3D_PREC_FP64
  nref = -1
  If (command_argument_count() == 11) Then
     Call get_command_argument(10, arg); Read(arg,*) prec
     Call get_command_argument(11, arg); Read(arg,*) nref
  End If
  Call get_command_argument(1, arg); Read(arg,*) nxa
  Call get_command_argument(2, arg); Read(arg,*) nya
  Call get_command_argument(3, arg); Read(arg,*) nza
  Call get_command_argument(4, arg); Read(arg,*) gtype
  Call get_command_argument(5, arg); Read(arg,*) alpha
  Call get_command_argument(6, arg); Read(arg,*) dt_in
  Call get_command_argument(7, arg); Read(arg,*) rk
  Call get_command_argument(8, fin)
  Call get_command_argument(9, fout)

  Call c3d_initialize(nxa, nya, nza, gtype, 8d0*Atan(1d0), 4d0*Atan(1d0), 2d0, alpha, 1d0, &
                      prec=prec)
  If (nref >= 0) Call c3d_check(c3d_set_refinement(ctx, nref), 'c3d_set_refinement')
  dt = dt_in
  rk_step = rk

  Allocate(U(nx, nyg, nzg), V(nxg, ny, nzg), W(nxg, nyg, nz))
  Allocate(P(nxg, nyg, nzg), Pint(2:nxg-1, 2:nyg-1, 2:nzg))
  Open(newunit=iu, file=fin, access='stream', form='unformatted', status='old')
  Read(iu) U, V, W
  Close(iu)
  P = 0d0

  Call c3d_upload(U, V, W, P)
  Call apply_boundary_conditions_resident
  Call compute_non_IB_projection
  Call c3d_download(U, V, W, P, Pint)
  Call apply_boundary_conditions_resident
  Call check_divergence(max_div)

  Open(newunit=iu, file=fout, access='stream', form='unformatted', status='replace')
  Write(iu) U, V, W, P, Pint, max_div
  Close(iu)
  Call c3d_finalize
End Program projection_driver
