!------------------------------------------------------------------------------!
! Channel between moving walls on the GPU: the time loop of app/main.f90:78-119 !
! with a standing or traveling wave body (setup_IB_geometry,                  !
! src/immersed_boundary_geometry.f90:176-390) re-set-up at every Runge-Kutta   !
! stage, on the buffered tanh grid (grid_type 2), a fixed time step and a      !
! constant driving pressure gradient.                                          !
!                                                                            !
! usage: waves_fd nx ny nz alpha min_buffer_width nu dt dPdx nsteps body_type   !
!                 nxb nzb body_param_1 body_param_2 body_param_3 ramp t0 in out !
!   in  : raw FP64 U, V, W (reference layout)                                 !
!   out : raw FP64 U, V, W, fb (4nb), yb (nb), then t, max|div|, max slip and  !
!         the summed BiCGSTAB iterations (FP64)                               !
!------------------------------------------------------------------------------!
Program waves_fd
  Use iso_fortran_env, Only : Int32, Int64, error_unit, output_unit
  Use c3d_global
  Use c3d_flow
  Use boundary_conditions
  Use immersed_boundary_geometry
  Use time_integration
  Use monitor
  Implicit None

  Integer(Int32) :: nxa, nya, nza, nsteps, nxb, nzb, istep, iu, nb
  Real(Int64)    :: alpha, mbw, nu_in, dt_in, p1, p2, p3, ramp, t0, max_div, max_slip
  Character(256) :: arg, body, fin, fout
  Real(Int64), Allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), fb(:), xb(:), yb(:), zb(:), &
                              normals(:), sb(:), ub(:)
  Integer(c_int), Allocatable :: yref(:)

  If (command_argument_count() /= 19) Then
     Write(error_unit,*) 'usage: waves_fd nx ny nz alpha min_buffer_width nu dt dPdx nsteps ', &
          'body_type nxb nzb body_param_1 body_param_2 body_param_3 ramp t0 in out'
     Error Stop 2
  End If
  Call get_command_argument(1, arg);  Read(arg,*) nxa
  Call get_command_argument(2, arg);  Read(arg,*) nya
  Call get_command_argument(3, arg);  Read(arg,*) nza
  Call get_command_argument(4, arg);  Read(arg,*) alpha
  Call get_command_argument(5, arg);  Read(arg,*) mbw
  Call get_command_argument(6, arg);  Read(arg,*) nu_in
  Call get_command_argument(7, arg);  Read(arg,*) dt_in
  Call get_command_argument(8, arg);  Read(arg,*) dPdx
  Call get_command_argument(9, arg);  Read(arg,*) nsteps
  Call get_command_argument(10, body)
  Call get_command_argument(11, arg); Read(arg,*) nxb
  Call get_command_argument(12, arg); Read(arg,*) nzb
  Call get_command_argument(13, arg); Read(arg,*) p1
  Call get_command_argument(14, arg); Read(arg,*) p2
  Call get_command_argument(15, arg); Read(arg,*) p3
  Call get_command_argument(16, arg); Read(arg,*) ramp
  Call get_command_argument(17, arg); Read(arg,*) t0
  Call get_command_argument(18, fin)
  Call get_command_argument(19, fout)

  Call c3d_initialize(nxa, nya, nza, 2, 8d0*Atan(1d0), 4d0*Atan(1d0), 2d0, alpha, nu_in, mbw)
  x_mass_cte = 0
  dt = dt_in
  t = t0

  Allocate(U(nx, nyg, nzg), V(nxg, ny, nzg), W(nxg, nyg, nz))
  Open(newunit=iu, file=fin, access='stream', form='unformatted', status='old')
  Read(iu) U, V, W
  Close(iu)
  Call c3d_upload(U, V, W)
  Call apply_boundary_conditions_resident

  Call setup_IB_geometry(body, nxb, nzb, p1, p2, p3, ramp)

  Do istep = 1, nsteps
     Call compute_time_step_RK3
     Call check_divergence(max_div)
     Call check_slip(max_slip)
     Write(output_unit,'(A,I6,A,ES12.5,A,ES10.3,A,ES10.3,A,I5)') 'step ', istep, '  t ', t, &
          '  max|div| ', max_div, '  max slip ', max_slip, '  cg iters ', cg_last_iter
  End Do

  nb = 2 * nxb * nzb
  Allocate(fb(4*nb), xb(nb), yb(nb), zb(nb), normals(3*nb), sb(nb), ub(3*nb), yref(nb))
  Call check_divergence(max_div)
  Call check_slip(max_slip)
  Call c3d_download(U, V, W)
  Call c3d_check(c3d_ib_get_force(ctx, fb), 'ib_get_force')
  Call c3d_check(c3d_ib_get_markers(ctx, xb, yb, zb, normals, sb, ub, yref), 'ib_get_markers')
  Open(newunit=iu, file=fout, access='stream', form='unformatted', status='replace')
  Write(iu) U, V, W, fb, yb, t, max_div, max_slip, Real(cg_accum_iter, Int64)
  Close(iu)
  Call c3d_finalize
End Program waves_fd
