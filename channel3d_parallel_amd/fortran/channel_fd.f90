!------------------------------------------------------------------------------!
! Channel-flow driver on the GPU: the time loop of app/main.f90:102-119      !
! (compute_dt, compute_time_step_RK3, monitor) with the initial mass flow    !
! of app/main.f90:90 and constant x mass flow (x_mass_cte = 1).             !
!                                                                            !
! usage: channel_fd nx ny nz grid_type alpha nu cfl dPdx nsteps nmonitor in out [snapshot]
!   in : raw FP64 U, V, W (reference Fortran layout, boundary conditions     !
!        applied by the driver), or snapshot:<file> to restart from a       !
!        snapshot with read_input_data (U, V, W, P, t, dPdx restored)        !
!   out: raw FP64 U, V, W, P, then dPdx, t, max|div|, utau                  !
!   snapshot (optional): the final state written by output_data             !
!        (src/input_output.f90:498-728 format)                               !
!------------------------------------------------------------------------------!
Program channel_fd
  Use iso_fortran_env, Only : Int32, Int64, error_unit, output_unit
  Use c3d_global
  Use c3d_flow
  Use boundary_conditions
  Use mass_flow
  Use time_integration
  Use monitor
  Use statistics
  Use input_output
  Implicit None

  Integer(Int32) :: nxa, nya, nza, gtype, nsteps, nmonitor, istep, iu
  Real(Int64)    :: alpha, nu_in, max_div, q, utau, wtau
  Character(256) :: arg, fin, fout, fsnap
  Real(Int64), Allocatable :: U(:,:,:), V(:,:,:), W(:,:,:), P(:,:,:)

  If (command_argument_count() /= 12 .And. command_argument_count() /= 13) Then
     Write(error_unit,*) 'usage: channel_fd nx ny nz grid_type alpha nu cfl dPdx nsteps nmonitor ', &
          'in out [snapshot]'
     Error Stop 2
  End If
  Call get_command_argument(1, arg);  Read(arg,*) nxa
  Call get_command_argument(2, arg);  Read(arg,*) nya
  Call get_command_argument(3, arg);  Read(arg,*) nza
  Call get_command_argument(4, arg);  Read(arg,*) gtype
  Call get_command_argument(5, arg);  Read(arg,*) alpha
  Call get_command_argument(6, arg);  Read(arg,*) nu_in
  Call get_command_argument(7, arg);  Read(arg,*) CFL
  Call get_command_argument(8, arg);  Read(arg,*) dPdx
  Call get_command_argument(9, arg);  Read(arg,*) nsteps
  Call get_command_argument(10, arg); Read(arg,*) nmonitor
  Call get_command_argument(11, fin)
  Call get_command_argument(12, fout)
  fsnap = ''
  If (command_argument_count() == 13) Call get_command_argument(13, fsnap)

  Call c3d_initialize(nxa, nya, nza, gtype, 8d0*Atan(1d0), 4d0*Atan(1d0), 2d0, alpha, nu_in)
  x_mass_cte = 1

  Allocate(U(nx, nyg, nzg), V(nxg, ny, nzg), W(nxg, nyg, nz), P(nxg, nyg, nzg))
  If (fin(1:9) == 'snapshot:') Then
     ! restart (init_type = 0): the flow rate to hold is the snapshot's
     Call read_input_data(Trim(fin(10:)))
  Else
     Open(newunit=iu, file=fin, access='stream', form='unformatted', status='old')
     Read(iu) U, V, W
     Close(iu)
     P = 0d0
     Call c3d_upload(U, V, W, P)
     Call apply_boundary_conditions_resident
  End If

  ! initial mass flow (app/main.f90:90)
  Call compute_mean_mass_flow_U(Qflow_x_0)

  Do istep = 1, nsteps
     Call compute_dt
     Call compute_time_step_RK3
     If (nmonitor > 0) Then
        If (Mod(istep, nmonitor) == 0) Then
           Call check_divergence(max_div)
           Call compute_mean_mass_flow_U(q)
           Write(output_unit,'(A,I6,A,ES12.5,A,ES12.5,A,ES10.3,A,ES12.5)') 'step ', istep, &
                '  t ', t, '  dt ', dt, '  max|div| ', max_div, '  dPdx ', dPdx
        End If
     End If
  End Do

  Call check_divergence(max_div)
  Call compute_wall_shear(utau, wtau)
  Call c3d_download(U, V, W, P)
  Open(newunit=iu, file=fout, access='stream', form='unformatted', status='replace')
  Write(iu) U, V, W, P, dPdx, t, max_div, utau
  Close(iu)
  If (Len_Trim(fsnap) > 0) Call output_data(fsnap)
  Call c3d_finalize
End Program channel_fd
