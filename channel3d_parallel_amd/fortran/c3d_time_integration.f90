!------------------------------------------------------------------------------!
! RK3 host loop over the gfx950 library, keeping the reference's module and   !
! procedure names and call order:                                            !
!   equations        :: compute_rhs  (src/equations.f90, u, v and w at once) !
!                       compute_rhs_u/v/w on host arrays                     !
!   mass_flow        :: compute_mean_mass_flow_U,                            !
!                       compute_dPx_for_constant_mass_flow (src/mass_flow.f90)!
!   time_integration :: compute_time_step_RK3, compute_dt                    !
!                       (src/time_integration.f90:86-267)                    !
! The velocity stays resident on the GPU; dPdx lives in c3d_flow and is      !
! passed to the library before every right-hand side.                       !
!------------------------------------------------------------------------------!
Module c3d_flow
  Use iso_fortran_env, Only : Int32, Int64
  Implicit None
  Real(Int64)    :: t = 0d0
  Real(Int64)    :: CFL = 0.5d0
  Real(Int64)    :: dPdx = 0d0, dPdy = 0d0, dPdz = 0d0, dPdx_prev = 0d0
  Real(Int64)    :: Qflow_x_0 = 0d0
  Integer(Int32) :: x_mass_cte = 0
  ! Wray RK3 stage times (src/initialization.f90:408-410)
  Real(Int64), Parameter :: rk_t(3) = [8d0/15d0, 2d0/3d0, 1d0]
End Module c3d_flow

!------------------------------------------------------------------------------!
Module equations
  Use c3d_global
  Use c3d_flow
  Implicit None
Contains
  ! compute_rhs_u/v/w(U,V,W,F<slot>) of the resident velocity
  ! (src/equations.f90:24-423).
  Subroutine compute_rhs(slot)
    Integer(Int32), Intent(In) :: slot
    Call c3d_check(c3d_set_pressure_gradient(ctx, dPdx, dPdy, dPdz), 'set_pressure_gradient')
    Call c3d_check(c3d_compute_rhs(ctx, Int(slot, c_int)), 'compute_rhs')
  End Subroutine compute_rhs

  ! compute_rhs_u(U_, V_, W_, rhs_u), src/equations.f90:24-153, of the
  ! caller's arrays; rhs_u has the interior shape (nx-2, nyg-2, nzg-2).
  Subroutine compute_rhs_u(U_, V_, W_, rhs_u)
    Real(Int64), Contiguous, Intent(In)          :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Real(Int64), Contiguous, Intent(Out), Target :: rhs_u(2:, 2:, 2:)
    Call c3d_check(c3d_set_pressure_gradient(ctx, dPdx, dPdy, dPdz), 'set_pressure_gradient')
    Call c3d_check(c3d_compute_rhs_arrays(ctx, U_, V_, W_, c_loc(rhs_u), c_null_ptr, c_null_ptr), &
                   'compute_rhs_u')
  End Subroutine compute_rhs_u

  ! compute_rhs_v(U_, V_, W_, rhs_v), :160-288; rhs_v (nxg-2, ny-2, nzg-2).
  Subroutine compute_rhs_v(U_, V_, W_, rhs_v)
    Real(Int64), Contiguous, Intent(In)          :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Real(Int64), Contiguous, Intent(Out), Target :: rhs_v(2:, 2:, 2:)
    Call c3d_check(c3d_set_pressure_gradient(ctx, dPdx, dPdy, dPdz), 'set_pressure_gradient')
    Call c3d_check(c3d_compute_rhs_arrays(ctx, U_, V_, W_, c_null_ptr, c_loc(rhs_v), c_null_ptr), &
                   'compute_rhs_v')
  End Subroutine compute_rhs_v

  ! compute_rhs_w(U_, V_, W_, rhs_w), :295-423; rhs_w (nxg-2, nyg-2, nz-2).
  Subroutine compute_rhs_w(U_, V_, W_, rhs_w)
    Real(Int64), Contiguous, Intent(In)          :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Real(Int64), Contiguous, Intent(Out), Target :: rhs_w(2:, 2:, 2:)
    Call c3d_check(c3d_set_pressure_gradient(ctx, dPdx, dPdy, dPdz), 'set_pressure_gradient')
    Call c3d_check(c3d_compute_rhs_arrays(ctx, U_, V_, W_, c_null_ptr, c_null_ptr, c_loc(rhs_w)), &
                   'compute_rhs_w')
  End Subroutine compute_rhs_w
End Module equations

!------------------------------------------------------------------------------!
Module mass_flow
  Use c3d_global
  Use c3d_flow
  Implicit None
  ! Each procedure has the reference's form on a host array U(nx,nyg,nzg)
  ! and a resident form without the array argument.
  Interface compute_mean_mass_flow_U
     Module Procedure mean_mass_flow_U_array, mean_mass_flow_U_resident
  End Interface compute_mean_mass_flow_U
  Interface compute_dPx_for_constant_mass_flow
     Module Procedure dPx_constant_mass_flow_array, dPx_constant_mass_flow_resident
  End Interface compute_dPx_for_constant_mass_flow
Contains
  ! compute_mean_mass_flow_U(U, Qflow), src/mass_flow.f90:24-53
  Subroutine mean_mass_flow_U_array(U, Qflow)
    Real(Int64), Contiguous, Intent(In) :: U(:,:,:)
    Real(Int64), Intent(Out) :: Qflow
    Call c3d_check(c3d_mean_mass_flow_u_array(ctx, U, Qflow), 'compute_mean_mass_flow_U')
  End Subroutine mean_mass_flow_U_array

  Subroutine mean_mass_flow_U_resident(Qflow)
    Real(Int64), Intent(Out) :: Qflow
    Call c3d_check(c3d_mean_mass_flow_u(ctx, Qflow), 'compute_mean_mass_flow_U')
  End Subroutine mean_mass_flow_U_resident

  ! compute_dPx_for_constant_mass_flow(U, dPdx), src/mass_flow.f90:144-175:
  ! dPdx = Qflow_x_0 - (H-masked bulk velocity of U)
  Subroutine dPx_constant_mass_flow_array(U, dPdx_out)
    Real(Int64), Contiguous, Intent(In) :: U(:,:,:)
    Real(Int64), Intent(Out) :: dPdx_out
    Real(Int64) :: q
    Call c3d_check(c3d_mean_mass_flow_u_array(ctx, U, q), 'compute_dPx_for_constant_mass_flow')
    dPdx_out = Qflow_x_0 - q
  End Subroutine dPx_constant_mass_flow_array

  Subroutine dPx_constant_mass_flow_resident(dPdx_out)
    Real(Int64), Intent(Out) :: dPdx_out
    Real(Int64) :: q
    Call c3d_check(c3d_mean_mass_flow_u(ctx, q), 'compute_dPx_for_constant_mass_flow')
    dPdx_out = Qflow_x_0 - q
  End Subroutine dPx_constant_mass_flow_resident
End Module mass_flow

!------------------------------------------------------------------------------!
Module input_output
  Use c3d_global
  Use c3d_flow
  Implicit None
  Integer(c_int), Parameter :: RESTORE_ALL = 7  ! P, t and dPdx, fb (C3D_RESTORE_ALL)
Contains
  ! output_data (src/input_output.f90:498-728): the snapshot fname in the
  ! reference's big-endian format, with the current t and dPdx.
  Subroutine output_data(fname)
    Character(*), Intent(In) :: fname
    Call c3d_check(c3d_set_time(ctx, t), 'set_time')
    Call c3d_check(c3d_set_pressure_gradient(ctx, dPdx, dPdy, dPdz), 'set_pressure_gradient')
    Call c3d_check(c3d_write_snapshot(ctx, Trim(fname)//c_null_char, dt, c_null_ptr, c_null_ptr), &
                   'output_data')
  End Subroutine output_data

  ! read_input_data (src/input_output.f90:335-489): U, V, W as the reference
  ! restores them, plus P, t, dPdx and fb (a body set up first) unless
  ! reference_only is present and true.
  Subroutine read_input_data(fname, reference_only)
    Character(*), Intent(In) :: fname
    Logical, Intent(In), Optional :: reference_only
    Integer(c_int) :: what
    Real(c_double) :: dt_file, dy, dz
    what = RESTORE_ALL
    If (.Not. ib_active) what = 3
    If (Present(reference_only)) Then
       If (reference_only) what = 0
    End If
    Call c3d_check(c3d_read_snapshot(ctx, Trim(fname)//c_null_char, what, dt_file), &
                   'read_input_data')
    If (Iand(what, 2) /= 0) Then
       Call c3d_check(c3d_get_time(ctx, t), 'get_time')
       Call c3d_check(c3d_get_pressure_gradient(ctx, dPdx, dy, dz), 'get_pressure_gradient')
    End If
  End Subroutine read_input_data
End Module input_output

!------------------------------------------------------------------------------!
Module immersed_boundary_geometry
  Use c3d_global
  Use c3d_flow
  Implicit None
Contains
  ! setup_IB_geometry + setup_IB_operators + compute_heaviside for the moving
  ! walls (src/immersed_boundary_geometry.f90:176-390): body_type
  ! 'standing_wave' or 'traveling_wave_x', nxb x nzb markers per wall, at
  ! the current time t; everything is built on the device.
  Subroutine setup_IB_geometry(body_type, nxb, nzb, body_param_1, body_param_2, body_param_3, &
                               body_ramp_up_time)
    Character(*), Intent(In)   :: body_type
    Integer(Int32), Intent(In) :: nxb, nzb
    Real(Int64), Intent(In)    :: body_param_1, body_param_2, body_param_3, body_ramp_up_time
    Integer(c_int) :: body
    Select Case (Trim(body_type))
    Case ('standing_wave')
       body = 1
    Case ('traveling_wave_x')
       body = 2
    Case Default
       Write(error_unit,*) 'setup_IB_geometry: unsupported moving body ', Trim(body_type)
       Error Stop 1
    End Select
    Call c3d_check(c3d_ib_setup_wave(ctx, body, Int(nxb, c_int), Int(nzb, c_int), body_param_1, &
                                     body_param_2, body_param_3, body_ramp_up_time, t), &
                   'setup_IB_geometry')
    Call c3d_check(c3d_ib_set_solver(ctx, cg_tol, Int(cg_max_iter, c_int), Int(cg_guess, c_int)), &
                   'ib_set_solver')
    ib_active = .True.
    moving_body = .True.
  End Subroutine setup_IB_geometry

  ! The per-stage update of src/time_integration.f90:110-115 at time t.
  Subroutine move_IB_geometry
    Call c3d_check(c3d_ib_move(ctx, t), 'setup_IB_geometry (move)')
  End Subroutine move_IB_geometry

  ! remove_mean_per_body(f_), src/immersed_boundary_geometry.f90:400-442, in
  ! place on a host vector of nb markers.
  Subroutine remove_mean_per_body(f_)
    Real(Int64), Contiguous, Intent(InOut) :: f_(:)
    Call c3d_check(c3d_ib_remove_mean(ctx, f_), 'remove_mean_per_body')
  End Subroutine remove_mean_per_body
End Module immersed_boundary_geometry

!------------------------------------------------------------------------------!
Module time_integration
  Use c3d_global
  Use c3d_flow
  Use equations
  Use mass_flow
  Use boundary_conditions
  Use projection
  Use immersed_boundary_geometry
  Implicit None
Contains
  ! compute_time_step_RK3, src/time_integration.f90:86-213
  Subroutine compute_time_step_RK3
    Real(Int64) :: to
    Integer(Int32) :: stage
    to = t
    Call c3d_check(c3d_save_state(ctx), 'save_state')                   ! Uo = U
    Do stage = 1, 3
       rk_step = stage
       Call compute_rhs(stage)                                         ! F_stage
       Call c3d_check(c3d_rk3_update(ctx, Int(stage, c_int), dt), 'rk3_update')
       t = to + rk_t(stage)*dt
       If (moving_body) Call move_IB_geometry                          ! :110-115
       Call apply_boundary_conditions_resident
       Call compute_non_IB_projection
       If (ib_active) Then                                             ! :119-122
          Call apply_boundary_conditions_resident
          Call compute_IB_projection
       End If
       Call apply_boundary_conditions_resident
    End Do
    If (x_mass_cte == 1) Then                                          ! :187-193
       dPdx_prev = dPdx
       Call compute_dPx_for_constant_mass_flow(dPdx)
       Call c3d_check(c3d_add_to_u(ctx, dPdx), 'add_to_u')
       dPdx = 0.5d0*(dPdx_prev + dPdx/dt)
       Call apply_boundary_conditions_resident
    End If
  End Subroutine compute_time_step_RK3

  ! compute_dt, src/time_integration.f90:219-267
  Subroutine compute_dt
    Call c3d_check(c3d_compute_dt(ctx, CFL, dt), 'compute_dt')
  End Subroutine compute_dt
End Module time_integration
