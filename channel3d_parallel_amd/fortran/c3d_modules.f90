!------------------------------------------------------------------------------!
! Fortran host layer over the gfx950 library.  Module and procedure names     !
! follow the reference so its callers can switch by relinking:               !
!   poisson                 :: solve_poisson_equation   (src/poisson.f90:33)   !
!   projection              :: compute_non_IB_projection (src/projection.f90:22)!
!   differential_operators  :: divergence, gradient (src/differential_operators.f90) !
!   boundary_conditions     :: apply_boundary_conditions (src/boundary_conditions.f90:20) !
!   monitor                 :: check_divergence (src/monitor.f90:210)          !
! The reference keeps its state in module `global`; here c3d_global holds    !
! the library context, the sizes and the time-step scalars, and the velocity  !
! stays resident on the GPU between calls.                                   !
!------------------------------------------------------------------------------!
Module c3d_global

  Use, Intrinsic :: iso_c_binding
  Use iso_fortran_env, Only : error_unit, Int32, Int64
  Use c3d_bindings
  Implicit None

  Type(c_ptr)    :: ctx = c_null_ptr
  Integer(Int32) :: nx, ny, nz, nxg, nyg, nzg
  Integer(Int32) :: rk_step = 1
  Real(Int64)    :: dt = 0d0, nu = 0d0

Contains

  Subroutine c3d_check(rc, where)
    Integer(c_int), Intent(In) :: rc
    Character(*), Intent(In) :: where
    If (rc /= C3D_OK) Then
       Write(error_unit,*) 'c3d error in ', where, ' (', rc, '): ', Trim(c3d_error_message(ctx))
       Error Stop 1
    End If
  End Subroutine c3d_check

  ! Counterpart of `initialize` for the projection state
  ! (src/initialization.f90:24-460): grid, Dyy, wavenumbers, resident fields.
  Subroutine c3d_initialize(nx_global, ny_global, nz_global, grid_type, Lxp, Lzp, Ly_channel, &
                            alpha_stretch, nu_in)
    Integer(Int32), Intent(In) :: nx_global, ny_global, nz_global, grid_type
    Real(Int64), Intent(In)    :: Lxp, Lzp, Ly_channel, alpha_stretch, nu_in
    Type(c3d_desc) :: d
    Integer(c_int) :: rc
    d%nx = nx_global; d%ny = ny_global; d%nz = nz_global; d%grid_type = grid_type
    d%Lxp = Lxp; d%Lzp = Lzp; d%Ly = Ly_channel; d%alpha = alpha_stretch; d%nu = nu_in
    d%precision = C3D_PREC_FP64
    rc = c3d_create(d, ctx)
    If (rc /= C3D_OK) Then
       Write(error_unit,*) 'c3d_create failed (', rc, ') for ', nx_global, ny_global, nz_global
       Error Stop 1
    End If
    nx = nx_global; ny = ny_global; nz = nz_global
    nxg = nx + 1; nyg = ny + 1; nzg = nz + 1
    nu = nu_in
  End Subroutine c3d_initialize

  Subroutine c3d_finalize
    Integer(c_int) :: rc
    rc = c3d_destroy(ctx)
    ctx = c_null_ptr
  End Subroutine c3d_finalize

  ! Host -> device copy of the staggered velocity (and optionally P).
  Subroutine c3d_upload(U, V, W, P)
    Real(Int64), Target, Contiguous, Intent(In) :: U(:,:,:), V(:,:,:), W(:,:,:)
    Real(Int64), Target, Contiguous, Intent(In), Optional :: P(:,:,:)
    Type(c_ptr) :: pp
    pp = c_null_ptr
    If (Present(P)) pp = c_loc(P)
    Call c3d_check(c3d_set_fields(ctx, c_loc(U), c_loc(V), c_loc(W), pp), 'c3d_upload')
  End Subroutine c3d_upload

  ! Device -> host copy of U, V, W, P and P_interim (rhs_p shape).
  Subroutine c3d_download(U, V, W, P, P_interim)
    Real(Int64), Target, Contiguous, Intent(InOut) :: U(:,:,:), V(:,:,:), W(:,:,:)
    Real(Int64), Target, Contiguous, Intent(InOut), Optional :: P(:,:,:), P_interim(:,:,:)
    Type(c_ptr) :: pp, pi
    pp = c_null_ptr; pi = c_null_ptr
    If (Present(P)) pp = c_loc(P)
    If (Present(P_interim)) pi = c_loc(P_interim)
    Call c3d_check(c3d_get_fields(ctx, c_loc(U), c_loc(V), c_loc(W), pp, pi), 'c3d_download')
  End Subroutine c3d_download

End Module c3d_global

!------------------------------------------------------------------------------!
Module boundary_conditions
  Use c3d_global
  Implicit None
Contains
  ! apply_boundary_conditions(U_,V_,W_), src/boundary_conditions.f90:20-55:
  ! uploads the arrays, applies the conditions on the GPU, copies them back.
  ! The arrays become the resident velocity.
  Subroutine apply_boundary_conditions(U_, V_, W_)
    Real(Int64), Contiguous, Intent(InOut) :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Call c3d_upload(U_, V_, W_)
    Call c3d_check(c3d_apply_boundary_conditions(ctx), 'apply_boundary_conditions')
    Call c3d_download(U_, V_, W_)
  End Subroutine apply_boundary_conditions

  ! Same on the resident velocity, no host copies.
  Subroutine apply_boundary_conditions_resident
    Call c3d_check(c3d_apply_boundary_conditions(ctx), 'apply_boundary_conditions')
  End Subroutine apply_boundary_conditions_resident
End Module boundary_conditions

!------------------------------------------------------------------------------!
Module differential_operators
  Use c3d_global
  Implicit None
Contains
  ! divergence(div_, U_, V_, W_), src/differential_operators.f90:54-74.
  ! The velocity arrays become the resident velocity.
  Subroutine divergence(div_, U_, V_, W_)
    Real(Int64), Contiguous, Intent(Out) :: div_(2:, 2:, 2:)
    Real(Int64), Contiguous, Intent(In)  :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Call c3d_upload(U_, V_, W_)
    Call c3d_check(c3d_divergence(ctx, div_), 'divergence')
  End Subroutine divergence

  ! gradient(Gpx_, Gpy_, Gpz_, p_), src/differential_operators.f90:16-52.
  Subroutine gradient(Gpx_, Gpy_, Gpz_, p_)
    Real(Int64), Contiguous, Intent(In)  :: p_(2:, 2:, 2:)
    Real(Int64), Contiguous, Intent(Out) :: Gpx_(:,:,:), Gpy_(:,:,:), Gpz_(:,:,:)
    Call c3d_check(c3d_gradient(ctx, p_, Gpx_, Gpy_, Gpz_), 'gradient')
  End Subroutine gradient
End Module differential_operators

!------------------------------------------------------------------------------!
Module poisson
  Use c3d_global
  Implicit None
Contains
  ! solve_poisson_equation(rhs), src/poisson.f90:33-101; uses rk_step and dt
  ! of c3d_global as the reference uses those of `global`.
  Subroutine solve_poisson_equation(rhs)
    Real(Int64), Dimension(2:nxg-1, 2:nyg-1, 2:nzg), Intent(InOut) :: rhs
    Call c3d_check(c3d_solve_poisson_equation(ctx, rhs, Int(rk_step, c_int), dt), &
                   'solve_poisson_equation')
  End Subroutine solve_poisson_equation
End Module poisson

!------------------------------------------------------------------------------!
Module projection
  Use c3d_global
  Implicit None
Contains
  ! compute_non_IB_projection, src/projection.f90:22-46, on the resident
  ! velocity (upload with c3d_upload, read back with c3d_download).
  Subroutine compute_non_IB_projection
    Call c3d_check(c3d_compute_non_ib_projection(ctx, dt, Int(rk_step, c_int)), &
                   'compute_non_IB_projection')
  End Subroutine compute_non_IB_projection
End Module projection

!------------------------------------------------------------------------------!
Module monitor
  Use c3d_global
  Implicit None
Contains
  ! check_divergence(max_divergence), src/monitor.f90:210-221 (resident velocity).
  Subroutine check_divergence(max_divergence)
    Real(Int64), Intent(Out) :: max_divergence
    Call c3d_check(c3d_check_divergence(ctx, max_divergence), 'check_divergence')
  End Subroutine check_divergence
End Module monitor
