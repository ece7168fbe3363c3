!------------------------------------------------------------------------------!
! Fortran host layer over the gfx950 library.  Module and procedure names     !
! follow the reference so its callers can switch by relinking:               !
!   poisson                 :: solve_poisson_equation   (src/poisson.f90:33)   !
!   projection              :: compute_non_IB_projection (src/projection.f90:22)!
!   differential_operators  :: divergence, gradient (src/differential_operators.f90) !
!   boundary_conditions     :: apply_boundary_conditions (src/boundary_conditions.f90:20) !
!   immersed_boundary_operators :: regT, regT_1n, regTc_1n, regu, regv, regw   !
!                              (src/immersed_boundary_operators.f90:296-400)   !
!   heaviside               :: compute_heaviside (src/heaviside.f90:20)        !
!   monitor                 :: check_divergence, check_slip, monitor_fields     !
!                              (src/monitor.f90:27-59, 210-231)                 !
!   statistics              :: compute_wall_shear, compute_plane_statistics     !
!                              (src/statistics.f90:41-131)                     !
!   projection              :: compute_IB_projection, schur, bicgstab          !
!                              (src/projection.f90:48-186)                     !
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
  ! immersed boundary: a body is set up (body_type /= 'none'), BiCGSTAB
  ! settings (src/global.f90, namelist &params) and iteration counters
  Logical        :: ib_active = .False.
  Logical        :: moving_body = .False.  ! wall body re-set-up every stage
  Real(Int64)    :: cg_tol = 1d-10
  Integer(Int32) :: cg_max_iter = 1000
  Integer(Int32) :: cg_guess = 0   ! 0 previous force (reference), 1 zero
  Integer(Int32) :: cg_last_iter = 0
  Integer(Int64) :: cg_accum_iter = 0

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
  ! prec (optional): C3D_PREC_FP64 (default, the reference's arithmetic)
  ! or C3D_PREC_MIXED.
  Subroutine c3d_initialize(nx_global, ny_global, nz_global, grid_type, Lxp, Lzp, Ly_channel, &
                            alpha_stretch, nu_in, min_buffer_width, prec)
    Integer(Int32), Intent(In) :: nx_global, ny_global, nz_global, grid_type
    Real(Int64), Intent(In)    :: Lxp, Lzp, Ly_channel, alpha_stretch, nu_in
    Real(Int64), Intent(In), Optional :: min_buffer_width  ! grid_type 2 only
    Integer(Int32), Intent(In), Optional :: prec
    Type(c3d_desc) :: d
    Integer(c_int) :: rc
    d%nx = nx_global; d%ny = ny_global; d%nz = nz_global; d%grid_type = grid_type
    d%Lxp = Lxp; d%Lzp = Lzp; d%Ly = Ly_channel; d%alpha = alpha_stretch; d%nu = nu_in
    d%precision = C3D_PREC_FP64
    If (Present(prec)) d%precision = prec
    d%min_buffer_width = 0d0
    If (Present(min_buffer_width)) d%min_buffer_width = min_buffer_width
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
  ! apply_boundary_conditions(U_,V_,W_), src/boundary_conditions.f90:20-55,
  ! on the caller's arrays (the reference applies it to U, V, W, U_reg and
  ! the Heaviside fields alike): staged in the context's device scratch, the
  ! resident velocity is not touched.
  Subroutine apply_boundary_conditions(U_, V_, W_)
    Real(Int64), Contiguous, Intent(InOut) :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Call c3d_check(c3d_apply_boundary_conditions_arrays(ctx, U_, V_, W_), &
                   'apply_boundary_conditions')
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
  ! divergence(div_, U_, V_, W_), src/differential_operators.f90:54-74, of
  ! the caller's arrays (device scratch; the resident velocity is untouched).
  Subroutine divergence(div_, U_, V_, W_)
    Real(Int64), Contiguous, Intent(Out) :: div_(2:, 2:, 2:)
    Real(Int64), Contiguous, Intent(In)  :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Call c3d_check(c3d_divergence_arrays(ctx, U_, V_, W_, div_), 'divergence')
  End Subroutine divergence

  ! divergence of the resident velocity (no host copies of the velocity).
  Subroutine divergence_resident(div_)
    Real(Int64), Contiguous, Intent(Out) :: div_(2:, 2:, 2:)
    Call c3d_check(c3d_divergence(ctx, div_), 'divergence')
  End Subroutine divergence_resident

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

  ! compute_IB_projection, src/projection.f90:48-96, on the resident state.
  ! Non-convergence is a warning as in the reference (:182-185).
  Subroutine compute_IB_projection
    Integer(c_int) :: rc, it
    Real(c_double) :: err
    Integer(c_long_long) :: acc
    rc = c3d_compute_ib_projection(ctx, dt, Int(rk_step, c_int), it, err)
    If (rc == -4) Then
       Write(error_unit,'(A,I5,A,ES10.3,A,ES10.3)') 'WARNING: BiCGSTAB reached max iterations (', &
            it, ') but residual norm (', Sqrt(err), ') exceeds requested tolerance ', cg_tol
    Else
       Call c3d_check(rc, 'compute_IB_projection')
    End If
    Call c3d_check(c3d_ib_stats(ctx, it, acc, err), 'ib_stats')
    cg_last_iter = it
    cg_accum_iter = acc
  End Subroutine compute_IB_projection

  ! schur(Sf_, f_), src/projection.f90:98-144 (host arrays).
  Subroutine schur(Sf_, f_)
    Real(Int64), Contiguous, Intent(In)  :: f_(:)
    Real(Int64), Contiguous, Intent(Out) :: Sf_(:)
    Call c3d_check(c3d_ib_schur(ctx, f_, Sf_, dt, Int(rk_step, c_int)), 'schur')
  End Subroutine schur

  ! bicgstab(bcg_x, bcg_b), src/projection.f90:146-186 (host arrays).
  Subroutine bicgstab(bcg_x, bcg_b)
    Real(Int64), Contiguous, Intent(InOut) :: bcg_x(:)
    Real(Int64), Contiguous, Intent(In)    :: bcg_b(:)
    Integer(c_int) :: rc, it
    Real(c_double) :: err
    rc = c3d_ib_bicgstab(ctx, bcg_x, bcg_b, cg_tol, Int(cg_max_iter, c_int), dt, &
                         Int(rk_step, c_int), it, err)
    If (rc /= -4) Call c3d_check(rc, 'bicgstab')
    cg_accum_iter = cg_accum_iter + it
  End Subroutine bicgstab
End Module projection

!------------------------------------------------------------------------------!
Module immersed_boundary_operators
  Use c3d_global
  Implicit None
Contains
  ! setup_IB_operators (src/immersed_boundary_operators.f90:19-256) with the
  ! Heaviside fields of compute_heaviside (src/heaviside.f90:20-54) for a
  ! static body; markers as in src/global.f90 (normals, ub: 3*nb).
  Subroutine setup_IB_operators(xb, yb, zb, normals, sb, ub, nzb, nxb, nxb1)
    Real(Int64), Contiguous, Intent(In) :: xb(:), yb(:), zb(:), normals(:), sb(:), ub(:)
    Integer(Int32), Intent(In) :: nzb, nxb, nxb1
    Call c3d_check(c3d_ib_setup(ctx, Int(Size(xb), c_int), xb, yb, zb, normals, sb, ub, &
                                Int(nzb, c_int), Int(nxb, c_int), Int(nxb1, c_int)), &
                   'setup_IB_operators')
    Call c3d_check(c3d_ib_set_solver(ctx, cg_tol, Int(cg_max_iter, c_int), Int(cg_guess, c_int)), &
                   'ib_set_solver')
    ib_active = .True.
  End Subroutine setup_IB_operators

  ! regT(Eu_, U_, V_, W_), src/immersed_boundary_operators.f90:296-318:
  ! E of the caller's arrays, Eu_(3*nb) component-major.
  Subroutine regT(Eu_, U_, V_, W_)
    Real(Int64), Contiguous, Intent(In)  :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Real(Int64), Contiguous, Intent(Out) :: Eu_(:)
    Call c3d_check(c3d_ib_interpolate_arrays(ctx, 0_c_int, U_, V_, W_, Eu_), 'regT')
  End Subroutine regT

  ! regT_1n(E1nu_, U_, V_, W_), :320-342 (the normal-derivative weights).
  Subroutine regT_1n(E1nu_, U_, V_, W_)
    Real(Int64), Contiguous, Intent(In)  :: U_(:,:,:), V_(:,:,:), W_(:,:,:)
    Real(Int64), Contiguous, Intent(Out) :: E1nu_(:)
    Call c3d_check(c3d_ib_interpolate_arrays(ctx, 1_c_int, U_, V_, W_, E1nu_), 'regT_1n')
  End Subroutine regT_1n

  ! regTc_1n(E1np_, P_), :344-357, of a centre field with rhs_p bounds.
  Subroutine regTc_1n(E1np_, P_)
    Real(Int64), Contiguous, Intent(In)  :: P_(2:, 2:, 2:)
    Real(Int64), Contiguous, Intent(Out) :: E1np_(:)
    Call c3d_check(c3d_ib_interpolate_c(ctx, P_, E1np_), 'regTc_1n')
  End Subroutine regTc_1n

  ! regu / regv / regw(F_, f_), :359-400: R f_ into the whole field F_
  ! (zero off the support, as the reference zeroes it first).
  Subroutine regu(U_, f_)
    Real(Int64), Contiguous, Intent(In)  :: f_(:)
    Real(Int64), Contiguous, Intent(Out) :: U_(:,:,:)
    Call c3d_check(c3d_ib_spread(ctx, 0_c_int, f_, U_), 'regu')
  End Subroutine regu

  Subroutine regv(V_, f_)
    Real(Int64), Contiguous, Intent(In)  :: f_(:)
    Real(Int64), Contiguous, Intent(Out) :: V_(:,:,:)
    Call c3d_check(c3d_ib_spread(ctx, 1_c_int, f_, V_), 'regv')
  End Subroutine regv

  Subroutine regw(W_, f_)
    Real(Int64), Contiguous, Intent(In)  :: f_(:)
    Real(Int64), Contiguous, Intent(Out) :: W_(:,:,:)
    Call c3d_check(c3d_ib_spread(ctx, 2_c_int, f_, W_), 'regw')
  End Subroutine regw
End Module immersed_boundary_operators

!------------------------------------------------------------------------------!
Module heaviside
  Use c3d_global
  Implicit None
Contains
  ! compute_heaviside, src/heaviside.f90:20-54, for the current markers: the
  ! interior masks, Hc_exterior and E1nHc_exterior / E1nH_exterior stay on
  ! the device (setup_IB_operators has already run it once).
  Subroutine compute_heaviside
    Call c3d_check(c3d_ib_compute_heaviside(ctx), 'compute_heaviside')
  End Subroutine compute_heaviside
End Module heaviside

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

  ! check_slip(max_slip), src/monitor.f90:223-231 (resident state).
  Subroutine check_slip(max_slip)
    Real(Int64), Intent(Out) :: max_slip
    Call c3d_check(c3d_check_slip(ctx, max_slip), 'check_slip')
  End Subroutine check_slip

  ! Field summaries of output_monitor, src/monitor.f90:27-39, 57-59 (resident
  ! velocity): maxima and "Mean U/V/W" as the reference prints them, and the
  ! number of non-finite entries (a NaN guard for the time loop).
  Subroutine monitor_fields(maxU, maxV, maxW, meanU, meanV, meanW, nonfinite)
    Real(Int64), Intent(Out) :: maxU, maxV, maxW, meanU, meanV, meanW
    Integer(Int64), Intent(Out) :: nonfinite
    Real(Int64) :: r(7)
    Call c3d_check(c3d_monitor(ctx, r), 'monitor_fields')
    maxU = r(1); maxV = r(2); maxW = r(3)
    meanU = r(4); meanV = r(5); meanW = r(6)
    nonfinite = Int(r(7), Int64)
  End Subroutine monitor_fields
End Module monitor

!------------------------------------------------------------------------------!
Module statistics
  Use c3d_global
  Implicit None
Contains
  ! Wall shear of compute_statistics, src/statistics.f90:41-131 (resident
  ! velocity): friction velocities utau, wtau and, optionally, the plane-mean
  ! profiles Umean(1:nyg), Wmean(1:nyg).
  Subroutine compute_wall_shear(utau, wtau, Umean, Wmean)
    Real(Int64), Intent(Out) :: utau, wtau
    Real(Int64), Intent(Out), Optional, Target :: Umean(:), Wmean(:)
    Real(Int64), Allocatable, Target :: um(:), wm(:)
    Allocate(um(nyg), wm(nyg))
    Call c3d_check(c3d_wall_shear(ctx, utau, wtau, um, wm), 'compute_wall_shear')
    If (Present(Umean)) Umean = um
    If (Present(Wmean)) Wmean = wm
    Deallocate(um, wm)
  End Subroutine compute_wall_shear

  ! Plane-mean profiles of compute_statistics, src/statistics.f90:41-111
  ! (resident velocity); each argument has nyg entries.
  Subroutine compute_plane_statistics(Umean, Vmean, Wmean, U2mean, V2mean, W2mean, UVmean)
    Real(Int64), Intent(Out) :: Umean(:), Vmean(:), Wmean(:), U2mean(:), V2mean(:), W2mean(:), &
                                UVmean(:)
    Real(Int64), Allocatable :: m(:)
    Allocate(m(7 * nyg))
    Call c3d_check(c3d_plane_statistics(ctx, m), 'compute_plane_statistics')
    Umean  = m(        1:    nyg)
    Vmean  = m(  nyg + 1:2 * nyg)
    Wmean  = m(2*nyg + 1:3 * nyg)
    U2mean = m(3*nyg + 1:4 * nyg)
    V2mean = m(4*nyg + 1:5 * nyg)
    W2mean = m(5*nyg + 1:6 * nyg)
    UVmean = m(6*nyg + 1:7 * nyg)
    Deallocate(m)
  End Subroutine compute_plane_statistics
End Module statistics
