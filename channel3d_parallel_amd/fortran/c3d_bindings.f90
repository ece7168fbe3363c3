!------------------------------------------------------------------------------!
! iso_c_binding interfaces to the gfx950 projection library (include/c3d.h). !
! Fortran hosts use these through the modules in c3d_modules.f90, which keep !
! the reference's module and procedure names.                                !
!------------------------------------------------------------------------------!
Module c3d_bindings

  Use, Intrinsic :: iso_c_binding
  Implicit None

  Integer(c_int), Parameter :: C3D_OK = 0
  Integer(c_int), Parameter :: C3D_PREC_FP64 = 0
  Integer(c_int), Parameter :: C3D_PREC_MIXED = 1

  ! struct c3d_desc
  Type, Bind(C) :: c3d_desc
     Integer(c_int) :: nx, ny, nz, grid_type
     Real(c_double) :: Lxp, Lzp, Ly, alpha, nu
     Integer(c_int) :: precision
     Real(c_double) :: min_buffer_width
  End Type c3d_desc

  Interface

     Integer(c_int) Function c3d_create(desc, ctx) Bind(C, name="c3d_create")
       Import :: c_int, c_ptr, c3d_desc
       Type(c3d_desc), Intent(In) :: desc
       Type(c_ptr), Intent(Out) :: ctx
     End Function c3d_create

     Integer(c_int) Function c3d_destroy(ctx) Bind(C, name="c3d_destroy")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
     End Function c3d_destroy

     Type(c_ptr) Function c3d_last_error(ctx) Bind(C, name="c3d_last_error")
       Import :: c_ptr
       Type(c_ptr), Value :: ctx
     End Function c3d_last_error

     Integer(c_int) Function c3d_grid_get(ctx, what, out, n) Bind(C, name="c3d_grid_get")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: what, n
       Real(c_double) :: out(*)
     End Function c3d_grid_get

     Integer(c_int) Function c3d_set_fields(ctx, U, V, W, P) Bind(C, name="c3d_set_fields")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx, U, V, W, P
     End Function c3d_set_fields

     Integer(c_int) Function c3d_get_fields(ctx, U, V, W, P, Pint) Bind(C, name="c3d_get_fields")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx, U, V, W, P, Pint
     End Function c3d_get_fields

     Integer(c_int) Function c3d_apply_boundary_conditions(ctx) &
          Bind(C, name="c3d_apply_boundary_conditions")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
     End Function c3d_apply_boundary_conditions

     Integer(c_int) Function c3d_divergence(ctx, div) Bind(C, name="c3d_divergence")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: div(*)
     End Function c3d_divergence

     Integer(c_int) Function c3d_gradient(ctx, p, Gx, Gy, Gz) Bind(C, name="c3d_gradient")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: p(*)
       Real(c_double) :: Gx(*), Gy(*), Gz(*)
     End Function c3d_gradient

     Integer(c_int) Function c3d_solve_poisson_equation(ctx, rhs, rk_step, dt) &
          Bind(C, name="c3d_solve_poisson_equation")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: rhs(*)
       Integer(c_int), Value :: rk_step
       Real(c_double), Value :: dt
     End Function c3d_solve_poisson_equation

     Integer(c_int) Function c3d_compute_non_ib_projection(ctx, dt, rk_step) &
          Bind(C, name="c3d_compute_non_ib_projection")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: dt
       Integer(c_int), Value :: rk_step
     End Function c3d_compute_non_ib_projection

     Integer(c_int) Function c3d_check_divergence(ctx, max_div) Bind(C, name="c3d_check_divergence")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: max_div
     End Function c3d_check_divergence

     Integer(c_int) Function c3d_wall_shear(ctx, utau, wtau, Umean, Wmean) &
          Bind(C, name="c3d_wall_shear")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: utau, wtau
       Real(c_double) :: Umean(*), Wmean(*)
     End Function c3d_wall_shear

     Integer(c_int) Function c3d_plane_statistics(ctx, means) &
          Bind(C, name="c3d_plane_statistics")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: means(*)
     End Function c3d_plane_statistics

     Integer(c_int) Function c3d_monitor(ctx, out) Bind(C, name="c3d_monitor")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: out(7)
     End Function c3d_monitor

     Integer(c_int) Function c3d_set_pressure_gradient(ctx, dPdx, dPdy, dPdz) &
          Bind(C, name="c3d_set_pressure_gradient")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: dPdx, dPdy, dPdz
     End Function c3d_set_pressure_gradient

     Integer(c_int) Function c3d_get_pressure_gradient(ctx, dPdx, dPdy, dPdz) &
          Bind(C, name="c3d_get_pressure_gradient")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: dPdx, dPdy, dPdz
     End Function c3d_get_pressure_gradient

     Integer(c_int) Function c3d_save_state(ctx) Bind(C, name="c3d_save_state")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
     End Function c3d_save_state

     Integer(c_int) Function c3d_compute_rhs(ctx, slot) Bind(C, name="c3d_compute_rhs")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: slot
     End Function c3d_compute_rhs

     Integer(c_int) Function c3d_rk3_update(ctx, rk_step, dt) Bind(C, name="c3d_rk3_update")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: rk_step
       Real(c_double), Value :: dt
     End Function c3d_rk3_update

     Integer(c_int) Function c3d_compute_dt(ctx, cfl, dt) Bind(C, name="c3d_compute_dt")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: cfl
       Real(c_double), Intent(Out) :: dt
     End Function c3d_compute_dt

     Integer(c_int) Function c3d_mean_mass_flow_u(ctx, q) Bind(C, name="c3d_mean_mass_flow_u")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: q
     End Function c3d_mean_mass_flow_u

     Integer(c_int) Function c3d_add_to_u(ctx, d) Bind(C, name="c3d_add_to_u")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: d
     End Function c3d_add_to_u

     ! ---- immersed boundary ----
     Integer(c_int) Function c3d_ib_setup(ctx, nb, xb, yb, zb, normals, sb, ub, nzb, nxb, nxb1) &
          Bind(C, name="c3d_ib_setup")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: nb, nzb, nxb, nxb1
       Real(c_double), Intent(In) :: xb(*), yb(*), zb(*), normals(*), sb(*), ub(*)
     End Function c3d_ib_setup

     Integer(c_int) Function c3d_ib_setup_ref(ctx, nb, xb, yb, zb, normals, sb, ub, yref, nzb, &
          nxb, nxb1) Bind(C, name="c3d_ib_setup_ref")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: nb, nzb, nxb, nxb1
       Real(c_double), Intent(In) :: xb(*), yb(*), zb(*), normals(*), sb(*), ub(*)
       Integer(c_int), Intent(In) :: yref(*)
     End Function c3d_ib_setup_ref

     Integer(c_int) Function c3d_ib_setup_wave(ctx, body, nxb, nzb, p1, p2, p3, ramp, t) &
          Bind(C, name="c3d_ib_setup_wave")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: body, nxb, nzb
       Real(c_double), Value :: p1, p2, p3, ramp, t
     End Function c3d_ib_setup_wave

     Integer(c_int) Function c3d_ib_move(ctx, t) Bind(C, name="c3d_ib_move")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: t
     End Function c3d_ib_move

     Integer(c_int) Function c3d_ib_get_markers(ctx, xb, yb, zb, normals, sb, ub, yref) &
          Bind(C, name="c3d_ib_get_markers")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: xb(*), yb(*), zb(*), normals(*), sb(*), ub(*)
       Integer(c_int), Intent(Out) :: yref(*)
     End Function c3d_ib_get_markers

     Integer(c_int) Function c3d_write_snapshot(ctx, path, dt, tangents_1, tangents_2) &
          Bind(C, name="c3d_write_snapshot")
       Import :: c_int, c_ptr, c_double, c_char
       Type(c_ptr), Value :: ctx
       Character(kind=c_char), Intent(In) :: path(*)
       Real(c_double), Value :: dt
       Type(c_ptr), Value :: tangents_1, tangents_2
     End Function c3d_write_snapshot

     Integer(c_int) Function c3d_read_snapshot(ctx, path, what, dt_out) &
          Bind(C, name="c3d_read_snapshot")
       Import :: c_int, c_ptr, c_double, c_char
       Type(c_ptr), Value :: ctx
       Character(kind=c_char), Intent(In) :: path(*)
       Integer(c_int), Value :: what
       Real(c_double), Intent(Out) :: dt_out
     End Function c3d_read_snapshot

     Integer(c_int) Function c3d_set_time(ctx, t) Bind(C, name="c3d_set_time")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: t
     End Function c3d_set_time

     Integer(c_int) Function c3d_get_time(ctx, t) Bind(C, name="c3d_get_time")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: t
     End Function c3d_get_time

     Integer(c_int) Function c3d_ib_set_solver(ctx, tol, maxit, guess_mode) &
          Bind(C, name="c3d_ib_set_solver")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: tol
       Integer(c_int), Value :: maxit, guess_mode
     End Function c3d_ib_set_solver

     ! Refinement sweeps per projection (mixed precision; FP64 defect
     ! correction on an FP64 context) and the residual entering each sweep.
     Integer(c_int) Function c3d_set_refinement(ctx, nref) Bind(C, name="c3d_set_refinement")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: nref
     End Function c3d_set_refinement

     Integer(c_int) Function c3d_refinement_residuals(ctx, rmax, n) &
          Bind(C, name="c3d_refinement_residuals")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: rmax(*)
       Integer(c_int), Value :: n
     End Function c3d_refinement_residuals

     Integer(c_int) Function c3d_ib_set_force(ctx, fb) Bind(C, name="c3d_ib_set_force")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: fb(*)
     End Function c3d_ib_set_force

     Integer(c_int) Function c3d_ib_get_force(ctx, fb) Bind(C, name="c3d_ib_get_force")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: fb(*)
     End Function c3d_ib_get_force

     Integer(c_int) Function c3d_compute_ib_projection(ctx, dt, rk_step, iters, err) &
          Bind(C, name="c3d_compute_ib_projection")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: dt
       Integer(c_int), Value :: rk_step
       Integer(c_int), Intent(Out) :: iters
       Real(c_double), Intent(Out) :: err
     End Function c3d_compute_ib_projection

     Integer(c_int) Function c3d_ib_schur(ctx, f, Sf, dt, rk_step) Bind(C, name="c3d_ib_schur")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: f(*)
       Real(c_double) :: Sf(*)
       Real(c_double), Value :: dt
       Integer(c_int), Value :: rk_step
     End Function c3d_ib_schur

     Integer(c_int) Function c3d_ib_bicgstab(ctx, x, b, tol, maxit, dt, rk_step, iters, err) &
          Bind(C, name="c3d_ib_bicgstab")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: x(*)
       Real(c_double), Intent(In) :: b(*)
       Real(c_double), Value :: tol, dt
       Integer(c_int), Value :: maxit, rk_step
       Integer(c_int), Intent(Out) :: iters
       Real(c_double), Intent(Out) :: err
     End Function c3d_ib_bicgstab

     Integer(c_int) Function c3d_check_slip(ctx, max_slip) Bind(C, name="c3d_check_slip")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: max_slip
     End Function c3d_check_slip

     Integer(c_int) Function c3d_ib_stats(ctx, last_iters, accum_iters, last_err) &
          Bind(C, name="c3d_ib_stats")
       Import :: c_int, c_ptr, c_double, c_long_long
       Type(c_ptr), Value :: ctx
       Integer(c_int), Intent(Out) :: last_iters
       Integer(c_long_long), Intent(Out) :: accum_iters
       Real(c_double), Intent(Out) :: last_err
     End Function c3d_ib_stats

     Integer(c_int) Function c3d_ib_set_exterior_pressure_gradient(ctx, flag) &
          Bind(C, name="c3d_ib_set_exterior_pressure_gradient")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: flag
     End Function c3d_ib_set_exterior_pressure_gradient

     ! host-array forms (device scratch of the context; resident state untouched)
     Integer(c_int) Function c3d_apply_boundary_conditions_arrays(ctx, U, V, W) &
          Bind(C, name="c3d_apply_boundary_conditions_arrays")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: U(*), V(*), W(*)
     End Function c3d_apply_boundary_conditions_arrays

     Integer(c_int) Function c3d_divergence_arrays(ctx, U, V, W, div) &
          Bind(C, name="c3d_divergence_arrays")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Real(c_double) :: div(*)
     End Function c3d_divergence_arrays

     Integer(c_int) Function c3d_compute_rhs_arrays(ctx, U, V, W, Fu, Fv, Fw) &
          Bind(C, name="c3d_compute_rhs_arrays")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Type(c_ptr), Value :: Fu, Fv, Fw
     End Function c3d_compute_rhs_arrays

     Integer(c_int) Function c3d_mean_mass_flow_u_array(ctx, U, q) &
          Bind(C, name="c3d_mean_mass_flow_u_array")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: U(*)
       Real(c_double), Intent(Out) :: q
     End Function c3d_mean_mass_flow_u_array

     Integer(c_int) Function c3d_ib_interpolate_arrays(ctx, with_dxn, U, V, W, out) &
          Bind(C, name="c3d_ib_interpolate_arrays")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: with_dxn
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Real(c_double) :: out(*)
     End Function c3d_ib_interpolate_arrays

     Integer(c_int) Function c3d_ib_interpolate_c(ctx, p, out) Bind(C, name="c3d_ib_interpolate_c")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(In) :: p(*)
       Real(c_double) :: out(*)
     End Function c3d_ib_interpolate_c

     Integer(c_int) Function c3d_ib_spread(ctx, comp, fvec, field) Bind(C, name="c3d_ib_spread")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: comp
       Real(c_double), Intent(In) :: fvec(*)
       Real(c_double) :: field(*)
     End Function c3d_ib_spread

     Integer(c_int) Function c3d_ib_remove_mean(ctx, f) Bind(C, name="c3d_ib_remove_mean")
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double) :: f(*)
     End Function c3d_ib_remove_mean

     Integer(c_int) Function c3d_ib_compute_heaviside(ctx) Bind(C, name="c3d_ib_compute_heaviside")
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: ctx
     End Function c3d_ib_compute_heaviside

  End Interface

Contains

  ! Library error message as a Fortran string.
  Function c3d_error_message(ctx) Result(msg)
    Type(c_ptr), Intent(In) :: ctx
    Character(len=512) :: msg
    Character(kind=c_char), Pointer :: cs(:)
    Integer :: i
    msg = ' '
    Call c_f_pointer(c3d_last_error(ctx), cs, [512])
    Do i = 1, 512
       If (cs(i) == c_null_char) Exit
       msg(i:i) = cs(i)
    End Do
  End Function c3d_error_message

End Module c3d_bindings
