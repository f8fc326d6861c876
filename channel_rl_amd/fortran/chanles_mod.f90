!------------------------------------------------------------------------------!
! chanles_mod.f90 -- Fortran host layer over the C ABI of include/chanles.h.   !
!                                                                              !
! Module chanles_c   : iso_c_binding interfaces, one per C entry point.       !
! Module chanles_gpu : the reference's hot-path subroutines with their own    !
!                      names and argument shapes (code/time_integration.f90,  !
!                      code/equations.f90, code/subgrid.f90,                  !
!                      code/wallmodel.f90, code/projection.f90,               !
!                      code/mass_flow.f90, code/boundary_conditions.f90), so  !
!                      a host that `Use`s it instead of those modules keeps   !
!                      its call sites.  A non-zero status becomes `Stop`, the !
!                      reference's error convention.                          !
!------------------------------------------------------------------------------!
Module chanles_c
  Use, Intrinsic :: iso_c_binding
  Implicit None

  Integer(c_int), Parameter :: CHAN_FP64 = 0, CHAN_FP32_BULK = 1
  Integer(c_int), Parameter :: CHAN_U = 0, CHAN_V = 1, CHAN_W = 2, CHAN_NU_T = 3, CHAN_P = 4, &
                               CHAN_ALPHA_X = 5, CHAN_RHS_U = 6, CHAN_RHS_V = 7, CHAN_RHS_W = 8
  Integer(c_int), Parameter :: CHAN_WALL_GIVEN = 0, CHAN_WALL_EQUILIBRIUM = 1

  Type, Bind(c) :: chan_config
     Integer(c_int) :: nx, ny, nz
     Type(c_ptr)    :: x, y, z
     Real(c_double) :: nu, CFL, dPdx
     Integer(c_int) :: x_mass_cte, LES_model, Dirichlet_nu_t, precision, device
  End Type chan_config

  Interface
     Integer(c_int) Function chan_create(cfg, ctx) Bind(c, name='chan_create')
       Import :: chan_config, c_ptr, c_int
       Type(chan_config), Intent(In) :: cfg
       Type(c_ptr), Intent(Out) :: ctx
     End Function chan_create
     Subroutine chan_destroy(ctx) Bind(c, name='chan_destroy')
       Import :: c_ptr
       Type(c_ptr), Value :: ctx
     End Subroutine chan_destroy
     Type(c_ptr) Function chan_last_error(ctx) Bind(c, name='chan_last_error')
       Import :: c_ptr
       Type(c_ptr), Value :: ctx
     End Function chan_last_error
     Integer(c_int) Function chan_set_field(ctx, which, src) Bind(c, name='chan_set_field')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: which
       Real(c_double), Intent(In) :: src(*)
     End Function chan_set_field
     Integer(c_int) Function chan_get_field(ctx, which, dst) Bind(c, name='chan_get_field')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: which
       Real(c_double), Intent(Out) :: dst(*)
     End Function chan_get_field
     Integer(c_int) Function chan_set_scalars(ctx, dt, t, dPdx, Q) Bind(c, name='chan_set_scalars')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: dt, t, dPdx, Q
     End Function chan_set_scalars
     Integer(c_int) Function chan_get_scalars(ctx, out4) Bind(c, name='chan_get_scalars')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: out4(4)
     End Function chan_get_scalars
     Integer(c_int) Function chan_apply_boundary_conditions(ctx) Bind(c, name='chan_apply_boundary_conditions')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_apply_boundary_conditions
     Integer(c_int) Function chan_compute_eddy_viscosity(ctx) Bind(c, name='chan_compute_eddy_viscosity')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_compute_eddy_viscosity
     Integer(c_int) Function chan_compute_wall_model(ctx) Bind(c, name='chan_compute_wall_model')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_compute_wall_model
     Integer(c_int) Function chan_compute_rhs(ctx) Bind(c, name='chan_compute_rhs')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_compute_rhs
     Integer(c_int) Function chan_compute_projection_step(ctx) Bind(c, name='chan_compute_projection_step')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_compute_projection_step
     Integer(c_int) Function chan_initialize_mass_flow(ctx) Bind(c, name='chan_initialize_mass_flow')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_initialize_mass_flow
     Integer(c_int) Function chan_compute_dPx_for_constant_mass_flow(ctx, dPdx) &
          Bind(c, name='chan_compute_dPx_for_constant_mass_flow')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: dPdx
     End Function chan_compute_dPx_for_constant_mass_flow
     Integer(c_int) Function chan_compute_mean_mass_flow_U(ctx, Q) Bind(c, name='chan_compute_mean_mass_flow_U')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: Q
     End Function chan_compute_mean_mass_flow_U
     Integer(c_int) Function chan_compute_dt(ctx, dt) Bind(c, name='chan_compute_dt')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: dt
     End Function chan_compute_dt
     Integer(c_int) Function chan_compute_time_step_RK3(ctx) Bind(c, name='chan_compute_time_step_RK3')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_compute_time_step_RK3
     Integer(c_int) Function chan_check_divergence(ctx, mx, ma) Bind(c, name='chan_check_divergence')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: mx, ma
     End Function chan_check_divergence
     Integer(c_int) Function chan_set_wall_model(ctx, mode, h_wm, kappa, B) Bind(c, name='chan_set_wall_model')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: mode
       Real(c_double), Value :: h_wm, kappa, B
     End Function chan_set_wall_model
     Integer(c_int) Function chan_plane_statistics(ctx, out) Bind(c, name='chan_plane_statistics')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: out(*)
     End Function chan_plane_statistics
     Integer(c_int) Function chan_compute_pressure(ctx) Bind(c, name='chan_compute_pressure')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_compute_pressure
     Integer(c_int) Function chan_wall_shear(ctx, out4) Bind(c, name='chan_wall_shear')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Intent(Out) :: out4(4)
     End Function chan_wall_shear
     Integer(c_int) Function chan_synchronize(ctx) Bind(c, name='chan_synchronize')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_synchronize
     Integer(c_int) Function chan_solve_poisson_equation(ctx) Bind(c, name='chan_solve_poisson_equation')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: ctx
     End Function chan_solve_poisson_equation
     Integer(c_int) Function chan_compute_equilibrium_wall_stress(ctx, h_wm, kappa, B) &
          Bind(c, name='chan_compute_equilibrium_wall_stress')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Real(c_double), Value :: h_wm, kappa, B
     End Function chan_compute_equilibrium_wall_stress
     Integer(c_int) Function chan_rl_observe(ctx, nagents, dPdx_ref, draws, states, rewards) &
          Bind(c, name='chan_rl_observe')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: nagents
       Real(c_double), Value :: dPdx_ref
       Real(c_double), Intent(In) :: draws(*)
       Real(c_double), Intent(Out) :: states(*), rewards(*)
     End Function chan_rl_observe
     Integer(c_int) Function chan_rl_act(ctx, nagents, factors) Bind(c, name='chan_rl_act')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: ctx
       Integer(c_int), Value :: nagents
       Real(c_double), Intent(In) :: factors(*)
     End Function chan_rl_act
     Integer(c_size_t) Function c_strlen(s) Bind(c, name='strlen')
       Import :: c_ptr, c_size_t
       Type(c_ptr), Value :: s
     End Function c_strlen
  End Interface

End Module chanles_c

Module chanles_gpu
  Use, Intrinsic :: iso_c_binding
  Use chanles_c
  Implicit None

  ! sizes (code/global.f90:49-61) and the scalars the reference keeps global
  Integer :: nx, ny, nz, nxg, nyg, nzg
  Real(c_double) :: dt = 0d0, t = 0d0, dPdx = 0d0, Qflow_x_0 = 0d0
  ! the reference's module arrays that its hot-path subroutines read or
  ! write besides their arguments (code/global.f90:85-186,
  ! code/initialization.f90:159,467,476): the eddy viscosity compute_rhs_*
  ! reads, the wall stress compute_wall_model reads, the pseudo-pressure
  ! solve_poisson_equation solves in place
  Real(c_double), Allocatable, Target :: nu_t(:,:,:), alpha_x(:,:,:), rhs_p(:,:,:)
  ! ctx holds the resident simulation (compute_time_step_RK3 and the calls
  ! without array arguments act on it); wctx is a work context of the same
  ! channel on which the wrappers that take arrays compute, so that a call
  ! between steps never replaces the resident state
  Type(c_ptr) :: ctx = c_null_ptr, wctx = c_null_ptr
  Real(c_double), Allocatable, Target :: xf(:), yf(:), zf(:)

Contains

  Subroutine check(status, what)
    Integer(c_int), Intent(In) :: status
    Character(*), Intent(In) :: what
    Character(len=512) :: msg
    Character(kind=c_char), Pointer :: cmsg(:)
    Integer :: n, i
    If (status == 0) Return
    msg = ''
    n = Int(c_strlen(chan_last_error(ctx)))
    Call c_f_pointer(chan_last_error(ctx), cmsg, [n])
    Do i = 1, Min(n, 512)
       msg(i:i) = cmsg(i)
    End Do
    Write(*,*) 'Error in ', what, ': ', Trim(msg)
    Stop 'Error: GPU channel call failed'
  End Subroutine check

  ! check() for a call on another context than ctx
  Subroutine check_ctx(status, c, what)
    Integer(c_int), Intent(In) :: status
    Type(c_ptr), Intent(In) :: c
    Character(*), Intent(In) :: what
    Type(c_ptr) :: keep
    If (status == 0) Return
    keep = ctx
    ctx = c
    Call check(status, what)
    ctx = keep
  End Subroutine check_ctx

  ! code/initialization.f90:23 initialize (the GPU part): grid and parameters
  Subroutine chanles_initialize(nx_, ny_, nz_, x, y, z, nu, CFL, dPdx_, x_mass_cte, LES_model, &
                                Dirichlet_nu_t, precision)
    Integer, Intent(In) :: nx_, ny_, nz_, x_mass_cte, LES_model, Dirichlet_nu_t, precision
    Real(c_double), Intent(In) :: x(nx_), y(ny_), z(nz_), nu, CFL, dPdx_
    Type(chan_config) :: cfg
    nx = nx_; ny = ny_; nz = nz_
    nxg = nx + 1; nyg = ny + 1; nzg = nz + 1
    Allocate(xf(nx), yf(ny), zf(nz))
    xf = x; yf = y; zf = z
    cfg%nx = nx; cfg%ny = ny; cfg%nz = nz
    cfg%x = c_loc(xf); cfg%y = c_loc(yf); cfg%z = c_loc(zf)
    cfg%nu = nu; cfg%CFL = CFL; cfg%dPdx = dPdx_
    cfg%x_mass_cte = x_mass_cte; cfg%LES_model = LES_model; cfg%Dirichlet_nu_t = Dirichlet_nu_t
    cfg%precision = precision; cfg%device = 0
    dPdx = dPdx_
    Call check(chan_create(cfg, ctx), 'chan_create')
    Call check_ctx(chan_create(cfg, wctx), wctx, 'chan_create (work context)')
    Allocate(nu_t(nxg,nyg,nzg), alpha_x(nx,2,nzg), rhs_p(2:nxg-1,2:nyg-1,2:nzg))
    nu_t = 0d0
    alpha_x = 0d0
    rhs_p = 0d0
  End Subroutine chanles_initialize

  Subroutine chanles_finalize
    If (c_associated(ctx)) Call chan_destroy(ctx)
    If (c_associated(wctx)) Call chan_destroy(wctx)
    ctx = c_null_ptr
    wctx = c_null_ptr
    If (Allocated(xf)) Deallocate(xf, yf, zf)
    If (Allocated(nu_t)) Deallocate(nu_t, alpha_x, rhs_p)
  End Subroutine chanles_finalize

  ! the module's alpha_x as the wall stress of the resident simulation's steps
  ! (code/wallmodel.f90:30-35 reads the global alpha_x at every substep)
  Subroutine chanles_set_alpha_x
    Call check(chan_set_field(ctx, CHAN_ALPHA_X, alpha_x), 'set alpha_x')
  End Subroutine chanles_set_alpha_x

  ! host -> device state, then Qflow_x_0 (code/initialization.f90:442)
  Subroutine chanles_set_state(U, V, W)
    Real(c_double), Intent(In) :: U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz)
    Real(c_double) :: s(4)
    Call check(chan_set_field(ctx, CHAN_U, U), 'set U')
    Call check(chan_set_field(ctx, CHAN_V, V), 'set V')
    Call check(chan_set_field(ctx, CHAN_W, W), 'set W')
    Call check(chan_initialize_mass_flow(ctx), 'initialize_mass_flow')
    Call check(chan_get_scalars(ctx, s), 'get_scalars')
    Qflow_x_0 = s(4)
    Call check(chan_set_scalars(ctx, dt, t, dPdx, Qflow_x_0), 'set_scalars')
  End Subroutine chanles_set_state

  Subroutine chanles_get_state(U, V, W, P, nu_t)
    Real(c_double), Intent(Out) :: U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz)
    Real(c_double), Intent(Out) :: P(nxg,nyg,nzg), nu_t(nxg,nyg,nzg)
    Real(c_double) :: s(4)
    Call check(chan_get_field(ctx, CHAN_U, U), 'get U')
    Call check(chan_get_field(ctx, CHAN_V, V), 'get V')
    Call check(chan_get_field(ctx, CHAN_W, W), 'get W')
    Call check(chan_get_field(ctx, CHAN_P, P), 'get P')
    Call check(chan_get_field(ctx, CHAN_NU_T, nu_t), 'get nu_t')
    Call check(chan_get_scalars(ctx, s), 'get_scalars')
    dt = s(1); t = s(2); dPdx = s(3); Qflow_x_0 = s(4)
  End Subroutine chanles_get_state

  ! code/time_integration.f90:102
  Subroutine compute_dt
    Call check(chan_compute_dt(ctx, dt), 'compute_dt')
  End Subroutine compute_dt

  ! code/time_integration.f90:24 (whole step resident on the GPU)
  Subroutine compute_time_step_RK3
    Call check(chan_compute_time_step_RK3(ctx), 'compute_time_step_RK3')
  End Subroutine compute_time_step_RK3

  ! code/boundary_conditions.f90:20 on the device-resident U,V,W
  Subroutine apply_boundary_conditions
    Call check(chan_apply_boundary_conditions(ctx), 'apply_boundary_conditions')
  End Subroutine apply_boundary_conditions

  ! code/projection.f90:20 on the device-resident U,V,W
  Subroutine compute_projection_step
    Call check(chan_compute_projection_step(ctx), 'compute_projection_step')
  End Subroutine compute_projection_step

  ! the wrappers below that take arrays compute on the work context wctx
  Subroutine w_set(which, a, what)
    Integer(c_int), Intent(In) :: which
    Real(c_double), Intent(In) :: a(*)
    Character(*), Intent(In) :: what
    Call check_ctx(chan_set_field(wctx, which, a), wctx, what)
  End Subroutine w_set

  Subroutine w_velocity(U_, V_, W_)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Call w_set(CHAN_U, U_, 'set U')
    Call w_set(CHAN_V, V_, 'set V')
    Call w_set(CHAN_W, W_, 'set W')
  End Subroutine w_velocity

  ! code/subgrid.f90:23 compute_eddy_viscosity(U_,V_,W_,nu_t_)
  Subroutine compute_eddy_viscosity(U_, V_, W_, nu_t_)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Real(c_double), Intent(Out) :: nu_t_(nxg,nyg,nzg)
    Call w_velocity(U_, V_, W_)
    Call check_ctx(chan_compute_eddy_viscosity(wctx), wctx, 'compute_eddy_viscosity')
    Call check_ctx(chan_get_field(wctx, CHAN_NU_T, nu_t_), wctx, 'get nu_t')
  End Subroutine compute_eddy_viscosity

  ! code/wallmodel.f90:23 compute_wall_model(U_,V_,W_,nu_t_) with the
  ! module's alpha_x (code/wallmodel.f90:30-35)
  Subroutine compute_wall_model(U_, V_, W_, nu_t_)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Real(c_double), Intent(InOut) :: nu_t_(nxg,nyg,nzg)
    Call w_set(CHAN_U, U_, 'set U')
    Call w_set(CHAN_NU_T, nu_t_, 'set nu_t')
    Call w_set(CHAN_ALPHA_X, alpha_x, 'set alpha_x')
    Call check_ctx(chan_compute_wall_model(wctx), wctx, 'compute_wall_model')
    Call check_ctx(chan_get_field(wctx, CHAN_NU_T, nu_t_), wctx, 'get nu_t')
  End Subroutine compute_wall_model

  ! the three right-hand sides of U_,V_,W_ with the module's nu_t and dPdx
  Subroutine w_rhs(U_, V_, W_)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Call w_velocity(U_, V_, W_)
    Call w_set(CHAN_NU_T, nu_t, 'set nu_t')
    Call check_ctx(chan_set_scalars(wctx, dt, t, dPdx, Qflow_x_0), wctx, 'set_scalars')
    Call check_ctx(chan_compute_rhs(wctx), wctx, 'compute_rhs')
  End Subroutine w_rhs

  ! code/equations.f90:24 compute_rhs_u(U_,V_,W_,rhs_u)
  Subroutine compute_rhs_u(U_, V_, W_, rhs_u)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Real(c_double), Intent(Out) :: rhs_u(2:nx-1,2:nyg-1,2:nzg-1)
    Call w_rhs(U_, V_, W_)
    Call check_ctx(chan_get_field(wctx, CHAN_RHS_U, rhs_u), wctx, 'get rhs_u')
  End Subroutine compute_rhs_u

  ! code/equations.f90:174 compute_rhs_v(U_,V_,W_,rhs_v)
  Subroutine compute_rhs_v(U_, V_, W_, rhs_v)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Real(c_double), Intent(Out) :: rhs_v(2:nxg-1,2:ny-1,2:nzg-1)
    Call w_rhs(U_, V_, W_)
    Call check_ctx(chan_get_field(wctx, CHAN_RHS_V, rhs_v), wctx, 'get rhs_v')
  End Subroutine compute_rhs_v

  ! code/equations.f90:327 compute_rhs_w(U_,V_,W_,rhs_w)
  Subroutine compute_rhs_w(U_, V_, W_, rhs_w)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg), V_(nxg,ny,nzg), W_(nxg,nyg,nz)
    Real(c_double), Intent(Out) :: rhs_w(2:nxg-1,2:nyg-1,2:nz-1)
    Call w_rhs(U_, V_, W_)
    Call check_ctx(chan_get_field(wctx, CHAN_RHS_W, rhs_w), wctx, 'get rhs_w')
  End Subroutine compute_rhs_w

  ! code/mass_flow.f90:24 compute_mean_mass_flow_U(U,Qflow)
  Subroutine compute_mean_mass_flow_U(U_, Qflow)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg)
    Real(c_double), Intent(Out) :: Qflow
    Call w_set(CHAN_U, U_, 'set U')
    Call check_ctx(chan_compute_mean_mass_flow_U(wctx, Qflow), wctx, 'compute_mean_mass_flow_U')
  End Subroutine compute_mean_mass_flow_U

  ! code/mass_flow.f90:113 compute_dPx_for_constant_mass_flow(U,dPdx) with the
  ! module's Qflow_x_0
  Subroutine compute_dPx_for_constant_mass_flow(U_, dPdx_)
    Real(c_double), Intent(In) :: U_(nx,nyg,nzg)
    Real(c_double), Intent(Out) :: dPdx_
    Call w_set(CHAN_U, U_, 'set U')
    Call check_ctx(chan_set_scalars(wctx, dt, t, dPdx, Qflow_x_0), wctx, 'set_scalars')
    Call check_ctx(chan_compute_dPx_for_constant_mass_flow(wctx, dPdx_), wctx, &
                   'compute_dPx_for_constant_mass_flow')
  End Subroutine compute_dPx_for_constant_mass_flow

  ! code/projection.f90:74 solve_poisson_equation: the module's rhs_p
  ! (2:nxg-1,2:nyg-1,2:nzg) in place, with its periodic images in x and z
  ! (code/projection.f90:133-150); on one domain the extra plane nzg is
  ! left as it was (code/projection.f90:156-180 receives nothing there)
  Subroutine solve_poisson_equation
    Real(c_double), Allocatable :: P(:,:,:)
    Allocate(P(nxg,nyg,nzg))
    P = 0d0
    P(2:nxg-1,2:nyg-1,2:nzg) = rhs_p
    Call w_set(CHAN_P, P, 'set P')
    Call check_ctx(chan_solve_poisson_equation(wctx), wctx, 'solve_poisson_equation')
    Call check_ctx(chan_get_field(wctx, CHAN_P, P), wctx, 'get P')
    rhs_p(2:nxg-1,2:nyg-1,2:nzg-1) = P(2:nxg-1,2:nyg-1,2:nzg-1)
    Deallocate(P)
  End Subroutine solve_poisson_equation

  ! code/projection.f90:223 check_divergence(max_divergence) of the resident
  ! state: the signed maximum, started at 0 as the reference does
  Subroutine check_divergence(max_divergence)
    Real(c_double), Intent(Out) :: max_divergence
    Real(c_double) :: max_abs
    Call check(chan_check_divergence(ctx, max_divergence, max_abs), 'check_divergence')
  End Subroutine check_divergence

  ! the largest |divergence| of the resident state (no reference counterpart:
  ! the round-off check the signed maximum cannot give)
  Subroutine check_divergence_magnitude(max_abs_divergence)
    Real(c_double), Intent(Out) :: max_abs_divergence
    Real(c_double) :: max_div
    Call check(chan_check_divergence(ctx, max_div, max_abs_divergence), 'check_divergence')
  End Subroutine check_divergence_magnitude

  Subroutine set_equilibrium_wall_model(h_wm)
    Real(c_double), Intent(In) :: h_wm
    Call check(chan_set_wall_model(ctx, CHAN_WALL_EQUILIBRIUM, h_wm, 0.41d0, 5.2d0), 'set_wall_model')
  End Subroutine set_equilibrium_wall_model

  ! code/statistics.f90:52-64 plane means, rows 1..ny
  Subroutine compute_plane_statistics(stats)
    Real(c_double), Intent(Out) :: stats(ny,7)
    Call check(chan_plane_statistics(ctx, stats), 'plane_statistics')
  End Subroutine compute_plane_statistics

  ! code/statistics.f90:95-130: u_tau and Re_tau from the wall shear
  Subroutine compute_wall_shear(utau_, Retau_)
    Real(c_double), Intent(Out) :: utau_, Retau_
    Real(c_double) :: w(4)
    Call check(chan_wall_shear(ctx, w), 'wall_shear')
    utau_ = w(3)
    Retau_ = w(4)
  End Subroutine compute_wall_shear

  ! code/pressure.f90:21 compute_pressure: the true pressure into the
  ! device P (read back with chanles_get_state)
  Subroutine compute_pressure
    Call check(chan_compute_pressure(ctx), 'compute_pressure')
  End Subroutine compute_pressure

End Module chanles_gpu
