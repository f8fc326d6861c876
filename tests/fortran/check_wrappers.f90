!------------------------------------------------------------------------------!
! check_wrappers.f90 -- drives every subroutine of the Fortran drop-in layer   !
! (chanles_gpu, smarties_gpu) with the reference's names and arguments on the  !
! inputs a test writes, and writes what they return for comparison with the    !
! oracle.  Usage: check_wrappers <input> <output>                              !
!                                                                              !
! input (stream): int32 nx ny nz LES_model x_mass_cte Dirichlet_nu_t nagents   !
!   precision; float64 nu CFL dPdx dPdx_ref Qflow_x_0; x(nx) y(ny) z(nz);      !
!   U V W nu_t alpha_x rhs_p(2:nxg-1,2:nyg-1,2:nzg)                            !
! output (stream, float64): rhs_u rhs_v rhs_w, compute_eddy_viscosity's nu_t,  !
!   compute_wall_model's nu_t, Qflow, dPdx, the solved rhs_p, the signed max   !
!   divergence; then per send_recv_state_action at istep 1, 1000, 1200 and     !
!   nsteps = 1400: draws(2,nagents,2), states(6,nagents,2),                    !
!   rewards(nagents,2) and the resident alpha_x(nx,2,nzg); aa                  !
!------------------------------------------------------------------------------!
Program check_wrappers
  Use, Intrinsic :: iso_c_binding
  Use chanles_gpu
  Use smarties_gpu
  Implicit None
  Character(400) :: fin, fout
  Integer(c_int) :: hdr(8)
  Real(c_double) :: par(5), Q, dpx, md
  Real(c_double), Allocatable :: x(:), y(:), z(:), U(:,:,:), V(:,:,:), W(:,:,:), nut_in(:,:,:)
  Real(c_double), Allocatable :: a_in(:,:,:), p_in(:,:,:), ru(:,:,:), rv(:,:,:), rw(:,:,:), ne(:,:,:), nw(:,:,:)
  Real(c_double), Allocatable :: a_out(:,:,:)
  Integer :: n1, n2, n3, k, steps(4)
  Interface
     Subroutine smarties_double_dump(path) Bind(c, name='smarties_double_dump')
       Import :: c_char
       Character(kind=c_char), Intent(In) :: path(*)
     End Subroutine smarties_double_dump
  End Interface

  Call Get_command_argument(1, fin)
  Call Get_command_argument(2, fout)
  Open(1, file=fin, access='stream', form='unformatted', action='read')
  Read(1) hdr
  Read(1) par
  n1 = hdr(1); n2 = hdr(2); n3 = hdr(3)
  Allocate(x(n1), y(n2), z(n3))
  Read(1) x, y, z
  Call chanles_initialize(n1, n2, n3, x, y, z, par(1), par(2), par(3), hdr(5), hdr(4), hdr(6), hdr(8))
  Allocate(U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz), nut_in(nxg,nyg,nzg), a_in(nx,2,nzg))
  Allocate(p_in(2:nxg-1,2:nyg-1,2:nzg))
  Read(1) U, V, W, nut_in, a_in, p_in
  Close(1)

  ! the module's globals, as the reference's `global` holds them
  nu_t = nut_in
  alpha_x = a_in
  rhs_p = p_in
  dPdx = par(3)
  Qflow_x_0 = par(5)

  Allocate(ru(2:nx-1,2:nyg-1,2:nzg-1), rv(2:nxg-1,2:ny-1,2:nzg-1), rw(2:nxg-1,2:nyg-1,2:nz-1))
  Allocate(ne(nxg,nyg,nzg), nw(nxg,nyg,nzg), a_out(nx,2,nzg))
  Call compute_rhs_u(U, V, W, ru)
  Call compute_rhs_v(U, V, W, rv)
  Call compute_rhs_w(U, V, W, rw)
  Call compute_eddy_viscosity(U, V, W, ne)
  nw = nut_in
  Call compute_wall_model(U, V, W, nw)
  Call compute_mean_mass_flow_U(U, Q)
  Call compute_dPx_for_constant_mass_flow(U, dpx)
  Call solve_poisson_equation

  Open(2, file=fout, access='stream', form='unformatted', action='write', status='replace')
  Write(2) ru, rv, rw, ne, nw, Q, dpx, rhs_p

  ! the resident simulation: the wrappers above must not have touched it
  Call chanles_set_state(U, V, W)
  Call check_divergence(md)
  Write(2) md

  ! the RL coupling on the resident channel
  nagents = hdr(7)
  dPdx_ref = par(4)
  steps = [1, 1000, 1200, 1400]
  nsteps = steps(4)
  Call Random_seed()
  Do k = 1, 4
     istep = steps(k)
     Call send_recv_state_action(c_null_ptr)
     Call check(chan_get_field(ctx, CHAN_ALPHA_X, a_out), 'get alpha_x')
     Write(2) draws, states, rewards, a_out
  End Do
  Write(2) aa
  Close(2)
  Call smarties_double_dump(Trim(fout) // '.smarties' // c_null_char)
  Call chanles_finalize
End Program check_wrappers
