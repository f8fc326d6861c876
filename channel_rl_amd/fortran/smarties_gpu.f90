!------------------------------------------------------------------------------!
! smarties_gpu.f90 -- the RL coupling of the reference                         !
! (code/smarties_stat.f90:24-258, send_recv_state_action) on the channel that  !
! chanles_gpu keeps resident on the GPU.                                       !
!                                                                              !
! The reference reads U rows 1..5 and nu_t rows 1..2 of its host arrays every  !
! step and writes alpha_x; here the states and rewards are built on the GPU    !
! from the resident U and nu_t (chan_rl_observe) and the agents' actions are   !
! interpolated into the resident alpha_x on the GPU (chan_rl_act), so only     !
! the 6 numbers per agent and wall that smarties exchanges cross to the host.  !
! The random draws (RANDOM_NUMBER, in the reference's order), the step         !
! schedule (istep 1, 1000, every 200th, nsteps), the rewards and the agent     !
! ids are the reference's.  The smarties entry points are those of            !
! code/smarties_extern.cpp:25-62 (C ABI); linking needs the smarties library. !
!                                                                              !
! The reference re-interpolates its stored actions into alpha_x at every step  !
! between receives (code/smarties_stat.f90:231-258); alpha_x does not change   !
! in between, so interpolating once per receive (chan_rl_act) leaves the same  !
! alpha_x.                                                                     !
!------------------------------------------------------------------------------!
Module smarties_gpu
  Use, Intrinsic :: iso_c_binding
  Use chanles_gpu
  Implicit None

  ! code/global.f90:189-192
  Integer, Parameter :: NUM_ACTIONS = 1, STATE_SIZE = 6
  Integer :: istep = 0, nsteps = 0, nagents = 1
  Real(c_double) :: dPdx_ref = 0d0
  ! the last observation: states(6, agent, wall), rewards(agent, wall), the
  ! draws it used, draws(2, agent, wall), and the factor aa of istep 1
  Real(c_double), Allocatable :: states(:,:,:), rewards(:,:), rewards_old(:,:), draws(:,:,:)
  Real(c_double) :: aa = 0d0

  Interface
     Subroutine smarties_sendInitState(comm, S, state_dim, agent_id) Bind(c, name='smarties_sendInitState')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: comm, S
       Integer(c_int), Value :: state_dim, agent_id
     End Subroutine smarties_sendInitState
     Subroutine smarties_sendState(comm, S, state_dim, R, agent_id) Bind(c, name='smarties_sendState')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: comm, S
       Integer(c_int), Value :: state_dim, agent_id
       Real(c_double), Value :: R
     End Subroutine smarties_sendState
     Subroutine smarties_sendLastState(comm, S, state_dim, R, agent_id) Bind(c, name='smarties_sendLastState')
       Import :: c_ptr, c_int, c_double
       Type(c_ptr), Value :: comm, S
       Integer(c_int), Value :: state_dim, agent_id
       Real(c_double), Value :: R
     End Subroutine smarties_sendLastState
     Subroutine smarties_recvAction(comm, A, action_dim, agent_id) Bind(c, name='smarties_recvAction')
       Import :: c_ptr, c_int
       Type(c_ptr), Value :: comm, A
       Integer(c_int), Value :: action_dim, agent_id
     End Subroutine smarties_recvAction
  End Interface

Contains

  ! states and rewards of every agent from the resident U and nu_t
  ! (code/smarties_stat.f90:59-178), the draws taken in the reference's order
  Subroutine observe
    Integer :: i
    If (.Not. Allocated(states)) Then
       Allocate(states(STATE_SIZE, nagents, 2), rewards(nagents, 2), rewards_old(nagents, 2), &
                draws(2, nagents, 2))
       rewards_old = 0d0
    End If
    Do i = 1, nagents
       Call Random_number(draws(1, i, 1))   ! sampling row 3 or 4, bottom wall
       Call Random_number(draws(2, i, 1))   ! its weight
       Call Random_number(draws(1, i, 2))   ! top wall
       Call Random_number(draws(2, i, 2))
    End Do
    Call check(chan_rl_observe(ctx, nagents, dPdx_ref, draws, states, rewards), 'rl_observe')
  End Subroutine observe

  ! factors(agent, wall) = each agent's action: alpha_x at the agents times
  ! the factors, interpolated over the walls on the GPU
  Subroutine act(factors)
    Real(c_double), Intent(In) :: factors(nagents, 2)
    Call check(chan_rl_act(ctx, nagents, factors), 'rl_act')
  End Subroutine act

  Subroutine send_recv_state_action(smarties_comm) Bind(c, name='send_recv_state_action')
    Type(c_ptr), Intent(In), Value :: smarties_comm
    Real(c_double), Target :: state(STATE_SIZE), action(NUM_ACTIONS)
    Real(c_double) :: reward, factors(nagents, 2)
    Integer :: iwall, na
    Integer(c_int) :: agent_id

    ! code/smarties_stat.f90:35-37: the first observation sees the SGS
    ! viscosity of the current state
    If (istep == 1) Call check(chan_compute_eddy_viscosity(ctx), 'compute_eddy_viscosity')
    Call observe

    If (istep < 1000) Then
       If (istep == 1) Then
          ! code/smarties_stat.f90:180-188
          Call Random_number(aa)
          alpha_x = dPdx_ref * (1d0 + (aa - 0.5d0))
          Call chanles_set_alpha_x
       End If
    Else If (istep == 1000) Then
       ! code/smarties_stat.f90:189-198
       Do iwall = 1, 2
          Do na = 1, nagents
             state = states(:, na, iwall)
             agent_id = 2 * (na - 1) + iwall - 1
             Call smarties_sendInitState(smarties_comm, c_loc(state), STATE_SIZE, agent_id)
             Call smarties_recvAction(smarties_comm, c_loc(action), NUM_ACTIONS, agent_id)
             factors(na, iwall) = action(1)
          End Do
       End Do
       rewards_old = rewards
       Call act(factors)
    Else If (istep < nsteps .And. Mod(istep, 200) == 0) Then
       ! code/smarties_stat.f90:199-215
       Do iwall = 1, 2
          Do na = 1, nagents
             state = states(:, na, iwall)
             reward = rewards(na, iwall) - rewards_old(na, iwall)
             If (rewards(na, iwall) > 0d0) reward = reward + rewards(na, iwall)
             agent_id = 2 * (na - 1) + iwall - 1
             Call smarties_sendState(smarties_comm, c_loc(state), STATE_SIZE, reward, agent_id)
             Call smarties_recvAction(smarties_comm, c_loc(action), NUM_ACTIONS, agent_id)
             factors(na, iwall) = action(1)
          End Do
       End Do
       rewards_old = rewards
       Call act(factors)
    Else If (istep == nsteps) Then
       ! code/smarties_stat.f90:216-228
       Do iwall = 1, 2
          Do na = 1, nagents
             state = states(:, na, iwall)
             reward = rewards(na, iwall) - rewards_old(na, iwall)
             If (rewards(na, iwall) > 0d0) reward = reward + rewards(na, iwall)
             agent_id = 2 * (na - 1) + iwall - 1
             Call smarties_sendLastState(smarties_comm, c_loc(state), STATE_SIZE, reward, agent_id)
          End Do
       End Do
    End If
  End Subroutine send_recv_state_action

End Module smarties_gpu
