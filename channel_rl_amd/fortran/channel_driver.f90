!------------------------------------------------------------------------------!
! channel_driver.f90 -- the reference's time loop (code/app_main.f90:100-133,  !
! without the smarties RL coupling) on the GPU through chanles_gpu.            !
!                                                                              !
! Usage: channel_driver <input_parameters> [precision] [h_wm]                  !
!   input_parameters : the reference's text format                             !
!                      (scripts/case_folder/input_parameters_*,                !
!                      code/input_output.f90:20-91)                            !
!   precision        : 0 FP64 (default), 1 FP32 bulk + FP64 Poisson            !
!   h_wm             : >= 0 selects the equilibrium log-law wall model at that !
!                      height (0: first off-wall cell); absent: alpha_x = 0    !
! The initial field is the .init file named in the parameters                  !
! (code/input_output.f90:390-570 format, read without the added noise).        !
! Snapshots are written every nsave steps in the same format                   !
! (code/input_output.f90:579-730), with the true pressure as P.                !
!------------------------------------------------------------------------------!
Program channel_driver
  Use, Intrinsic :: iso_c_binding
  Use iso_fortran_env, Only : Int32, Int64
  Use chanles_gpu
  Implicit None

  Character(200) :: filepar, filein, fileout, dummy_line, arg
  Integer(Int32) :: nxi, nyi, nzi, x_mass_cte, LES_model, nsteps, nsave, nstats, nmonitor
  Integer(Int32) :: nstep_init, random_init, Dirichlet_nu_t, precision, istep, j
  Real(Int64) :: CFL, nu, dPdx_in, h_wm, maxdiv, maxabs, utau, Retau
  Real(Int64), Allocatable :: x(:), y(:), z(:), U(:,:,:), V(:,:,:), W(:,:,:), P(:,:,:), nut(:,:,:)
  Real(Int64), Allocatable :: stats(:,:)
  Logical :: use_wm
  Integer(Int64) :: c0, c1, crate

  If (Command_argument_count() < 1) Stop 'usage: channel_driver <input_parameters> [precision] [h_wm]'
  Call Get_command_argument(1, filepar)
  precision = 0
  use_wm = .False.
  h_wm = 0d0
  If (Command_argument_count() >= 2) Then
     Call Get_command_argument(2, arg)
     Read(arg, *) precision
  End If
  If (Command_argument_count() >= 3) Then
     Call Get_command_argument(3, arg)
     Read(arg, *) h_wm
     use_wm = .True.
  End If

  Write(*,*) '----------------------------------------------------------------------'
  Write(*,*) '   channel LES time step on the GPU (MI355X), Fortran host'
  Write(*,*) '----------------------------------------------------------------------'
  Write(*,*) 'reading input paramaters...'
  ! code/input_output.f90:27-62
  Open(1, file=filepar, form='formatted', action='Read')
  Read(1,*) dummy_line
  Read(1,*) nxi, nyi, nzi
  Read(1,*) dummy_line
  Read(1,*) CFL
  Read(1,*) dummy_line
  Read(1,*) nu
  Read(1,*) dummy_line
  Read(1,*) dPdx_in, x_mass_cte
  Read(1,*) dummy_line
  Read(1,*) LES_model
  Read(1,*) dummy_line
  Read(1,*) nsteps, nsave, nstats, nmonitor
  Read(1,*) dummy_line
  Read(1,*) filein
  Read(1,*) dummy_line
  Read(1,*) fileout
  Read(1,*) dummy_line
  Read(1,*) nstep_init
  Read(1,*) dummy_line
  Read(1,*) random_init
  Read(1,*) dummy_line
  Read(1,*) Dirichlet_nu_t
  Close(1)
  If (random_init /= 0) Stop 'Error: this driver starts from an .init file (random_init = 0)'

  Allocate(x(nxi), y(nyi), z(nzi))
  Allocate(U(nxi, nyi+1, nzi+1), V(nxi+1, nyi, nzi+1), W(nxi+1, nyi+1, nzi))
  Allocate(P(nxi+1, nyi+1, nzi+1), nut(nxi+1, nyi+1, nzi+1), stats(nyi, 7))
  Call read_init(filein)

  Call chanles_initialize(nxi, nyi, nzi, x, y, z, nu, CFL, dPdx_in, x_mass_cte, LES_model, &
                          Dirichlet_nu_t, precision)
  If (use_wm) Call set_equilibrium_wall_model(h_wm)
  Call chanles_set_state(U, V, W)
  Write(*,*) 'grid', nxi, nyi, nzi, ' nu', nu, ' CFL', CFL, ' LES_model', LES_model
  Write(*,*) 'initial mass flow Qflow_x_0', Qflow_x_0

  Call System_clock(c0, crate)
  Do istep = 1, nsteps
     Call compute_dt
     Call compute_time_step_RK3
     If (Mod(istep, nmonitor) == 0 .Or. istep == nsteps) Then
        Call check_divergence(maxdiv)
        Call check_divergence_magnitude(maxabs)
        Call System_clock(c1)
        Write(*,'(A,I8,A,ES12.4,A,ES12.4,A,F10.4)') ' step', istep, '  dt', dt, '  max|div|', maxabs, &
             '  elapsed (s)', Real(c1 - c0, Int64) / Real(crate, Int64)
     End If
     If (Mod(istep, nstats) == 0) Then
        Call compute_plane_statistics(stats)
        If (Any(stats /= stats)) Stop 'Error: NaNs!'
        Call compute_wall_shear(utau, Retau)
        Write(*,'(A,ES14.6,A,ES14.6,A,ES14.6,A,F10.2)') '  U centre', stats(nyi/2+1, 1), &
             '  <uv> near wall', stats(3, 7), '  u_tau', utau, '  Re_tau', Retau
     End If
     If (Mod(istep, nsave) == 0) Then
        ! the snapshot holds the true pressure, as the reference's output
        ! (code/input_output.f90:589)
        Call compute_pressure
        Call chanles_get_state(U, V, W, P, nut)
        Call write_snapshot(istep + nstep_init)
     End If
  End Do
  Call compute_pressure
  Call chanles_get_state(U, V, W, P, nut)
  Call write_snapshot(nsteps + nstep_init)
  Call chanles_finalize
  Write(*,*) 'done'

Contains

  ! code/input_output.f90:390-570 (one reader, no added noise)
  Subroutine read_init(fname)
    Character(*), Intent(In) :: fname
    Integer(Int32) :: n, nn(3)
    Open(2, file=fname, access='stream', form='unformatted', action='Read', status='old')
    Read(2) n
    If (n /= nxi) Stop 'nx_f/=nx'
    Read(2) x
    Read(2) n
    If (n /= nyi) Stop 'ny_f/=ny'
    Read(2) y
    Read(2) n
    If (n /= nzi) Stop 'nz_f/=nz'
    Read(2) z
    Read(2) n
    If (n /= nxi-1) Stop 'nxm_f/=nxm'
    Call skip(n)
    Read(2) n
    If (n /= nyi-1) Stop 'nym_f/=nym'
    Call skip(n)
    Read(2) n
    If (n /= nzi-1) Stop 'nzm_f/=nzm'
    Call skip(n)
    Read(2) nn
    If (nn(1) /= nxi .Or. nn(2) /= nyi+1 .Or. nn(3) /= nzi+1) Stop 'Error! wrong size in input file (U)'
    Read(2) U
    Read(2) nn
    If (nn(1) /= nxi+1 .Or. nn(2) /= nyi .Or. nn(3) /= nzi+1) Stop 'Error! wrong size in input file (V)'
    Read(2) V
    Read(2) nn
    If (nn(1) /= nxi+1 .Or. nn(2) /= nyi+1 .Or. nn(3) /= nzi) Stop 'Error! wrong size in input file (W)'
    Read(2) W
    Close(2)
  End Subroutine read_init

  Subroutine skip(n)
    Integer(Int32), Intent(In) :: n
    Real(Int64) :: buf(n)
    Read(2) buf
  End Subroutine skip

  ! code/input_output.f90:579-730
  Subroutine write_snapshot(num)
    Integer(Int32), Intent(In) :: num
    Character(200) :: fname
    Character(12) :: ext
    Real(Int64) :: xm(nxi-1), ym(nyi-1), zm(nzi-1)
    Write(ext, '(I8)') num
    fname = Trim(Adjustl(fileout)) // '.' // Trim(Adjustl(ext))
    Write(*,*) 'writting ', Trim(fname)
    xm = 0.5d0 * (x(1:nxi-1) + x(2:nxi))
    ym = 0.5d0 * (y(1:nyi-1) + y(2:nyi))
    zm = 0.5d0 * (z(1:nzi-1) + z(2:nzi))
    Open(3, file=fname, access='stream', form='unformatted', action='write', status='replace')
    Write(3) Shape(x), x
    Write(3) Shape(y), y
    Write(3) Shape(z), z
    Write(3) Shape(xm), xm
    Write(3) Shape(ym), ym
    Write(3) Shape(zm), zm
    Write(3) nxi, nyi+1, nzi+1
    Write(3) U
    Write(3) nxi+1, nyi, nzi+1
    Write(3) V
    Write(3) nxi+1, nyi+1, nzi
    Write(3) W
    Write(3) nxi+1, nyi+1, nzi+1
    Write(3) P
    Write(3) nxi+1, nyi+1, nzi+1
    Write(3) nut
    Close(3)
  End Subroutine write_snapshot

End Program channel_driver
