!------------------------------------------------------------------------------
! mdcfg_gpu -- impacts as the benchmark configurations define them, driven
! from the Fortran host through the MD-step mode of the C-ABI (fib_md_*).
!
! R replica slabs of n x n x nz diamond cells (a = 5.431 A, (100) top, box
! z = slab + 40 A of vacuum, the bottom two atomic layers pinned,
! Maxwell-Boltzmann velocities at T without net momentum, M Ga ions parked
! 8 A above the surface); M sequential impacts per replica, each from a
! uniform random spot of the surface cell 5 A above the surface at E eV
! along (polar, azimuth), S steps per impact with dt = min(1 fs, 0.05 A /
! |v|max); atoms more than 10 A above the surface are removed; optional
! Berendsen shells on the sides and the bottom.  The single-impact
! configuration is R = 1, M = 1; the cumulative dose is R = 16, n = nz = 40,
! M = 200 at 2 keV with 3 A shells.
!
! After each impact it prints one line with the replicas' mean cumulative
! sputter yield, the energy ledger's drift since the start (kinetic +
! potential energy + what the shells took + what removed atoms carried out -
! what the ions brought in, in eV: mean and largest magnitude), the mean
! energy fired so far, the simulated time and the timestep of replica 1, and
! the wall time of the impact (firing and steps).
! With --frames, the first simulated replica is written to data/mdrun2.xyz
! and data/3dmol2.xyz in the reference's format, at the start and after each
! impact.  The frame of an impact is an asynchronous snapshot
! (fib_snapshot_begin), written while the next impact's steps run
! (fib_md_run_begin / fib_md_run_end); --sync-frames fetches and writes each
! frame before going on instead (same files).
!
! All replicas fire their next ion in one call (fib_md_fire_many: one list
! rebuild of the fired replicas).  --listfac sets the list radius factor
! (cell size and skin; the autotuner's choice for the dose batch is 1.15).
!
! --first F --count C simulate only replicas F .. F+C-1 of the R: every
! replica's random numbers are still drawn in the same order, and a
! replica's trajectory does not depend on the others in the batch, so runs
! over disjoint groups give exactly the replicas of one run over all R.
! A second line per impact lists each simulated replica's cumulative yield
! and ledger drift.
!
! Usage: mdcfg_gpu [--cells N] [--nz NZ] [--replicas R] [--impacts M]
!                  [--steps S] [--energy EV] [--polar DEG] [--azimuth DEG]
!                  [--temp K] [--shell A] [--seed N] [--listfac F] [--mixed]
!                  [--first F] [--count C] [--frames] [--sync-frames]
!------------------------------------------------------------------------------
program mdcfg_gpu
  use iso_c_binding
  use fibgpu
  use fibhost
  implicit none

  real(dp), parameter :: ang = 1.0E-10_dp, ev = 1.602176634E-19_dp
  real(dp), parameter :: a_cfg = 5.431E-10_dp, vacuum = 40.0E-10_dp, park = 8.0E-10_dp
  real(dp), parameter :: ga_speed = 1663.6334_dp   ! m/s per sqrt(eV) (knock.f90:29)

  integer :: n, nz, nrep, nimp, nsteps, k, nargs, r, m, ios, first, cnt, q
  integer(c_int32_t) :: prec, rc
  real(dp) :: energy, polar, azim, tempk, shell, surf, th, ph, spd, listfac
  real(dp) :: x(3), v(3), ymean, dmean, dmax, fmean, temp
  logical :: frames, sync_frames, pend
  integer(c_int64_t) :: plt
  real(dp) :: ptime
  character(len=64) :: arg
  type(siga_t) :: s
  type(state_t) :: st
  type(fib_params) :: prm
  type(fib_md_params) :: mp
  type(fib_md_report) :: rep
  type(fib_instance_desc), allocatable :: desc(:)
  type(c_ptr) :: h
  real(dp), allocatable, target :: X0(:,:), V0(:,:), Xo(:,:), Vo(:,:)
  integer(c_int32_t), allocatable, target :: at(:), xs(:)
  real(dp), allocatable, target :: surf_z(:), bot_z(:), epot(:), led0(:), led(:)
  real(dp), allocatable :: xf(:,:), vf(:,:)
  integer(c_int32_t), allocatable :: finst(:), fion(:)
  integer(c_int64_t) :: lt, c0, c1, crate
  integer, allocatable :: yrep(:)

  n = 4; nz = 4; nrep = 2; nimp = 2; nsteps = 1000
  first = 1; cnt = 0
  energy = 2000.0_dp; polar = 0.0_dp; azim = 0.0_dp; tempk = 300.0_dp; shell = 0.0_dp
  listfac = 0.0_dp
  prec = FIB_PREC_FP64
  frames = .false.
  sync_frames = .false.
  pend = .false.
  ranseed = -4242
  nargs = command_argument_count()
  k = 1
  do while (k <= nargs)
     call get_command_argument(k, arg)
     select case (trim(arg))
     case ('--mixed')
        prec = FIB_PREC_MIXED
     case ('--frames')
        frames = .true.
     case ('--sync-frames')
        frames = .true.
        sync_frames = .true.
     case ('--cells', '--nz', '--replicas', '--impacts', '--steps', '--seed', '--first', &
          '--count')
        k = k + 1
        call int_arg(trim(arg))
     case ('--energy', '--polar', '--azimuth', '--temp', '--shell', '--listfac')
        k = k + 1
        call real_arg(trim(arg))
     case default
        write(0, *) 'mdcfg_gpu: unknown argument ', trim(arg)
        stop 2
     end select
     k = k + 1
  end do
  if (n < 1 .or. nz < 1 .or. nrep < 1 .or. nimp < 1 .or. nsteps < 0) then
     write(0, *) 'mdcfg_gpu: sizes must be positive'
     stop 2
  end if
  if (cnt == 0) cnt = nrep - first + 1
  if (first < 1 .or. cnt < 1 .or. first + cnt - 1 > nrep) then
     write(0, *) 'mdcfg_gpu: --first/--count outside the replicas'
     stop 2
  end if

  ! ---- one slab (every replica starts from its own velocities) -----------
  call make_slab()
  allocate(desc(cnt), surf_z(cnt), bot_z(cnt), epot(cnt), led0(cnt), led(cnt), yrep(cnt))
  do r = 1, cnt
     desc(r)%natm = st%natm
     desc(r)%nsg = st%nsg
     desc(r)%nlj = st%nlj
     desc(r)%reserved = 0
     desc(r)%lb = st%lb
  end do
  surf_z = surf
  bot_z = 0.0_dp
  call fib_default_params(prm)
  if (listfac > 0.0_dp) prm%listfac = listfac
  rc = fib_create(prm, int(cnt, c_int32_t), desc, prec, h)
  if (rc /= FIB_OK) then
     write(0, *) 'mdcfg_gpu: fib_create failed, status ', rc
     stop 1
  end if
  allocate(Xo(st%natm,3), Vo(st%natm,3))
  do r = 1, nrep
     call thermal_velocities()
     if (r < first .or. r >= first + cnt) cycle
     if (r == first) st%V = V0
     call check(fib_set_state(h, int(r - first, c_int32_t), c_loc(X0), c_loc(V0), c_loc(at), &
          c_loc(xs), c_loc(X0)), 'fib_set_state')
  end do
  mp%dt_max = 1.0E-15_dp
  mp%dx_max = 0.05_dp * ang
  mp%remove_height = 10.0_dp * ang
  mp%shell_width = shell * ang
  mp%ttar = tempk
  mp%tau = 1.0E-13_dp
  mp%displaced_r = 1.5_dp * ang
  mp%use_graph = 1
  mp%reserved = 0
  call check(fib_md_begin(h, mp, surf_z, bot_z), 'fib_md_begin')
  call ledger(led0)
  print '(A,I3,A,I3,A,I3,A,I9,A,I9)', ' replicas', nrep, '  simulated', first, ' to', &
       first + cnt - 1, '  atoms per replica', st%natm, '  steps per impact', nsteps
  flush(6)
  if (frames) then
     call open_outputs(1, ios)
     if (ios /= 0) then
        write(0, *) 'mdcfg_gpu: cannot open data/mdrun2.xyz (is there a data/ folder?)'
        stop 1
     end if
     st%X = X0
     call writeatoms(s, st, 0_c_int64_t, 0.0_dp, host_temperature(st))
  end if

  ! ---- impacts ---------------------------------------------------------------
  th = polar * pi_ref() / 180.0_dp
  ph = azim * pi_ref() / 180.0_dp
  spd = ga_speed * sqrt(energy)
  v = [spd * sin(th) * cos(ph), spd * sin(th) * sin(ph), -spd * cos(th)]
  allocate(xf(3,cnt), vf(3,cnt), finst(cnt), fion(cnt))
  do m = 1, nimp
     do r = 1, nrep
        x = [ran1() * st%lb(1), ran1() * st%lb(2), surf + 5.0_dp * ang]
        if (r < first .or. r >= first + cnt) cycle
        q = r - first + 1
        xf(:,q) = x
        vf(:,q) = v
        finst(q) = int(q - 1, c_int32_t)
        fion(q) = int(m, c_int32_t)
     end do
     call system_clock(c0, crate)
     call check(fib_md_fire_many(h, int(cnt, c_int32_t), finst, fion, xf, vf), 'fib_md_fire_many')
     call check(fib_md_run_begin(h, int(nsteps, c_int32_t)), 'fib_md_run_begin')
     if (pend) then  ! the previous impact's frame, written while this one runs
        call check(fib_snapshot_wait(h, 0_c_int32_t, c_loc(Xo), c_loc(Vo)), 'fib_snapshot_wait')
        call write_frame(plt, ptime)
        pend = .false.
     end if
     call check(fib_md_run_end(h), 'fib_md_run')
     call system_clock(c1)
     call ledger(led)
     ymean = 0; fmean = 0
     do r = 1, cnt
        call check(fib_md_get_report(h, int(r - 1, c_int32_t), rep), 'fib_md_get_report')
        ymean = ymean + rep%removed
        fmean = fmean + rep%e_fired
        yrep(r) = int(rep%removed)
     end do
     ymean = ymean / cnt
     fmean = fmean / cnt / ev
     dmean = sum(led - led0) / cnt / ev
     dmax = maxval(abs(led - led0)) / ev
     call check(fib_md_get_report(h, 0_c_int32_t, rep), 'fib_md_get_report')
     print '(A,I5,A,F9.3,A,ES12.4,A,ES12.4,A,F12.3,A,ES12.4,A,ES12.4,A,F9.3)', ' impact', m, &
          '  yield', ymean, '  drift_eV', dmean, '  drift_max_eV', dmax, '  fired_eV', fmean, &
          '  time_s', rep%time, '  dt_s', rep%dt, '  wall_s', real(c1 - c0, dp) / real(crate, dp)
     print '(A,I5,*(1X,I0,1X,ES11.3))', ' replicas', m, (yrep(r), (led(r) - led0(r)) / ev, r = 1, cnt)
     flush(6)  ! two lines per impact, as it happens
     if (frames) then
        lt = int(m, c_int64_t) * nsteps
        if (sync_frames) then
           call check(fib_get_state(h, 0_c_int32_t, c_loc(Xo), c_loc(Vo), c_null_ptr), &
                'fib_get_state')
           call write_frame(lt, rep%time)
        else
           call check(fib_snapshot_begin(h, 0_c_int32_t, 0_c_int32_t), 'fib_snapshot_begin')
           pend = .true.
           plt = lt
           ptime = rep%time
        end if
     end if
  end do
  if (pend) then
     call check(fib_snapshot_wait(h, 0_c_int32_t, c_loc(Xo), c_loc(Vo)), 'fib_snapshot_wait')
     call write_frame(plt, ptime)
  end if
  if (frames) call close_outputs(1)
  rc = fib_destroy(h)
  print *, 'DONE'

contains

  subroutine check(code, what)
    integer(c_int32_t), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= FIB_OK) then
       write(0, *) 'mdcfg_gpu: ', what, ' failed, status ', code, ': ', fib_error_string(h)
       stop 1
    end if
  end subroutine check

  subroutine int_arg(name)
    character(len=*), intent(in) :: name
    integer :: val
    character(len=64) :: a
    call get_command_argument(k, a)
    read(a, *) val
    select case (name)
    case ('--cells');    n = val
    case ('--nz');       nz = val
    case ('--replicas'); nrep = val
    case ('--impacts');  nimp = val
    case ('--steps');    nsteps = val
    case ('--seed');     ranseed = -abs(val)
    case ('--first');    first = val
    case ('--count');    cnt = val
    end select
  end subroutine int_arg

  subroutine real_arg(name)
    character(len=*), intent(in) :: name
    real(dp) :: val
    character(len=64) :: a
    call get_command_argument(k, a)
    read(a, *) val
    select case (name)
    case ('--energy');  energy = val
    case ('--polar');   polar = val
    case ('--azimuth'); azim = val
    case ('--temp');    tempk = val
    case ('--shell');   shell = val
    case ('--listfac'); listfac = val
    end select
  end subroutine real_arg

  ! Diamond sites in the loop order x, y, z, basis; pinned bottom layers;
  ! the ions parked at random spots above the surface.
  subroutine make_slab()
    real(dp) :: base(8,3)
    integer :: i, j, l, b, q, nsg
    base(1,:) = [0.00_dp, 0.00_dp, 0.00_dp]
    base(2,:) = [0.00_dp, 0.50_dp, 0.50_dp]
    base(3,:) = [0.25_dp, 0.25_dp, 0.75_dp]
    base(4,:) = [0.25_dp, 0.75_dp, 0.25_dp]
    base(5,:) = [0.50_dp, 0.00_dp, 0.50_dp]
    base(6,:) = [0.50_dp, 0.50_dp, 0.00_dp]
    base(7,:) = [0.75_dp, 0.25_dp, 0.25_dp]
    base(8,:) = [0.75_dp, 0.75_dp, 0.75_dp]
    nsg = 8 * n * n * nz
    st%nsg = nsg
    st%nlj = nimp
    st%natm = nsg + nimp
    st%lb = [n * a_cfg, n * a_cfg, nz * a_cfg + vacuum]
    allocate(X0(st%natm,3), V0(st%natm,3), at(st%natm), xs(st%natm))
    allocate(st%X(st%natm,3), st%V(st%natm,3), st%mass(st%natm), st%atype(st%natm))
    q = 0
    do i = 0, n - 1
       do j = 0, n - 1
          do l = 0, nz - 1
             do b = 1, 8
                q = q + 1
                X0(q,:) = ([real(i, dp), real(j, dp), real(l, dp)] + base(b,:)) * a_cfg
             end do
          end do
       end do
    end do
    surf = maxval(X0(1:nsg,3))
    do q = nsg + 1, st%natm
       X0(q,:) = [ran1() * st%lb(1), ran1() * st%lb(2), surf + park]
    end do
    at(1:nsg) = 1
    at(nsg+1:) = 3
    xs = 0
    do q = 1, nsg
       if (X0(q,3) < 0.3_dp * a_cfg) xs(q) = q
    end do
    st%atype = at
    st%mass(1:nsg) = mass_si
    st%mass(nsg+1:) = mass_ga
    st%X = X0
  end subroutine make_slab

  ! Maxwell-Boltzmann at tempk (Box-Muller on ran1), mobile atoms without
  ! net momentum, pinned atoms and parked ions at rest.
  subroutine thermal_velocities()
    integer :: q, c
    real(dp) :: u1, u2, p(3), msum
    V0 = 0.0_dp
    p = 0.0_dp
    msum = 0.0_dp
    do q = 1, st%nsg
       if (xs(q) /= 0) cycle
       do c = 1, 3
          u1 = max(ran1(), 1.0E-300_dp)
          u2 = ran1()
          V0(q,c) = sqrt(-2.0_dp * log(u1)) * cos(2.0_dp * pi_ref() * u2) * &
               sqrt(kB * tempk / mass_si)
       end do
       p = p + mass_si * V0(q,:)
       msum = msum + mass_si
    end do
    do q = 1, st%nsg
       if (xs(q) == 0) V0(q,:) = V0(q,:) - p / msum
    end do
  end subroutine thermal_velocities

  ! frame of the first simulated replica from Xo, Vo (reference format)
  subroutine write_frame(step, t)
    integer(c_int64_t), intent(in) :: step
    real(dp), intent(in) :: t
    st%X = Xo
    st%V = Vo
    call writeatoms(s, st, step, t, host_temperature(st))
  end subroutine write_frame

  ! kinetic + potential + e_bath + e_removed - e_fired per replica (J)
  subroutine ledger(out)
    real(dp), intent(out) :: out(:)
    integer :: i
    call check(fib_md_energy(h, epot), 'fib_md_energy')
    do i = 1, cnt
       call check(fib_md_get_report(h, int(i - 1, c_int32_t), rep), 'fib_md_get_report')
       out(i) = rep%ke + epot(i) + rep%e_bath + rep%e_removed - rep%e_fired
    end do
  end subroutine ledger

end program mdcfg_gpu
