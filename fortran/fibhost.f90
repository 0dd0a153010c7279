!------------------------------------------------------------------------------
! fibhost.f90 -- host side of the MI355X FIB_MD path, in Fortran.
!
! Everything the reference does on the host around its force-and-integrate
! loop, restated: the siga.in reader (prms.f90:221-277), the pureepi lattice
! with its si.dat text round trip (prms.f90:484-557, data.f90:131-143), the
! ran1 generator (ran1.f:10-36, Numerical Recipes' minimal-standard generator
! with a Bays-Durham shuffle), the Gaussian impact spot with its fwhm.dat
! round trip (data.f90:263-285, 146-161), initial ions and velocities
! (data.f90:169-260), the knock position/velocity (knock.f90:21-55) and the
! data/ writers (out.f90:63-100).  The time integration itself runs on the
! GPU through the C-ABI (module fibgpu).
!
! Arithmetic is double precision throughout, as the reference's
! -fdefault-real-8 build; expressions keep the reference's operation order.
!------------------------------------------------------------------------------
module fibhost
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: dp = kind(1.0d0)

  ! the 42 siga.in values in read order (prms.f90:223-275)
  type, public :: siga_t
     integer  :: amorcrys = 0, whichic = 4, startimpact = 0
     real(dp) :: eV = 0
     integer  :: nc(3) = 0, npc(3) = 1
     integer  :: nlj = 0, nt = 0
     real(dp) :: ts_i = 0, ts_r = 0
     real(dp) :: ttar1 = 0, ke_limit_1 = 0, ke_limit_2 = 0
     real(dp) :: tau_i = 0, tau_r = 0, teps = 0
     real(dp) :: uo = 0
     integer  :: ntlist_i = 1, ntlist_r = 1, atlist_i = 1, atlist_r = 1, dslist_i = 1, dslist_r = 1
     integer  :: atm_out_i = 0, atm_out_r = 0, res_out = 0, eng_out = 0, tmp_out = 0
     integer  :: nrand = 0
     real(dp) :: outz(4) = 0, atomz(2) = 0
     real(dp) :: ionz = 0, knockz = 0, sidewidth = 0, dti = 0
     real(dp) :: fwhm = 0
     real(dp) :: phiz = 0, phixy = 0
  end type siga_t

  ! constants of prms.f90:172-216 used on the host
  real(dp), parameter, public :: mass_si = 46.637063E-27_dp, mass_ge = 120.6049350E-27_dp, &
       mass_ga = 1.157777417E-25_dp
  real(dp), parameter, public :: sigma_si = 2.0951E-10_dp, kB = 1.381E-23_dp
  real(dp), parameter, public :: a_lat = 5.431073E-10_dp   ! pureepi (prms.f90:494)

  ! ran1 state: the generator's seed word, the shuffle table and its output
  integer, public :: ranseed = -1
  integer :: sh_y = 0
  integer :: sh_v(32) = 0

  ! run state
  type, public :: state_t
     integer :: natm = 0, nsg = 0, nlj = 0
     real(dp) :: lb(3) = 0
     real(dp), allocatable :: X(:,:), Xi(:,:), V(:,:), mass(:), xrand(:,:)
     integer, allocatable :: atype(:), xsides(:)
  end type state_t

  public :: read_siga, ran1, pureepi, initranions, initdata, knock_draw, writeatoms, &
       open_outputs, close_outputs, host_temperature, proc_of, pi_ref, write_restart, &
       read_restart

  integer, parameter :: atm_unit = 41, snap_unit = 42, res_unit = 12

contains

  real(dp) function pi_ref()
    pi_ref = 4.0_dp * atan(1.0_dp)   ! prms.f90:172
  end function pi_ref

  ! --------------------------------------------------------------------------
  ! siga.in: one list-directed read per record, so a value may be followed by
  ! any comment text, and blank records are skipped (prms.f90:221-277).
  subroutine read_siga(fname, s, ios)
    character(len=*), intent(in) :: fname
    type(siga_t), intent(out) :: s
    integer, intent(out) :: ios
    integer :: u
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) return
    read(u, *, iostat=ios) s%amorcrys;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%whichic;    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%startimpact; if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%eV;         if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%nc;         if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%npc;        if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%nlj;        if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%nt;         if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ts_i;       if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ts_r;       if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ttar1;      if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ke_limit_1; if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ke_limit_2; if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%tau_i;      if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%tau_r;      if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%teps;       if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%uo;         if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ntlist_i;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ntlist_r;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%atlist_i;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%atlist_r;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%dslist_i;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%dslist_r;   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%atm_out_i;  if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%atm_out_r;  if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%res_out;    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%eng_out;    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%tmp_out;    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%nrand;      if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%outz(1);    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%outz(2);    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%outz(3);    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%outz(4);    if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%atomz(1);   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%atomz(2);   if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%ionz;       if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%knockz;     if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%sidewidth;  if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%dti;        if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%fwhm;       if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%phiz;       if (ios /= 0) goto 9
    read(u, *, iostat=ios) s%phixy
9   close(u)
  end subroutine read_siga

  ! --------------------------------------------------------------------------
  ! ran1: uniform deviate in (0,1).  Lehmer generator x <- 16807 x mod
  ! (2^31 - 1), evaluated with Schrage's factorisation (q = 127773,
  ! r = 2836), whose outputs pass through a 32-entry Bays-Durham shuffle;
  ! a call with ranseed <= 0 (re)initialises it, skipping 8 warm-up draws.
  ! Outputs are capped below 1 at 1 - 1.2e-7.
  real(dp) function ran1()
    integer, parameter :: im = 2147483647, ia = 16807, iq = 127773, ir = 2836, ntab = 32
    integer, parameter :: ndiv = 1 + (im - 1) / ntab
    real(dp), parameter :: am = 1.0_dp / im, rnmx = 1.0_dp - 1.2E-7_dp
    integer :: j, k
    if (ranseed <= 0 .or. sh_y == 0) then
       ranseed = max(-ranseed, 1)
       do j = ntab + 8, 1, -1
          call lehmer()
          if (j <= ntab) sh_v(j) = ranseed
       end do
       sh_y = sh_v(1)
    end if
    call lehmer()
    j = 1 + sh_y / ndiv
    sh_y = sh_v(j)
    sh_v(j) = ranseed
    ran1 = min(am * sh_y, rnmx)
  contains
    subroutine lehmer()
      k = ranseed / iq
      ranseed = ia * (ranseed - k * iq) - ir * k
      if (ranseed < 0) ranseed = ranseed + im
    end subroutine lehmer
  end function ran1

  ! --------------------------------------------------------------------------
  ! Diamond lattice, growth direction z (prms.f90:484-557): 8-site basis,
  ! cells in x, y, z order; positions are written to si.dat as E20.10 text
  ! (and read back by initdata).  Box: a*nc in x, y and 10*a*nc3 in z.  Side
  ! atoms (on the min/max x or y planes) get xsides(i) = i.
  subroutine pureepi(nc, nlj, st)
    integer, intent(in) :: nc(3), nlj
    type(state_t), intent(inout) :: st
    real(dp) :: base(8,3), ax(3), maxx, maxy
    real(dp), allocatable :: P(:,:)
    integer :: i, j, k, l, m, u, n
    ax = a_lat
    base(1,:) = [0.0_dp,        0.0_dp,        0.0_dp]
    base(2,:) = [0.0_dp,        0.50_dp*ax(2), 0.50_dp*ax(3)]
    base(3,:) = [0.250_dp*ax(1), 0.250_dp*ax(2), 0.750_dp*ax(3)]
    base(4,:) = [0.250_dp*ax(1), 0.750_dp*ax(2), 0.250_dp*ax(3)]
    base(5,:) = [0.50_dp*ax(1),  0.0_dp*ax(2),   0.50_dp*ax(3)]
    base(6,:) = [0.50_dp*ax(1),  0.50_dp*ax(2),  0.0_dp*ax(3)]
    base(7,:) = [0.750_dp*ax(1), 0.250_dp*ax(2), 0.250_dp*ax(3)]
    base(8,:) = [0.750_dp*ax(1), 0.750_dp*ax(2), 0.750_dp*ax(3)]
    n = 8 * nc(1) * nc(2) * nc(3)
    allocate(P(n,3))
    m = 0
    do i = 1, nc(1)
       do j = 1, nc(2)
          do k = 1, nc(3)
             do l = 1, 8
                m = m + 1
                P(m,1) = base(l,1) + ax(1) * real(i - 1, dp)
                P(m,2) = base(l,2) + ax(2) * real(j - 1, dp)
                P(m,3) = base(l,3) + ax(3) * real(k - 1, dp)
             end do
          end do
       end do
    end do
    open(newunit=u, file='si.dat', status='replace', action='write')
    do i = 1, n
       write(u, "(3E20.10,I10)") P(i,:), 1
    end do
    close(u)
    st%lb(1) = ax(1) * nc(1)
    st%lb(2) = ax(2) * nc(2)
    st%lb(3) = ax(3) * nc(3) * 10
    st%nsg = n
    st%nlj = nlj
    st%natm = n + nlj
    maxx = maxval(P(:,1))
    maxy = maxval(P(:,2))
    allocate(st%xsides(st%natm))
    st%xsides = 0
    do i = 1, n
       if (P(i,1) < 1.0E-10_dp .or. P(i,2) < 1.0E-10_dp .or. P(i,1) > maxx - 1.0E-10_dp &
            .or. P(i,2) > maxy - 1.0E-10_dp) st%xsides(i) = i
    end do
  end subroutine pureepi

  ! --------------------------------------------------------------------------
  ! Gaussian impact spot (data.f90:263-285): Box-Muller pairs with the two
  ! uniforms swapped between the cosine and sine branch, sigma = fwhm/2.35
  ! nm, written to fwhm.dat as E12.6 text.
  subroutine initranions(s)
    type(siga_t), intent(in) :: s
    real(dp) :: u1, u2, g1, g2, pi
    integer :: i, u
    pi = pi_ref()
    open(newunit=u, file='fwhm.dat', status='replace', action='write')
    do i = 1, s%nrand
       u1 = ran1()
       u2 = ran1()
       g1 = sqrt(-2 * log(u1)) * cos(2 * pi * u2) * s%fwhm / 2.35_dp * 1.E-9_dp
       g2 = sqrt(-2 * log(u2)) * sin(2 * pi * u1) * s%fwhm / 2.35_dp * 1.E-9_dp
       write(u, "(E12.6,A1,E12.6)") g1, ' ', g2
    end do
    close(u)
  end subroutine initranions

  ! --------------------------------------------------------------------------
  ! Initial state of a from-scratch run, whichic = 4 (data.f90:48-260): SW
  ! atoms from si.dat, ions in the storage layer at ionz with lateral
  ! separation >= sigma_Si, impact points from fwhm.dat, masses, uniform
  ! velocities in [-Uo/2, Uo/2) with the mean momentum per atom removed.
  subroutine initdata(s, st, ios)
    type(siga_t), intent(in) :: s
    type(state_t), intent(inout) :: st
    integer, intent(out) :: ios
    integer :: i, j, u, t
    real(dp) :: d(2), mm(3)
    allocate(st%X(st%natm,3), st%Xi(st%natm,3), st%V(st%natm,3), st%mass(st%natm), &
         st%atype(st%natm), st%xrand(s%nrand,2))
    st%V = 0
    ! si.dat, read back as text (the positions carry E20.10 rounding)
    open(newunit=u, file='si.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do i = 1, st%nsg
       read(u, *, iostat=ios) st%X(i,:), t
       if (ios /= 0) then
          close(u)
          return
       end if
       st%atype(i) = t
    end do
    close(u)
    ! ions in the storage layer; a too-close ion is redrawn against the same
    ! partner until it clears it, then checking goes on with the next one
    do i = st%nsg + 1, st%natm
       st%X(i,1) = st%lb(1) * ran1()
       st%X(i,2) = st%lb(2) * ran1()
       st%X(i,3) = s%ionz * st%lb(3) / 10.0_dp
       st%atype(i) = 3
    end do
    do i = st%nsg + 1, st%natm
       do j = st%nsg + 1, i - 1
          do
             d = st%X(i,1:2) - st%X(j,1:2)
             if (.not. (sqrt(d(1)**2 + d(2)**2) < sigma_si)) exit
             st%X(i,1) = st%lb(1) * ran1()
             st%X(i,2) = st%lb(2) * ran1()
          end do
       end do
    end do
    st%Xi = st%X
    open(newunit=u, file='fwhm.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do i = 1, s%nrand
       read(u, *, iostat=ios) st%xrand(i,:)
       if (ios /= 0) then
          close(u)
          return
       end if
    end do
    close(u)
    do i = 1, st%natm
       select case (st%atype(i))
       case (1); st%mass(i) = mass_si
       case (2); st%mass(i) = mass_ge
       case default; st%mass(i) = mass_ga
       end select
    end do
    if (s%amorcrys == 0) then
       mm = 0
       do i = 1, st%nsg
          st%V(i,1) = s%uo * (ran1() - 0.5_dp)
          st%V(i,2) = s%uo * (ran1() - 0.5_dp)
          st%V(i,3) = s%uo * (ran1() - 0.5_dp)
          mm = mm + st%mass(i) * st%V(i,:)
       end do
       mm = mm / st%nsg
       do i = 1, st%nsg
          st%V(i,:) = st%V(i,:) - mm / st%mass(i)
       end do
    end if
    ios = 0
  end subroutine initdata

  ! --------------------------------------------------------------------------
  ! Position and velocity of ion knc (knock.f90:21-55).  The draw
  ! NINT(ran1*Nrand) can be 0 (one past the table in the reference); it is
  ! taken as 1 here.
  subroutine knock_draw(s, st, xion, vion)
    type(siga_t), intent(in) :: s
    type(state_t), intent(in) :: st
    real(dp), intent(out) :: xion(3), vion(3)
    real(dp) :: vk, dr, dx, dy, pi
    integer :: ri
    pi = pi_ref()
    vk = 1663.6334_dp * sqrt(s%eV)
    dr = tan(s%phiz * pi / 180) * (s%knockz * st%lb(3) / 10 - st%lb(3) / 10)
    dx = dr * cos(s%phixy * pi / 180) * sin(s%phiz * pi / 180)
    dy = dr * sin(s%phixy * pi / 180) * sin(s%phiz * pi / 180)
    ri = nint(ran1() * s%nrand)
    ri = max(ri, 1)
    vion(1) = vk * cos(s%phixy * pi / 180) * sin(s%phiz * pi / 180)
    vion(2) = vk * sin(s%phixy * pi / 180) * sin(s%phiz * pi / 180)
    vion(3) = -1.0_dp * vk * cos(s%phiz * pi / 180)
    xion(1) = st%xrand(ri,1) + st%lb(1) / 2.0_dp - dx
    xion(2) = st%xrand(ri,2) + st%lb(2) / 2.0_dp - dy
    xion(3) = s%knockz * st%lb(3) / 10.0_dp
  end subroutine knock_draw

  ! --------------------------------------------------------------------------
  ! 0-based cell of the reference's processor grid (par.f90:144-160) that
  ! holds position x: what the reference prints as the atom's rank.
  integer function proc_of(s, lb, x)
    type(siga_t), intent(in) :: s
    real(dp), intent(in) :: lb(3), x(3)
    real(dp) :: w(3)
    integer :: c(3), l
    w(1) = (lb(1) + 1.E-10_dp) / real(s%npc(1), dp)
    w(2) = (lb(2) + 1.E-10_dp) / real(s%npc(2), dp)
    w(3) = (lb(3) / 10.0_dp + 1.E-10_dp) / real(s%npc(3), dp)
    do l = 1, 3
       c(l) = int(abs(x(l)) / w(l)) + 1
       if (c(l) > s%npc(l)) c(l) = s%npc(l)
    end do
    proc_of = c(1) - 1 + (c(2) - 1) * s%npc(1) + (c(3) - 1) * s%npc(1) * s%npc(2)
  end function proc_of

  ! temperature (temp.f90:183-205): sum m v^2 over non-Ga atoms / (3 Nsg kB)
  real(dp) function host_temperature(st)
    type(state_t), intent(in) :: st
    integer :: i
    real(dp) :: sm
    sm = 0
    do i = 1, st%natm
       if (st%atype(i) < 3) sm = sm + st%mass(i) * (st%V(i,1)**2 + st%V(i,2)**2 + st%V(i,3)**2)
    end do
    host_temperature = sm / 3.0_dp / st%nsg / kB
  end function host_temperature

  ! --------------------------------------------------------------------------
  ! data/ outputs (initio, out.f90:111-125): mdrun2.xyz is opened once per
  ! run when the output cadence is positive.
  subroutine open_outputs(atm_out, ios)
    integer, intent(in) :: atm_out
    integer, intent(out) :: ios
    ios = 0
    if (atm_out > 0) open(atm_unit, file='data/mdrun2.xyz', iostat=ios)
  end subroutine open_outputs

  subroutine close_outputs(atm_out)
    integer, intent(in) :: atm_out
    if (atm_out > 0) close(atm_unit)
  end subroutine close_outputs

  ! One frame (writeatoms, out.f90:63-100): appended to data/mdrun2.xyz,
  ! data/3dmol2.xyz replaced, and a progress line on stdout.  Species 2
  ! keeps the previous row's label, as in the reference, which has no label
  ! for it; KE uses the reference's J -> eV factor 6.252e18.
  subroutine writeatoms(s, st, lt, time, temp)
    type(siga_t), intent(in) :: s
    type(state_t), intent(in) :: st
    integer(c_int64_t), intent(in) :: lt
    real(dp), intent(in) :: time, temp
    character(len=2) :: lab
    real(dp) :: coef, ke
    integer :: i
    write(atm_unit, "(I9)") st%natm
    write(atm_unit, "(1A12,F12.5,3A12,2A8,3A16,A8,F14.5)") 'Time (ps) = ', time * 1e12_dp, &
         'x (A)', 'y (A)', 'z (A)', 'type', 'proc#', 'KE (eV)'
    open(snap_unit, file='data/3dmol2.xyz', status='replace')
    write(snap_unit, "(I9)") st%natm
    write(snap_unit, "(5A12,F12.5)") 'type', 'x (A)', 'y (A)', 'z (A)', 'Time (ps) = ', &
         time * 1e12_dp
    coef = 6.252e18_dp
    lab = '  '
    do i = 1, st%natm
       if (st%atype(i) == 1) then
          lab = 'Si'
       else if (st%atype(i) == 3) then
          lab = 'Ga'
       end if
       ke = 0.5_dp * st%mass(i) * (st%V(i,1)**2 + st%V(i,2)**2 + st%V(i,3)**2) * coef
       write(atm_unit, "(3F12.4,A8,I8,1ES14.3)") st%X(i,:) * 1e10_dp, lab, &
            proc_of(s, st%lb, st%X(i,:)), ke
       write(snap_unit, "(A8,3F10.4)") lab, st%X(i,:) * 1e10_dp
    end do
    close(snap_unit)
    write(*, "(I10,F10.4,F10.2)") lt, time * 1e12_dp, temp
  end subroutine writeatoms

  ! --------------------------------------------------------------------------
  ! Restart records (writerestart, out.f90:34-59): D/restart.in_old and
  ! D/restart.in<impact:6>_<lt:9>, unformatted sequential, records
  ! (lt, restart_time), (X, V, mass, atype), (impact) -- as the reference
  ! writes them, restart_time being the simulated clock of step lt (onestep
  ! sets restart_time = time every step, timeint.f90:165-166).
  ! Beside each goes D/restart.rng<same suffix> with what the reference does
  ! not keep: the ran1 state and the impact spots, so that a restart carries
  ! on the random sequence where the run left it.
  subroutine write_restart(s, st, lt, restart_time, impact)
    type(siga_t), intent(in) :: s
    type(state_t), intent(in) :: st
    integer, intent(in) :: lt, impact
    real(dp), intent(in) :: restart_time
    character(len=64) :: fn(2), fr(2)
    integer :: k
    fn(1) = 'D/restart.in_old'
    fr(1) = 'D/restart.rng_old'
    write(fn(2), "('D/restart.in',I6.6,'_',I9.9)") impact, lt
    write(fr(2), "('D/restart.rng',I6.6,'_',I9.9)") impact, lt
    do k = 1, 2
       open(res_unit, file=trim(fn(k)), form='unformatted', status='replace')
       write(res_unit) lt, restart_time
       write(res_unit) st%X, st%V, st%mass, st%atype
       write(res_unit) impact
       close(res_unit)
       open(res_unit, file=trim(fr(k)), form='unformatted', status='replace')
       write(res_unit) ranseed, sh_y, sh_v
       write(res_unit) s%nrand, st%xrand
       close(res_unit)
    end do
  end subroutine write_restart

  ! Initial state of a restarted run, whichic = 0 (ic, data.f90:71-93): Xi
  ! and species from si.dat, then X, V, masses, species, lt, the clock and
  ! the number of ions already fired from the restart record
  ! (D/restart.in_old for startimpact 0, else the first D/restart.in
  ! <startimpact:6>* file, found through F/restart_list_000.dat as the
  ! reference does); ions keep their stored positions as sites, unfired ions
  ! rest; impact spots from fwhm.dat.  With the record's companion
  ! D/restart.rng* present, the ran1 state and the impact spots come from it.
  subroutine read_restart(s, st, nt0, restart_time, ions, ios)
    type(siga_t), intent(in) :: s
    type(state_t), intent(inout) :: st
    integer, intent(out) :: nt0, ions, ios
    real(dp), intent(out) :: restart_time
    character(len=28) :: fn
    character(len=64) :: fr, cmd
    integer :: i, u, t, nr
    logical :: have
    allocate(st%X(st%natm,3), st%Xi(st%natm,3), st%V(st%natm,3), st%mass(st%natm), &
         st%atype(st%natm), st%xrand(s%nrand,2))
    open(newunit=u, file='si.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do i = 1, st%nsg
       read(u, *, iostat=ios) st%Xi(i,:), t
       if (ios /= 0) then
          close(u)
          return
       end if
       st%atype(i) = t
    end do
    close(u)
    if (s%startimpact == 0) then
       fn = 'D/restart.in_old'
    else
       write(cmd, "('ls D/restart.in',I6.6,'* > F/restart_list_000.dat')") s%startimpact
       call execute_command_line(trim(cmd), exitstat=ios)
       if (ios /= 0) return
       open(newunit=u, file='F/restart_list_000.dat', status='old', action='read', iostat=ios)
       if (ios /= 0) return
       read(u, "(A28)", iostat=ios) fn
       close(u)
       if (ios /= 0) return
    end if
    open(res_unit, file=trim(fn), form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    read(res_unit, iostat=ios) nt0, restart_time
    if (ios == 0) read(res_unit, iostat=ios) st%X, st%V, st%mass, st%atype
    if (ios == 0) read(res_unit, iostat=ios) ions
    close(res_unit)
    if (ios /= 0) return
    st%Xi(st%nsg+1:st%natm,:) = st%X(st%nsg+1:st%natm,:)
    st%V(st%nsg+ions+1:st%natm,:) = 0
    open(newunit=u, file='fwhm.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) return
    do i = 1, s%nrand
       read(u, *, iostat=ios) st%xrand(i,:)
       if (ios /= 0) then
          close(u)
          return
       end if
    end do
    close(u)
    ! the companion file: same name with 'rng' for 'in'
    fr = 'D/restart.rng' // trim(fn(13:))
    inquire(file=trim(fr), exist=have)
    if (have) then
       open(res_unit, file=trim(fr), form='unformatted', status='old', action='read', iostat=ios)
       if (ios /= 0) return
       read(res_unit, iostat=ios) ranseed, sh_y, sh_v
       if (ios == 0) read(res_unit, iostat=ios) nr
       if (ios == 0 .and. nr == s%nrand) then
          backspace(res_unit)
          read(res_unit, iostat=ios) nr, st%xrand
       end if
       close(res_unit)
       if (ios /= 0) return
    end if
    ios = 0
  end subroutine read_restart

end module fibhost
