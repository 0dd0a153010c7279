!------------------------------------------------------------------------------
! fibhost_capi.f90 -- C entry points of the Fortran host routines (module
! fibhost), so that the CPU tests can drive the parser, the generators and
! the writers through ctypes.  Library: libfibhost.so (no GPU code).
!------------------------------------------------------------------------------
module fibhost_capi
  use iso_c_binding
  use fibhost
  implicit none
  type(siga_t), save :: s
  type(state_t), save :: st

contains

  subroutine cstr(p, f)
    character(kind=c_char), intent(in) :: p(*)
    character(len=:), allocatable, intent(out) :: f
    integer :: n
    n = 0
    do while (p(n + 1) /= c_null_char)
       n = n + 1
    end do
    allocate(character(len=n) :: f)
    f = transfer(p(1:n), f)
  end subroutine cstr

  ! siga.in -> the 46 numbers of its 42 records, in read order
  integer(c_int) function fh_read_siga(path, v) bind(C, name="fh_read_siga")
    character(kind=c_char), intent(in) :: path(*)
    real(c_double), intent(out) :: v(46)
    character(len=:), allocatable :: f
    integer :: ios
    call cstr(path, f)
    call read_siga(f, s, ios)
    fh_read_siga = ios
    if (ios /= 0) return
    v = [real(s%amorcrys, dp), real(s%whichic, dp), real(s%startimpact, dp), s%eV, &
         real(s%nc, dp), real(s%npc, dp), real(s%nlj, dp), real(s%nt, dp), s%ts_i, s%ts_r, &
         s%ttar1, s%ke_limit_1, s%ke_limit_2, s%tau_i, s%tau_r, s%teps, s%uo, &
         real(s%ntlist_i, dp), real(s%ntlist_r, dp), real(s%atlist_i, dp), &
         real(s%atlist_r, dp), real(s%dslist_i, dp), real(s%dslist_r, dp), &
         real(s%atm_out_i, dp), real(s%atm_out_r, dp), real(s%res_out, dp), &
         real(s%eng_out, dp), real(s%tmp_out, dp), real(s%nrand, dp), s%outz, s%atomz, &
         s%ionz, s%knockz, s%sidewidth, s%dti, s%fwhm, s%phiz, s%phixy]
  end function fh_read_siga

  subroutine fh_set_seed(seed) bind(C, name="fh_set_seed")
    integer(c_int), value :: seed
    ranseed = seed
  end subroutine fh_set_seed

  integer(c_int) function fh_get_seed() bind(C, name="fh_get_seed")
    fh_get_seed = ranseed
  end function fh_get_seed

  real(c_double) function fh_ran1() bind(C, name="fh_ran1")
    fh_ran1 = ran1()
  end function fh_ran1

  ! readprms + initran + initdata of a from-scratch run in the current
  ! directory (writes si.dat and fwhm.dat); returns natm, nsg and the box
  integer(c_int) function fh_setup(path, seed, natm, nsg, lb) bind(C, name="fh_setup")
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: seed
    integer(c_int), intent(out) :: natm, nsg
    real(c_double), intent(out) :: lb(3)
    character(len=:), allocatable :: f
    integer :: ios
    call cstr(path, f)
    call read_siga(f, s, ios)
    fh_setup = ios
    if (ios /= 0) return
    ranseed = seed
    if (allocated(st%X)) deallocate(st%X, st%Xi, st%V, st%mass, st%xrand, st%atype)
    if (allocated(st%xsides)) deallocate(st%xsides)
    call pureepi(s%nc, s%nlj, st)
    call initranions(s)
    call initdata(s, st, ios)
    fh_setup = ios
    natm = st%natm
    nsg = st%nsg
    lb = st%lb
  end function fh_setup

  subroutine fh_get_state(X, V, atype, xsides, Xi, xrand) bind(C, name="fh_get_state")
    real(c_double), intent(out) :: X(st%natm,3), V(st%natm,3), Xi(st%natm,3)
    integer(c_int), intent(out) :: atype(st%natm), xsides(st%natm)
    real(c_double), intent(out) :: xrand(s%nrand,2)
    X = st%X
    V = st%V
    Xi = st%Xi
    atype = st%atype
    xsides = st%xsides
    xrand = st%xrand
  end subroutine fh_get_state

  subroutine fh_set_xv(X, V) bind(C, name="fh_set_xv")
    real(c_double), intent(in) :: X(st%natm,3), V(st%natm,3)
    st%X = X
    st%V = V
  end subroutine fh_set_xv

  subroutine fh_set_atype(atype) bind(C, name="fh_set_atype")
    integer(c_int), intent(in) :: atype(st%natm)
    st%atype = atype
  end subroutine fh_set_atype

  subroutine fh_knock(xion, vion) bind(C, name="fh_knock")
    real(c_double), intent(out) :: xion(3), vion(3)
    call knock_draw(s, st, xion, vion)
  end subroutine fh_knock

  real(c_double) function fh_temperature() bind(C, name="fh_temperature")
    fh_temperature = host_temperature(st)
  end function fh_temperature

  integer(c_int) function fh_open_outputs(atm_out) bind(C, name="fh_open_outputs")
    integer(c_int), value :: atm_out
    integer :: ios
    call open_outputs(atm_out, ios)
    fh_open_outputs = ios
  end function fh_open_outputs

  subroutine fh_close_outputs(atm_out) bind(C, name="fh_close_outputs")
    integer(c_int), value :: atm_out
    call close_outputs(atm_out)
  end subroutine fh_close_outputs

  subroutine fh_write_restart(lt, restart_time, impact) bind(C, name="fh_write_restart")
    integer(c_int), value :: lt, impact
    real(c_double), value :: restart_time
    call write_restart(s, st, lt, restart_time, impact)
  end subroutine fh_write_restart

  ! whichic-0 initial state from the restart record (the state of fh_setup
  ! is replaced); returns the iostat
  integer(c_int) function fh_read_restart(nt0, restart_time, ions) &
       bind(C, name="fh_read_restart")
    integer(c_int), intent(out) :: nt0, ions
    real(c_double), intent(out) :: restart_time
    integer :: ios, n0, ni
    real(dp) :: rt
    if (allocated(st%X)) deallocate(st%X, st%Xi, st%V, st%mass, st%xrand, st%atype)
    call read_restart(s, st, n0, rt, ni, ios)
    nt0 = n0
    ions = ni
    restart_time = rt
    fh_read_restart = ios
  end function fh_read_restart

  subroutine fh_writeatoms(lt, time, temp) bind(C, name="fh_writeatoms")
    integer(c_int64_t), value :: lt
    real(c_double), value :: time, temp
    call writeatoms(s, st, lt, time, temp)
    flush(6)
  end subroutine fh_writeatoms

end module fibhost_capi
