!------------------------------------------------------------------------------
! mdrun_gpu -- FIB_MD run on one MI355X: the reference's main program
! (main.f90:13-90) and tint (timeint.f90:36-120) with the force-and-integrate
! loop on the GPU through the C-ABI.
!
! Run it in a directory holding siga.in and a data/ folder, like the
! reference.  It writes si.dat, fwhm.dat, data/mdrun2.xyz, data/3dmol2.xyz
! and the reference's progress lines.  A frame due during an impact is an
! asynchronous snapshot (fib_snapshot_begin): its copy to the host runs
! beside the next steps, and it is written when the steps stop next, before
! anything else is written to the frame files.
!
! Usage: mdrun_gpu [--seed N] [--mixed]
!   --seed N  ranseed (the reference draws -(seconds*milliseconds) from the
!             wall clock, prms.f90:220,296; that stays the default)
!   --mixed   mixed-precision forces (FP64 is the default)
!------------------------------------------------------------------------------
program mdrun_gpu
  use iso_c_binding
  use fibgpu
  use fibhost
  implicit none

  type(siga_t) :: s
  type(state_t) :: st
  type(fib_params) :: prm
  type(fib_instance_desc) :: desc(1)
  type(fib_run_params) :: rp
  type(fib_step_ctl) :: ctl
  type(fib_step_report) :: rep
  type(c_ptr) :: h
  integer :: ios, k, nargs, prec, impact, vals(8), nt0, ions
  real(dp) :: restart_time
  integer(c_int32_t) :: rc
  integer(c_int64_t) :: lt
  real(dp) :: time, temp, xion(3), vion(3)
  real(dp), allocatable, target :: X(:,:), V(:,:)
  integer(c_int32_t), allocatable, target :: at(:), xs(:)
  real(dp), allocatable, target :: Xi(:,:)
  character(len=64) :: arg
  logical :: have_seed, pend
  integer(c_int64_t) :: plt
  real(dp) :: ptime, ptemp

  prec = FIB_PREC_FP64
  have_seed = .false.
  pend = .false.
  nargs = command_argument_count()
  k = 1
  do while (k <= nargs)
     call get_command_argument(k, arg)
     select case (trim(arg))
     case ('--seed')
        k = k + 1
        call get_command_argument(k, arg)
        read(arg, *) ranseed
        have_seed = .true.
     case ('--mixed')
        prec = FIB_PREC_MIXED
     case default
        write(0, *) 'mdrun_gpu: unknown argument ', trim(arg)
        stop 2
     end select
     k = k + 1
  end do

  ! ---- readprms (prms.f90:203-357) ----------------------------------------
  call read_siga('siga.in', s, ios)
  if (ios /= 0) then
     write(0, *) 'mdrun_gpu: cannot read siga.in'
     stop 1
  end if
  if (s%amorcrys /= 0 .or. (s%whichic /= 4 .and. s%whichic /= 0)) then
     write(0, *) 'mdrun_gpu: crystalline runs only (amorcrys 0), from scratch (whichic 4) ', &
          'or from a restart record (whichic 0)'
     stop 1
  end if
  if (.not. have_seed) then
     call date_and_time(values=vals)
     ranseed = -vals(7) * vals(8)
  end if
  call pureepi(s%nc, s%nlj, st)
  print *, 'readprms done'
  ! ---- initran, initdata (main.f90:40-48) -----------------------------------
  call initranions(s)
  print *, 'initran done'
  ! initdata / ic (data.f90:48-93): from scratch, or from a restart record
  nt0 = 0
  ions = 0
  restart_time = 0
  if (s%whichic == 4) then
     call initdata(s, st, ios)
  else
     call read_restart(s, st, nt0, restart_time, ions, ios)
  end if
  if (ios /= 0) then
     write(0, *) 'mdrun_gpu: cannot read si.dat / fwhm.dat / the restart record'
     stop 1
  end if
  print *, 'initdata done'

  ! ---- the GPU handle: lists, tables, state --------------------------------
  call fib_default_params(prm)
  desc(1)%natm = st%natm
  desc(1)%nsg = st%nsg
  desc(1)%nlj = st%nlj
  desc(1)%reserved = 0
  desc(1)%lb = st%lb
  rc = fib_create(prm, 1, desc, prec, h)
  if (rc /= FIB_OK) then
     write(0, *) 'mdrun_gpu: fib_create failed, status ', rc
     stop 1
  end if
  allocate(X(st%natm,3), V(st%natm,3), Xi(st%natm,3), at(st%natm), xs(st%natm))
  X = st%X
  V = st%V
  Xi = st%Xi
  at = st%atype
  xs = st%xsides
  call check(fib_set_state(h, 0, c_loc(X), c_loc(V), c_loc(at), c_loc(xs), c_loc(Xi)), &
       'fib_set_state')
  call check(fib_build_nlist(h), 'fib_build_nlist')
  print *, 'si_nlist done'
  rp%ts_i = s%ts_i;  rp%ts_r = s%ts_r
  rp%tau_i = s%tau_i;  rp%tau_r = s%tau_r
  rp%ttar = s%ttar1;  rp%teps = s%teps;  rp%ke_limit_1 = s%ke_limit_1
  rp%dti = s%dti;  rp%sidewidth = s%sidewidth
  rp%outz = s%outz;  rp%atomz = s%atomz;  rp%ionz = s%ionz
  rp%ntlist_i = s%ntlist_i;  rp%ntlist_r = s%ntlist_r
  rp%atlist_i = s%atlist_i;  rp%atlist_r = s%atlist_r
  rp%dslist_i = s%dslist_i;  rp%dslist_r = s%dslist_r
  rp%atm_out_i = s%atm_out_i;  rp%atm_out_r = s%atm_out_r;  rp%res_out = s%res_out
  rp%reserved = 0
  call check(fib_set_run(h, 0, rp), 'fib_set_run')
  call open_outputs(s%atm_out_i, ios)
  if (ios /= 0) then
     write(0, *) 'mdrun_gpu: cannot open data/mdrun2.xyz (is there a data/ folder?)'
     stop 1
  end if
  write(*, *) 'DONE INITIALIZING'
  write(*, *) '--------------------------------'

  ! ---- tint (timeint.f90:36-120) --------------------------------------------
  lt = nt0
  time = restart_time
  temp = host_temperature(st)
  if (lt == 0) call writeatoms(s, st, lt, time, temp)
  call check(fib_tint_init(h, lt, time), 'fib_tint_init')

  ctl%max_steps = 0
  ctl%reserved = 0
  if (nt0 == 0) then
     print *, 'Initializing target temperature. . .'
     call check(fib_ti_begin(h, FIB_TI_EQUILIBRATE, 0), 'fib_ti_begin')
     ctl%whichctrl = FIB_TI_EQUILIBRATE
     ctl%stop_on_output = 0
     do
        call check(fib_step(h, ctl, rep), 'fib_step')
        if (rep%ended /= 0) exit
     end do
     temp = rep%temp
     lt = rep%lt
     time = rep%time
     print *, 'Equalized to temperature ', temp
  end if

  do impact = ions + 1, st%nlj
     call check(fib_ti_begin(h, FIB_TI_IMPACT, impact), 'fib_ti_begin')
     call knock_draw(s, st, xion, vion)
     print *, 'KNOCK    --=>*<=--    ion ID: ', st%nsg + impact
     call check(fib_knock(h, impact, xion, vion), 'fib_knock')
     ctl%whichctrl = FIB_TI_IMPACT
     ctl%stop_on_output = 1
     do
        call check(fib_step(h, ctl, rep), 'fib_step')
        call flush_frame()
        lt = rep%lt
        time = rep%time
        temp = rep%temp
        ! iostuff (out.f90:129-150): restart record, then the frame
        if (rep%output_due /= 0 .and. s%res_out > 0) then
           if (mod(lt, int(s%res_out, c_int64_t)) == 0) then
              call fetch()
              ! the record carries the clock of this step, as onestep sets
              ! restart_time = time every step (timeint.f90:165-166)
              call write_restart(s, st, int(lt), time, impact)
           end if
        end if
        if (rep%atoms_due /= 0 .and. s%atm_out_i > 0) then
           call check(fib_snapshot_begin(h, 0_c_int32_t, 0_c_int32_t), 'fib_snapshot_begin')
           pend = .true.
           plt = lt
           ptime = time
           ptemp = temp
        end if
        if (rep%ended /= 0) exit
     end do
     call flush_frame()
     call fetch()
     call writeatoms(s, st, lt, time, temp)
  end do

  call close_outputs(s%atm_out_i)
  rc = fib_destroy(h)

contains

  subroutine check(code, what)
    integer(c_int32_t), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= FIB_OK) then
       write(0, *) 'mdrun_gpu: ', what, ' failed, status ', code, ': ', fib_error_string(h)
       stop 1
    end if
  end subroutine check

  ! the frame whose snapshot is pending, once its copy has arrived
  subroutine flush_frame()
    if (.not. pend) return
    call check(fib_snapshot_wait(h, 0_c_int32_t, c_loc(X), c_loc(V)), 'fib_snapshot_wait')
    st%X = X
    st%V = V
    call writeatoms(s, st, plt, ptime, ptemp)
    pend = .false.
  end subroutine flush_frame

  ! current X, V from the device into the host state (original atom order)
  subroutine fetch()
    call check(fib_get_state(h, 0, c_loc(X), c_loc(V), c_null_ptr), 'fib_get_state')
    st%X = X
    st%V = V
  end subroutine fetch

end program mdrun_gpu
