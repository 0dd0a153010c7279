!------------------------------------------------------------------------------
! fibgpu_iface.f90 -- Fortran binding (iso_c_binding) of include/fibgpu.h.
!
! The derived types mirror the C structs field for field (bind(C)); the
! interfaces declare every entry point of the C-ABI.  A Fortran caller passes
! its X(Natm,3), V(Natm,3) arrays directly: they already have the column-major
! layout the library expects.
!------------------------------------------------------------------------------
module fibgpu
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: FIB_OK = 0, FIB_ERR_ARG = 1, FIB_ERR_HIP = 2, &
       FIB_ERR_OVERFLOW = 3, FIB_ERR_TABLE = 4, FIB_ERR_NONFINITE = 5, &
       FIB_ERR_STATE = 6, FIB_ERR_NODEVICE = 7
  integer(c_int32_t), parameter :: FIB_PREC_FP64 = 0, FIB_PREC_MIXED = 1
  integer(c_int32_t), parameter :: FIB_EVAL_FORCE = 1, FIB_EVAL_ENERGY = 2, &
       FIB_EVAL_VIRIAL = 4
  integer(c_int32_t), parameter :: FIB_TI_EQUILIBRATE = 0, FIB_TI_IMPACT = 1, &
       FIB_TI_FREE = 2
  integer, parameter :: FIB_NTIMERS = 16

  type, bind(C) :: fib_params
     ! sw_Acoef is C sw_A (Fortran names are case-blind: sw_A would clash with sw_a)
     real(c_double) :: sw_Acoef, sw_B, sw_p, sw_q, sw_a, sw_lambda_si, sw_lambda_ge, sw_gamma
     real(c_double) :: eps_si, eps_ge, sigma_si, sigma_ge
     real(c_double) :: mass_si, mass_ge, mass_ga
     real(c_double) :: z_si, z_ge, z_ga
     real(c_double) :: mol_c(3), mol_b(3)
     real(c_double) :: bohr, kB, eps0, ec
     real(c_double) :: listfac
     integer(c_int32_t) :: ntable
     real(c_double) :: mixed_fp64_moliere_r
  end type fib_params

  type, bind(C) :: fib_instance_desc
     integer(c_int32_t) :: natm, nsg, nlj, reserved
     real(c_double) :: lb(3)
  end type fib_instance_desc

  type, bind(C) :: fib_run_params
     real(c_double) :: ts_i, ts_r, tau_i, tau_r, ttar, teps, ke_limit_1, dti, sidewidth
     real(c_double) :: outz(4), atomz(2), ionz
     integer(c_int32_t) :: ntlist_i, ntlist_r, atlist_i, atlist_r, dslist_i, dslist_r
     integer(c_int32_t) :: atm_out_i, atm_out_r, res_out, reserved
  end type fib_run_params

  type, bind(C) :: fib_step_ctl
     integer(c_int32_t) :: whichctrl, max_steps, stop_on_output, reserved
  end type fib_step_ctl

  type, bind(C) :: fib_step_report
     integer(c_int64_t) :: lt
     real(c_double) :: time, temp, ts
     integer(c_int32_t) :: steps_done, ended, output_due, regime, nlist_builds, atoms_due
  end type fib_step_report

  type, bind(C) :: fib_timing
     real(c_double) :: ms(FIB_NTIMERS)
     integer(c_int32_t) :: calls(FIB_NTIMERS)
  end type fib_timing

  ! MD-step mode (the benchmark configurations' integrator)
  type, bind(C) :: fib_md_params
     real(c_double) :: dt_max, dx_max, remove_height, shell_width, ttar, tau, displaced_r
     integer(c_int32_t) :: use_graph, reserved
  end type fib_md_params

  type, bind(C) :: fib_md_report
     real(c_double) :: time, dt, ion_z, ke
     integer(c_int64_t) :: steps
     integer(c_int32_t) :: removed, ions_removed, displaced, builds
     real(c_double) :: e_bath, e_removed, e_fired, dt_min
  end type fib_md_report

  interface
     subroutine fib_default_params(p) bind(C, name="fib_default_params")
       import :: fib_params
       type(fib_params), intent(out) :: p
     end subroutine fib_default_params

     integer(c_int32_t) function fib_create(p, ninst, inst, precision, h) &
          bind(C, name="fib_create")
       import :: c_int32_t, c_ptr, fib_params, fib_instance_desc
       type(fib_params), intent(in) :: p
       integer(c_int32_t), value :: ninst, precision
       type(fib_instance_desc), intent(in) :: inst(*)
       type(c_ptr), intent(out) :: h
     end function fib_create

     integer(c_int32_t) function fib_destroy(h) bind(C, name="fib_destroy")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function fib_destroy

     type(c_ptr) function fib_last_error(h) bind(C, name="fib_last_error")
       import :: c_ptr
       type(c_ptr), value :: h
     end function fib_last_error

     integer(c_int32_t) function fib_get_cells(h, inst, nc) bind(C, name="fib_get_cells")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       integer(c_int32_t), intent(out) :: nc(3)
     end function fib_get_cells

     integer(c_int32_t) function fib_set_state(h, inst, X, V, atype, xsides, Xi) &
          bind(C, name="fib_set_state")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, X, V, atype, xsides, Xi
       integer(c_int32_t), value :: inst
     end function fib_set_state

     integer(c_int32_t) function fib_get_state(h, inst, X, V, F) bind(C, name="fib_get_state")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, X, V, F
       integer(c_int32_t), value :: inst
     end function fib_get_state

     integer(c_int32_t) function fib_set_impact(h, inst, impact) bind(C, name="fib_set_impact")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst, impact
     end function fib_set_impact

     integer(c_int32_t) function fib_bin_cells(h, inst, icell) bind(C, name="fib_bin_cells")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, icell
       integer(c_int32_t), value :: inst
     end function fib_bin_cells

     integer(c_int32_t) function fib_build_nlist(h) bind(C, name="fib_build_nlist")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function fib_build_nlist

     integer(c_int32_t) function fib_get_nlist(h, inst, nss, nlist2, len2, nsslj, nlistlj, &
          lenlj) bind(C, name="fib_get_nlist")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, nss, nlist2, len2, nsslj, nlistlj, lenlj
       integer(c_int32_t), value :: inst
     end function fib_get_nlist

     integer(c_int32_t) function fib_compute(h, flags) bind(C, name="fib_compute")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: flags
     end function fib_compute

     integer(c_int32_t) function fib_get_forces(h, inst, F, u, utot, virial) &
          bind(C, name="fib_get_forces")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, F, u, utot, virial
       integer(c_int32_t), value :: inst
     end function fib_get_forces

     integer(c_int32_t) function fib_force(h, inst, impact, F, u, utot, virial) &
          bind(C, name="fib_force")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, F, u, utot, virial
       integer(c_int32_t), value :: inst, impact
     end function fib_force

     integer(c_int32_t) function fib_set_run(h, inst, rp) bind(C, name="fib_set_run")
       import :: c_int32_t, c_ptr, fib_run_params
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       type(fib_run_params), intent(in) :: rp
     end function fib_set_run

     integer(c_int32_t) function fib_tint_init(h, lt0, time0) bind(C, name="fib_tint_init")
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int64_t), value :: lt0
       real(c_double), value :: time0
     end function fib_tint_init

     integer(c_int32_t) function fib_ti_begin(h, whichctrl, impact) bind(C, name="fib_ti_begin")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: whichctrl, impact
     end function fib_ti_begin

     integer(c_int32_t) function fib_knock(h, knc, xion, vion) bind(C, name="fib_knock")
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: knc
       real(c_double), intent(in) :: xion(3), vion(3)
     end function fib_knock

     integer(c_int32_t) function fib_step(h, ctl, rep) bind(C, name="fib_step")
       import :: c_int32_t, c_ptr, fib_step_ctl, fib_step_report
       type(c_ptr), value :: h
       type(fib_step_ctl), intent(in) :: ctl
       type(fib_step_report), intent(out) :: rep
     end function fib_step

     integer(c_int32_t) function fib_temperature(h, inst, temp) bind(C, name="fib_temperature")
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       real(c_double), intent(out) :: temp
     end function fib_temperature

     integer(c_int32_t) function fib_timing_enable(h, mask) bind(C, name="fib_timing_enable")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: mask
     end function fib_timing_enable

     integer(c_int32_t) function fib_get_timing(h, t) bind(C, name="fib_get_timing")
       import :: c_int32_t, c_ptr, fib_timing
       type(c_ptr), value :: h
       type(fib_timing), intent(out) :: t
     end function fib_get_timing

     type(c_ptr) function fib_timer_name(k) bind(C, name="fib_timer_name")
       import :: c_ptr, c_int32_t
       integer(c_int32_t), value :: k
     end function fib_timer_name

     integer(c_int32_t) function fib_synchronize(h) bind(C, name="fib_synchronize")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function fib_synchronize

     integer(c_int32_t) function fib_md_begin(h, p, surface_z, bottom_z) &
          bind(C, name="fib_md_begin")
       import :: c_int32_t, c_double, c_ptr, fib_md_params
       type(c_ptr), value :: h
       type(fib_md_params), intent(in) :: p
       real(c_double), intent(in) :: surface_z(*), bottom_z(*)
     end function fib_md_begin

     integer(c_int32_t) function fib_md_fire(h, inst, ion, x, v) bind(C, name="fib_md_fire")
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst, ion
       real(c_double), intent(in) :: x(3), v(3)
     end function fib_md_fire

     integer(c_int32_t) function fib_md_fire_many(h, count, inst, ion, x, v) &
          bind(C, name="fib_md_fire_many")
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: count
       integer(c_int32_t), intent(in) :: inst(*), ion(*)
       real(c_double), intent(in) :: x(3,*), v(3,*)
     end function fib_md_fire_many

     integer(c_int32_t) function fib_md_run(h, nsteps) bind(C, name="fib_md_run")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
     end function fib_md_run

     ! fib_md_run in two halves: the host works while the steps run
     integer(c_int32_t) function fib_md_run_begin(h, nsteps) bind(C, name="fib_md_run_begin")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nsteps
     end function fib_md_run_begin

     integer(c_int32_t) function fib_md_run_end(h) bind(C, name="fib_md_run_end")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function fib_md_run_end

     ! asynchronous snapshot of X, V (original order) into a slot (0 or 1)
     integer(c_int32_t) function fib_snapshot_begin(h, inst, slot) bind(C, name="fib_snapshot_begin")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst, slot
     end function fib_snapshot_begin

     integer(c_int32_t) function fib_snapshot_wait(h, slot, X, V) bind(C, name="fib_snapshot_wait")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h, X, V
       integer(c_int32_t), value :: slot
     end function fib_snapshot_wait

     integer(c_int32_t) function fib_md_get_report(h, inst, r) bind(C, name="fib_md_get_report")
       import :: c_int32_t, c_ptr, fib_md_report
       type(c_ptr), value :: h
       integer(c_int32_t), value :: inst
       type(fib_md_report), intent(out) :: r
     end function fib_md_get_report

     integer(c_int32_t) function fib_md_energy(h, epot) bind(C, name="fib_md_energy")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: epot(*)
     end function fib_md_energy
  end interface

contains

  ! Message of the last failed call on h, as a Fortran string.
  function fib_error_string(h) result(s)
    type(c_ptr), value :: h
    character(len=:), allocatable :: s
    type(c_ptr) :: p
    character(kind=c_char), pointer :: c(:)
    integer :: n
    interface
       integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
         import :: c_ptr, c_size_t
         type(c_ptr), value :: str
       end function c_strlen
    end interface
    p = fib_last_error(h)
    if (.not. c_associated(p)) then
       s = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, c, [n])
    allocate(character(len=n) :: s)
    s = transfer(c(1:n), s)
  end function fib_error_string

end module fibgpu
