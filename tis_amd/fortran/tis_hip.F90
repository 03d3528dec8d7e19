! tis_hip.F90 -- Fortran (iso_c_binding) interface to libtis_hip.so, the
! MI355X implementation of the DT15 force-and-integrate path.
!
! The derived types mirror include/tis_system.h and include/tis_bdhi.h;
! the interface block binds every entry point of include/tis_hip.h and
! include/tis_bdhi.h.  The wrappers below keep the calling convention of
! the reference's call sites (1-based replica index, (3,N) FP64 arrays,
! stop on error), so a host engine can route
!   neighbor(irep)                                   (neighbor.F90:13)
!   force_sumup(force_mp, irep)                      (force_sumup.F90:15-32)
!   energy_sumup(irep, velo_mp, energy, energy_unit) (energy_sumup.F90:17-36)
! to the device with one call each.
module tis_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: TIS_OK = 0, TIS_MODE_GRAPH = 0, TIS_MODE_EAGER = 1, &
       TIS_MODE_PLAIN = 2, TIS_E_MAX = 25

  type, bind(C), public :: tis_params_c
     real(c_double) :: exv_adjust, exv_dist, exv_dist_PS, exv_inf
     integer(c_int32_t) :: n_sep_P, n_sep_S, n_sep_B
     real(c_double) :: hb_cutoff_dist, neigh_margin
     integer(c_int32_t) :: ele_form, dh_cutoff_type
     real(c_double) :: cutoff_ele, ele_coef, ele_cdist, ewld_alpha
     integer(c_int32_t) :: ewld_hmax
     real(c_double) :: beta_hb
     integer(c_int32_t) :: periodic
     real(c_double) :: box(3)
  end type tis_params_c

  type, bind(C), public :: tis_system_c
     integer(c_int32_t) :: nmp, nunit
     type(c_ptr) :: unit, mptype, iclass, exv_rad, exv_eps, charge, rlist2_unit
     integer(c_int32_t) :: nbd
     type(c_ptr) :: bd_mp, bd_nat, bd_coef
     integer(c_int32_t) :: nba
     type(c_ptr) :: ba_mp, ba_nat, ba_coef
     integer(c_int32_t) :: nst
     type(c_ptr) :: st_mp, st_nat, st_coef
     integer(c_int32_t) :: ntst
     type(c_ptr) :: tst_mp, tst_nat, tst_coef, tst2st, tst_excl
     integer(c_int32_t) :: nhb
     type(c_ptr) :: hb_mp, hb_nat, hb_coef, hb_site, hb_tertiary
     integer(c_int32_t) :: nhbsite
     type(c_ptr) :: hbsite_valence
     type(c_ptr) :: mass, fric, radius
     real(c_double) :: tempk
     type(tis_params_c) :: p
  end type tis_system_c

  ! include/tis_bdhi.h
  type, bind(C), public :: tis_bd_system_c
     integer(c_int32_t) :: nmp
     type(c_ptr) :: radius
     integer(c_int32_t) :: nfene
     type(c_ptr) :: fene_mp, fene_nat, fene_coef, fene_dist2
     integer(c_int32_t) :: nwca
     type(c_ptr) :: wca_mp, wca_nat2, wca_coef
     real(c_double) :: cutoff_wca, tempk, viscosity
     integer(c_int32_t) :: hydro_tensor
  end type tis_bd_system_c

  public :: tis_ctx_create, tis_ctx_destroy, tis_last_error, tis_system_upload, tis_system_validate, &
       tis_set_system_T, tis_list_capacity, tis_term_status, tis_transfer_bytes, &
       tis_system_append, tis_system_retire, tis_arena_stats, tis_tune, tis_launch_shape, &
       tis_state_set, tis_state_get, tis_neighbor, tis_pairs_get, tis_force_sumup, tis_force_get, &
       tis_energy_sumup, tis_energy_sumup_set, tis_energy_unit, tis_langevin_init, tis_langevin_steps, tis_system_origin, tis_device_bytes, &
       tis_status_counters, tis_synchronize, tis_last_timing, tis_last_force_ms, tis_label_define, &
       tis_label_apply, tis_bdhi_setup, tis_bdhi_state_set, tis_bdhi_state_get, tis_bdhi_forces, &
       tis_bdhi_tensors, tis_bdhi_tensor_get, tis_bdhi_factor_error, tis_bdhi_steps, tis_bdhi_timing, &
       tis_bdhi_factor_noalloc
  public :: hip_check, hip_neighbor, hip_force_sumup, hip_energy_sumup

  interface
     integer(c_int32_t) function tis_ctx_create(device, ctx) bind(C, name="tis_ctx_create")
       import :: c_int32_t, c_ptr
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int32_t) function tis_ctx_destroy(ctx) bind(C, name="tis_ctx_destroy")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
     end function
     type(c_ptr) function tis_last_error(ctx) bind(C, name="tis_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int32_t) function tis_system_upload(ctx, nsys, systems, keys, seed) bind(C, name="tis_system_upload")
       import :: c_int32_t, c_ptr, c_int64_t, tis_system_c
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsys
       type(tis_system_c), intent(in) :: systems(*)
       type(c_ptr), value :: keys
       integer(c_int64_t), value :: seed
     end function
     integer(c_int32_t) function tis_system_validate(nsys, systems, msg, msg_cap) &
          bind(C, name="tis_system_validate")
       import :: c_int32_t, c_int64_t, c_char, tis_system_c
       integer(c_int32_t), value :: nsys
       type(tis_system_c), intent(in) :: systems(*)
       character(kind=c_char), intent(out) :: msg(*)
       integer(c_int64_t), value :: msg_cap
     end function
     integer(c_int32_t) function tis_transfer_bytes(ctx, bytes) bind(C, name="tis_transfer_bytes")
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: bytes(4)
     end function
     integer(c_int32_t) function tis_system_append(ctx, nadd, systems, keys, isys_out) &
          bind(C, name="tis_system_append")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx, systems, keys, isys_out
       integer(c_int32_t), value :: nadd
     end function
     integer(c_int32_t) function tis_system_retire(ctx, n, isys) bind(C, name="tis_system_retire")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       integer(c_int32_t), intent(in) :: isys(*)
     end function
     integer(c_int32_t) function tis_tune(ctx, reps, shape_out, ms_out) bind(C, name="tis_tune")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: reps
       integer(c_int32_t), intent(out) :: shape_out(3)
       real(c_double), intent(out) :: ms_out(8)
     end function
     integer(c_int32_t) function tis_launch_shape(ctx, set, get) bind(C, name="tis_launch_shape")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx, set, get
     end function
     integer(c_int32_t) function tis_arena_stats(ctx, stats) bind(C, name="tis_arena_stats")
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: stats(6)
     end function
     integer(c_int32_t) function tis_term_status(ctx, isys, hb_status, st_status) bind(C, name="tis_term_status")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys
       integer(c_int32_t), intent(out) :: hb_status(*), st_status(*)
     end function
     integer(c_int32_t) function tis_list_capacity(ctx, cap, cap_out) bind(C, name="tis_list_capacity")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: cap
       integer(c_int32_t), intent(out) :: cap_out
     end function
     integer(c_int32_t) function tis_set_system_T(ctx, isys, sys) bind(C, name="tis_set_system_T")
       import :: c_int32_t, c_ptr, tis_system_c
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys
       type(tis_system_c), intent(in) :: sys
     end function
     integer(c_int32_t) function tis_state_set(ctx, isys, which, buf) bind(C, name="tis_state_set")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys, which
       real(c_double), intent(in) :: buf(*)
     end function
     integer(c_int32_t) function tis_state_get(ctx, isys, which, buf) bind(C, name="tis_state_get")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys, which
       real(c_double), intent(out) :: buf(*)
     end function
     integer(c_int32_t) function tis_neighbor(ctx, isys) bind(C, name="tis_neighbor")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys
     end function
     integer(c_int32_t) function tis_pairs_get(ctx, isys, which, pairs, coef, cap, n) bind(C, name="tis_pairs_get")
       import :: c_int32_t, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, pairs, coef
       integer(c_int32_t), value :: isys, which
       integer(c_int64_t), value :: cap
       integer(c_int64_t), intent(out) :: n
     end function
     integer(c_int32_t) function tis_force_sumup(ctx, isys, force_out) bind(C, name="tis_force_sumup")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys
       real(c_double), intent(out) :: force_out(*)
     end function
     integer(c_int32_t) function tis_force_get(ctx, isys, force_out) bind(C, name="tis_force_get")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys
       real(c_double), intent(out) :: force_out(*)
     end function
     integer(c_int32_t) function tis_energy_sumup(ctx, isys, energy_out) bind(C, name="tis_energy_sumup")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys
       real(c_double), intent(out) :: energy_out(*)
     end function
     integer(c_int32_t) function tis_energy_sumup_set(ctx, n, isys, energy_out) bind(C, name="tis_energy_sumup_set")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       integer(c_int32_t), intent(in) :: isys(*)
       real(c_double), intent(out) :: energy_out(*)
     end function
     integer(c_int32_t) function tis_energy_unit(ctx, isys, style, energy_unit) bind(C, name="tis_energy_unit")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys, style
       real(c_double), intent(out) :: energy_unit(*)
     end function
     integer(c_int32_t) function tis_langevin_init(ctx, tstep, rebuild_every, step0, init_accel) &
          bind(C, name="tis_langevin_init")
       import :: c_int32_t, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), value :: tstep
       integer(c_int32_t), value :: rebuild_every, init_accel
       integer(c_int64_t), value :: step0
     end function
     integer(c_int32_t) function tis_langevin_steps(ctx, nsteps, mode) bind(C, name="tis_langevin_steps")
       import :: c_int32_t, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: nsteps
       integer(c_int32_t), value :: mode
     end function
     integer(c_int32_t) function tis_system_origin(ctx, isys, step_origin, init_accel) bind(C, name="tis_system_origin")
       import :: c_int32_t, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys, init_accel
       integer(c_int64_t), value :: step_origin
     end function
     integer(c_int32_t) function tis_device_bytes(ctx, bytes) bind(C, name="tis_device_bytes")
       import :: c_int32_t, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: bytes
     end function
     integer(c_int32_t) function tis_status_counters(ctx, counters) bind(C, name="tis_status_counters")
       import :: c_int32_t, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: counters(*)
     end function
     integer(c_int32_t) function tis_synchronize(ctx) bind(C, name="tis_synchronize")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int32_t) function tis_last_timing(ctx, total_ms, step_ms, pre_ms, nlaunch) bind(C, name="tis_last_timing")
       import :: c_int32_t, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: total_ms, step_ms, pre_ms
       integer(c_int64_t), intent(out) :: nlaunch
     end function
     integer(c_int32_t) function tis_last_force_ms(ctx, force_ms) bind(C, name="tis_last_force_ms")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: force_ms
     end function
     integer(c_int32_t) function tis_label_define(ctx, label, isys, sys) bind(C, name="tis_label_define")
       import :: c_int32_t, c_ptr, tis_system_c
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: label, isys
       type(tis_system_c), intent(in) :: sys
     end function
     integer(c_int32_t) function tis_label_apply(ctx, n, isys, label, vscale) bind(C, name="tis_label_apply")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx, vscale
       integer(c_int32_t), value :: n
       integer(c_int32_t), intent(in) :: isys(*), label(*)
     end function
     integer(c_int32_t) function tis_bdhi_setup(ctx, sys, key, seed) bind(C, name="tis_bdhi_setup")
       import :: c_int32_t, c_ptr, c_int64_t, tis_bd_system_c
       type(c_ptr), value :: ctx
       type(tis_bd_system_c), intent(in) :: sys
       integer(c_int32_t), value :: key
       integer(c_int64_t), value :: seed
     end function
     integer(c_int32_t) function tis_bdhi_state_set(ctx, which, buf) bind(C, name="tis_bdhi_state_set")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which
       real(c_double), intent(in) :: buf(*)
     end function
     integer(c_int32_t) function tis_bdhi_state_get(ctx, which, buf) bind(C, name="tis_bdhi_state_get")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which
       real(c_double), intent(out) :: buf(*)
     end function
     integer(c_int32_t) function tis_bdhi_forces(ctx, force_out) bind(C, name="tis_bdhi_forces")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: force_out(*)
     end function
     integer(c_int32_t) function tis_bdhi_tensors(ctx) bind(C, name="tis_bdhi_tensors")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int32_t) function tis_bdhi_tensor_get(ctx, which, buf) bind(C, name="tis_bdhi_tensor_get")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: which
       real(c_double), intent(out) :: buf(*)
     end function
     integer(c_int32_t) function tis_bdhi_factor_error(ctx, rel_err) bind(C, name="tis_bdhi_factor_error")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: rel_err
     end function
     integer(c_int32_t) function tis_bdhi_steps(ctx, tstep, nsteps, refactor_every, step0) bind(C, name="tis_bdhi_steps")
       import :: c_int32_t, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       real(c_double), value :: tstep
       integer(c_int64_t), value :: nsteps, step0
       integer(c_int32_t), value :: refactor_every
     end function
     integer(c_int32_t) function tis_bdhi_timing(ctx, total_ms, fill_ms, factor_ms) bind(C, name="tis_bdhi_timing")
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: total_ms, fill_ms, factor_ms
     end function
     integer(c_int32_t) function tis_bdhi_factor_noalloc(device, D, L, n3, stream, info) &
          bind(C, name="tis_bdhi_factor_noalloc")
       import :: c_int32_t, c_ptr
       integer(c_int32_t), value :: device, n3
       type(c_ptr), value :: D, L, stream, info
     end function
     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function
  end interface

contains

  ! status < 0 -> print the library's message and stop, as util_error
  ! (ERROR%STOP_ALL) does in the reference (util_error.F90:77,88)
  subroutine hip_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer(c_size_t) :: n
    if (rc >= 0) return
    p = tis_last_error(ctx)
    n = c_strlen(p)
    call c_f_pointer(p, msg, [n])
    write(0, '(a, a, i0, a, *(a))') what, ' failed (', rc, '): ', msg
    error stop 1
  end subroutine hip_check

  ! neighbor(irep), neighbor.F90:13
  subroutine hip_neighbor(ctx, irep)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: irep
    call hip_check(ctx, tis_neighbor(ctx, int(irep - 1, c_int32_t)), 'tis_neighbor')
  end subroutine hip_neighbor

  ! force_sumup(force_mp, irep), force_sumup.F90:15-32
  subroutine hip_force_sumup(ctx, force_mp, irep)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(out) :: force_mp(:,:)
    integer, intent(in) :: irep
    call hip_check(ctx, tis_force_sumup(ctx, int(irep - 1, c_int32_t), force_mp), 'tis_force_sumup')
  end subroutine hip_force_sumup

  ! energy_sumup(irep, velo_mp, energy, energy_unit), energy_sumup.F90:17-36;
  ! energy_unit(nunit, nunit, E_TYPE%MAX) with the post-processing of
  ! i_output_energy_style (style, default 0)
  subroutine hip_energy_sumup(ctx, irep, velo_mp, energy, energy_unit, style)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: irep
    real(c_double), intent(in) :: velo_mp(:,:)
    real(c_double), intent(out) :: energy(:)
    real(c_double), intent(out) :: energy_unit(:,:,:)
    integer, intent(in), optional :: style
    integer(c_int32_t) :: st
    st = 0_c_int32_t
    if (present(style)) st = int(style, c_int32_t)
    call hip_check(ctx, tis_state_set(ctx, int(irep - 1, c_int32_t), 1_c_int32_t, velo_mp), 'tis_state_set')
    call hip_check(ctx, tis_energy_sumup(ctx, int(irep - 1, c_int32_t), energy), 'tis_energy_sumup')
    call hip_check(ctx, tis_energy_unit(ctx, int(irep - 1, c_int32_t), st, energy_unit), 'tis_energy_unit')
  end subroutine hip_energy_sumup

end module tis_hip
