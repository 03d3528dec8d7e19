/*
 * tis_hip.h -- C-ABI of libtis_hip.so, the MI355X (gfx950) implementation of
 * the DT15 TIS per-step force-and-integrate path.
 *
 * One context owns a batch of systems (one system = one replica of one
 * molecule set, or one independent chain) resident in HBM.  Every entry point
 * returns int32 status: 0 ok, < 0 error (message via tis_last_error).  Plain
 * pointers and sizes only; no exceptions cross the ABI.  Calls on one context
 * are stream-ordered and not thread-safe; separate contexts are independent.
 *
 * Reference interfaces replaced (paths in the reference's src/):
 *   tis_neighbor       <- neighbor(irep)                 neighbor.F90:13
 *   tis_force_sumup    <- force_sumup(force_mp, irep)    force_sumup.F90:15-32
 *   tis_energy_sumup   <- energy_sumup(irep, velo_mp, energy, energy_unit)
 *                                                        energy_sumup.F90:17-36
 *   tis_langevin_*     <- the Langevin branch of time_integral
 *                        (time_integral_pre.F90:160-180, time_integral.F90:57-189)
 *   tis_set_system_T   <- simu_para2 / simu_ele_set / simu_set_dtrna refresh
 *                        (simu_para2.F90:8-108) for one replica
 */
#ifndef TIS_HIP_H
#define TIS_HIP_H

#include <stdint.h>
#include "tis_system.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tis_ctx tis_ctx;

/* status codes */
enum { TIS_OK = 0, TIS_ERR_ARG = -1, TIS_ERR_HIP = -2, TIS_ERR_STATE = -3,
       TIS_ERR_OVERFLOW = -4, TIS_ERR_NOMEM = -5 };

/* which list for tis_pairs_get */
enum { TIS_LIST_EXV = 0, TIS_LIST_ELE = 1, TIS_LIST_HB = 2 };

/* counters of tis_status_counters (TIS_CNT_REBUILDS: whole-batch rebuilds,
 * plus one per system rebuild fired by the displacement trigger;
 * TIS_CNT_NB_OVERFLOW and the slot after it are the run's stop flag and
 * completed-step count, zero between calls; TIS_CNT_RESUMES: runs resumed
 * after a list rebuild overflowed the per-bead capacity) */
enum { TIS_CNT_CLAMP_EXV = 0, TIS_CNT_CLAMP_ELE = 1, TIS_CNT_HB_DELETED = 2,
       TIS_CNT_REBUILDS = 3, TIS_CNT_STEPS = 4, TIS_CNT_NB_OVERFLOW = 5,
       TIS_CNT_RESUMES = 7, TIS_CNT_MAX = 8 };

/* Langevin launch modes for tis_langevin_steps */
enum { TIS_MODE_GRAPH = 0,   /* batched, hipGraph-captured, one step = 2 launches */
       TIS_MODE_EAGER = 1,   /* batched, launched eagerly                         */
       TIS_MODE_PLAIN = 2 }; /* one system per launch (plain GPU path)            */

/* Create a context on HIP device `device` (its own stream, default priority). */
int32_t tis_ctx_create(int32_t device, tis_ctx **out);
int32_t tis_ctx_destroy(tis_ctx *ctx);
/* Last error message of the context (static storage owned by ctx). */
const char *tis_last_error(const tis_ctx *ctx);

/* Upload a batch of nsys systems (replaces any previous batch).  keys[s] is the
 * stable identity of system s in the random streams (noise of a system depends
 * only on (seed, key, step, bead), not on its slot in the batch); NULL means
 * key = s.  Coordinates start at zero; set them with tis_state_set (which 0).
 * TIS_ERR_ARG if a system has more than 256 distinct (friction, mass) pairs
 * (its beads share Langevin constants by that pair). */
int32_t tis_system_upload(tis_ctx *ctx, int32_t nsys, const tis_system *systems,
                          const uint32_t *keys, uint64_t seed);

/* Check a batch description without a device call: every bead index of
 * every term in [0, nmp), units in [0, nunit), hbond sites in [-1, nhbsite),
 * tst2st in [-1, nst), tst_excl 0/1, no NULL array behind a non-zero
 * count.  TIS_ERR_ARG with the first offence in msg (msg_cap bytes, may be
 * NULL).  tis_system_upload runs the same check first and leaves the
 * previous batch in place when it fails.  No reference counterpart: the
 * reference trusts its .ninfo reader (read_nativeinfo.F90). */
int32_t tis_system_validate(int32_t nsys, const tis_system *systems, char *msg, int64_t msg_cap);

/* Add nadd systems to the batch without touching the device data of the
 * systems already in it (continuous batching; no reference counterpart: the
 * reference runs one job per process, mloop_simulator.F90).  The new
 * systems take indices nsys .. nsys + nadd - 1 (written to isys_out, may be
 * NULL); only their own topology and parameters are uploaded, after the
 * last system, when they fit the batch's capacities.  Otherwise the batch is
 * laid out again with room for twice what it then holds: retired systems'
 * storage is dropped and the live systems' positions, velocities and
 * accelerations are copied over on the device (their indices do not
 * change).  Lists, captured graphs and the Langevin set-up are dropped:
 * set the new systems' state, then tis_langevin_init (with per-system
 * origins from tis_system_origin) before the next steps. */
int32_t tis_system_append(tis_ctx *ctx, int32_t nadd, const tis_system *systems, const uint32_t *keys,
                          int32_t *isys_out);

/* Drop systems from every launch (their storage stays until the batch is
 * next laid out; their indices are not reused).  Calls on a retired system
 * fail with TIS_ERR_ARG.  As after tis_system_append, tis_langevin_init
 * comes before the next steps. */
int32_t tis_system_retire(tis_ctx *ctx, int32_t n, const int32_t *isys);

/* stats[6]: live systems, their beads, beads stored (live and retired since
 * the last layout), bead capacity, layouts so far, system indices issued. */
int32_t tis_arena_stats(tis_ctx *ctx, int64_t *stats);

/* Refresh the temperature-dependent coefficients of system isys (tempk,
 * st_coef U0, charge, ele_coef, ele_cdist, beta_hb) from `sys`, which must
 * have the same topology as the uploaded one.  Langevin constants follow. */
int32_t tis_set_system_T(tis_ctx *ctx, int32_t isys, const tis_system *sys);

/* State of system isys, (3,N) column-major FP64 like xyz_mp_rep(:,:,irep).
 * which: 0 coordinates, 1 velocities, 2 accelerations. */
int32_t tis_state_set(tis_ctx *ctx, int32_t isys, int32_t which, const double *buf);
int32_t tis_state_get(tis_ctx *ctx, int32_t isys, int32_t which, double *buf);

/* neighbor(irep): rebuild exv / ele / hb lists of system isys (-1: all) from
 * the current coordinates. */
int32_t tis_neighbor(tis_ctx *ctx, int32_t isys);

/* Copy a list of system isys out (parity checks).  exv/ele: pairs[3*k] =
 * (i, j, mirror) with i < j, 0-based, sorted by (i, j); coef[k] (ele only,
 * may be NULL) = q_i q_j ele_coef.  hb: pairs[k] = candidate index.  *n gets
 * the length; returns TIS_ERR_OVERFLOW if cap is too small. */
int32_t tis_pairs_get(tis_ctx *ctx, int32_t isys, int32_t which, int32_t *pairs,
                      double *coef, int64_t cap, int64_t *n);

/* force_sumup: forces on every bead of system isys at the current coordinates
 * and lists, out (3,N) FP64.  isys = -1 evaluates all systems in one launch
 * and writes nothing (use tis_force_get). */
int32_t tis_force_sumup(tis_ctx *ctx, int32_t isys, double *force_out);
int32_t tis_force_get(tis_ctx *ctx, int32_t isys, double *force_out);

/* energy_sumup: energy[TIS_E_MAX] (E_TYPE slots, TIS_E_TOTAL excludes VELO)
 * of system isys at the current coordinates, lists and last hbond state;
 * kinetic term from the stored velocities.  isys = -1: every system in one
 * launch, energy_out[nsys * TIS_E_MAX] (the replica loop of
 * energy_allrep.F90:123-125). */
int32_t tis_energy_sumup(tis_ctx *ctx, int32_t isys, double *energy_out);
/* energy_sumup of the n systems isys[0..n) in one launch (the replicas of
 * an exchange, energy_allrep.F90:123-216, when the batch holds other
 * systems too): energy_out[n * TIS_E_MAX], in the order of isys.  Repeats
 * are allowed; TIS_ERR_ARG for n < 1 or an index out of range. */
int32_t tis_energy_sumup_set(tis_ctx *ctx, int32_t n, const int32_t *isys, double *energy_out);
/* energy_unit of energy_sumup (energy_sumup.F90:17-36, 76-508) for one
 * system: energy_unit[(iunit, junit, type)] column-major (nunit, nunit,
 * TIS_E_MAX), each term added at the unit pair the reference's term routine
 * uses, then the post-processing of i_output_energy_style = style (0: half
 * of each upper off-diagonal entry to both diagonal entries; 1: symmetric
 * off-diagonal sums) and TOTAL over the types other than TOTAL and VELO.
 * The Ewald terms contribute nothing here, as in energy_ele_coulomb_ewld. */
int32_t tis_energy_unit(tis_ctx *ctx, int32_t isys, int32_t style, double *energy_unit);

/* Langevin dynamics for every system of the batch.
 * tis_langevin_init: time step tstep (tau); neighbour lists rebuilt every
 * rebuild_every steps (> 0) for the whole batch, or (rebuild_every = 0) per
 * system by the reference's displacement trigger: a system's lists are
 * rebuilt at the start of a step when the two largest displacements since
 * its last rebuild sum to more than neigh_margin (time_integral.F90:57-93);
 * first step index step0; if init_accel, the accelerations are set to
 * F/m + c1*xi(step0) (time_integral_pre.F90:176-179), if 0 the stored
 * accelerations are used, if -1 per system (tis_system_origin). */
int32_t tis_langevin_init(tis_ctx *ctx, double tstep, int32_t rebuild_every, uint64_t step0,
                          int32_t init_accel);
/* Runs nsteps steps.  When a list rebuild inside the run needs more
 * entries per bead than the lists hold, the run stops at the start of that
 * step (every later kernel of the run returns at once, the state is that
 * of the step's start), the capacity grows (entries already built are
 * kept) and the run resumes from that step: no step ever runs on truncated
 * lists, and the trajectory equals that of a run with the larger capacity
 * from the start.  TIS_CNT_RESUMES counts the resumes. */
int32_t tis_langevin_steps(tis_ctx *ctx, int64_t nsteps, int32_t mode);

/* Step origin of system isys: its own step count is (global step - origin),
 * which keys its noise and valence draws, so a system started at any point of
 * a run draws exactly what it would draw alone.  init_accel: when
 * tis_langevin_init is called with init_accel = -1, systems flagged here get
 * new accelerations and the others keep their stored ones. */
int32_t tis_system_origin(tis_ctx *ctx, int32_t isys, uint64_t step_origin, int32_t init_accel);

/* Temperature labels for replica exchange.  tis_label_define stores the
 * temperature-dependent coefficients of `sys` (charges, stack U0, ele_coef,
 * ele_cdist, beta_hb, tempk) as label `label`; sys must have the topology of
 * system isys.  tis_label_apply gives each system isys[k] the coefficients
 * of label[k] in one launch, refreshes its Langevin noise amplitude, and
 * multiplies its velocities by vscale[k] (NULL: no scaling).  Together they
 * replace the per-replica parameter refresh of energy_allrep.F90:150-216,
 * simu_replica_exchange.F90:272-300 and time_integral_post.F90:264-370. */
int32_t tis_label_define(tis_ctx *ctx, int32_t label, int32_t isys, const tis_system *sys);
int32_t tis_label_apply(tis_ctx *ctx, int32_t n, const int32_t *isys, const int32_t *label,
                        const double *vscale);

/* hbond and stack states of system isys, as a restart file records them
 * (the DTRNA15 block of write_rst.F90:96-106): hb_status[nhb] (may be
 * NULL) of the last force evaluation, 1 = formed and kept by the valence
 * pass; st_status[nst] (may be NULL) at the current coordinates, 0 = the
 * consecutive stack is switched off by an exclusive tertiary stack
 * (force_dtrna_stack_nlocal.F90:83-92). */
int32_t tis_term_status(tis_ctx *ctx, int32_t isys, int32_t *hb_status, int32_t *st_status);

/* Entries per bead of the excluded-volume lists (the lists are stored
 * nb[entry][bead]).  cap > 0 sets it (the entries built so far are kept;
 * TIS_ERR_ARG if the current lists hold more than cap entries for a bead),
 * cap <= 0 only reads it; *cap_out (may be NULL) gets the capacity in
 * force.  List builds grow it as they need (tis_neighbor,
 * tis_langevin_init, and the resume of tis_langevin_steps).  No reference
 * counterpart: the reference sizes its list by MXMPNEIGHBOR per bead
 * (const_maxsize.F90, allocate_neighbor.F90). */
int32_t tis_list_capacity(tis_ctx *ctx, int32_t cap, int32_t *cap_out);

/* Payload bytes moved since the context was created, bytes[4]:
 * [0] host to device, topology, parameters and launch tables;
 * [1] host to device, dynamic state (tis_state_set);
 * [2] device to host, dynamic state (tis_state_get);
 * [3] device to device, positions, velocities and accelerations carried by
 *     a new arena layout (tis_system_append).
 * Status words (counters, flags) are not counted.  Serving reports them per
 * admission and completion. */
int32_t tis_transfer_bytes(tis_ctx *ctx, int64_t *bytes);

/* Launch-shape tuner of the force pass (no reference counterpart): times
 * the force pass of the whole batch at the current coordinates under 8
 * launch shapes -- electrostatic chunk pairs per wave {1, 2} x bonded terms
 * per lane {1, 2} x off-diagonal chunk pairs first or in system order --
 * in reps rounds that time each candidate once (after warm-up passes), and
 * keeps the one with the smallest median for this batch.  No result
 * depends on the shape (forces and trajectories are bitwise the same) and
 * no state changes (the counters the passes bump are restored).
 * shape_out[3] = (pairs per wave, terms per lane, off-diagonal first),
 * ms_out[8] = median force-pass ms per candidate in the order above, pairs
 * per wave varying fastest (either may be NULL). */
int32_t tis_tune(tis_ctx *ctx, int32_t reps, int32_t *shape_out, double *ms_out);
/* Set (set != NULL) and/or read (get != NULL) the launch shape, 3 values as
 * in tis_tune. */
int32_t tis_launch_shape(tis_ctx *ctx, const int32_t *set, int32_t *get);

/* Device memory held by the context's batch, in bytes (admission control). */
int32_t tis_device_bytes(tis_ctx *ctx, int64_t *bytes);

/* Counters (TIS_CNT_*), accumulated since upload. */
int32_t tis_status_counters(tis_ctx *ctx, int64_t *counters);
int32_t tis_synchronize(tis_ctx *ctx);

/* Device-side timing of the last tis_langevin_steps call: total milliseconds
 * and, for eager runs, per-step averages of the gather/integrate kernel
 * (k_main, step_kernel_ms) and of everything before it (pre_kernel_ms), from
 * HIP events on the context's stream. */
int32_t tis_last_timing(tis_ctx *ctx, double *total_ms, double *step_kernel_ms,
                        double *pre_kernel_ms, int64_t *nlaunch);
/* Of the last eager tis_langevin_steps call: average duration of the force
 * pass kernel (k_force: hbonds and tertiary stacks, electrostatic pairs,
 * bonded terms) per step, ms, from events bound to its dispatch. */
int32_t tis_last_force_ms(tis_ctx *ctx, double *force_ms);

#ifdef __cplusplus
}
#endif
#endif /* TIS_HIP_H */
