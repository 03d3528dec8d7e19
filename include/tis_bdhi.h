/*
 * tis_bdhi.h -- C-ABI of the Brownian-dynamics-with-hydrodynamic-interaction
 * path of libtis_hip.so (Ermak-McCammon step with a Rotne-Prager-Yamakawa
 * diffusion tensor factored by blocked Cholesky on the GPU).
 *
 * Reference interfaces replaced (paths in the reference's src/):
 *   tis_bdhi_tensors  <- simu_hydro_tensors(irep, tempk)   simu_hydro_tensors.F90:17-324
 *                        (tensor fill :50-302, Cholesky :304-324)
 *   tis_bdhi_steps    <- the SIM%BROWNIAN_HI branch of time_integral
 *                        (time_integral.F90:429-457, constants time_integral_pre.F90:313-317)
 *   tis_bdhi_forces   <- force_sumup restricted to force_fene (force_fene.F90:6-59)
 *                        and force_wca (force_wca.F90:7-100)
 *
 * One context holds at most one BD-HI system beside its DT15 batch.  Arrays
 * are (3,n) column-major FP64 coordinates and (3n,3n) column-major tensors.
 */
#ifndef TIS_BDHI_H
#define TIS_BDHI_H

#include <stdint.h>
#include "tis_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* i_hydro_tensor (const_index.F90:39-45) */
enum { TIS_HI_RPY = 1, TIS_HI_RPY_OVER = 2, TIS_HI_ERMAK_OVER = 3, TIS_HI_ZUK_RPY = 4 };

typedef struct tis_bd_system {
  int32_t nmp;
  const double *radius;        /* [nmp] hydrodynamic radius [A] */
  int32_t nfene;               /* FENE bonds: ifene2mp, fene_nat, coef_fene, dist2_fene */
  const int32_t *fene_mp;      /* [2*nfene] 0-based */
  const double *fene_nat, *fene_coef, *fene_dist2;
  int32_t nwca;                /* WCA pairs: iwca2mp, wca_nat2, coef_wca(:,1:2) */
  const int32_t *wca_mp;       /* [2*nwca] 0-based */
  const double *wca_nat2;      /* [nwca] */
  const double *wca_coef;      /* [2*nwca]: repulsive, attractive */
  double cutoff_wca;           /* inpro%cutoff_wca */
  double tempk;
  double viscosity;            /* inpara%viscosity after setp_mass_fric (internal units) */
  int32_t hydro_tensor;        /* TIS_HI_* */
} tis_bd_system;

/* Upload the system (replaces a previous one); key/seed select its noise. */
int32_t tis_bdhi_setup(tis_ctx *ctx, const tis_bd_system *sys, uint32_t key, uint64_t seed);
/* which: 0 coordinates (set/get), 1 forces at the current coordinates (get). */
int32_t tis_bdhi_state_set(tis_ctx *ctx, int32_t which, const double *buf);
int32_t tis_bdhi_state_get(tis_ctx *ctx, int32_t which, double *buf);
/* Forces at the current coordinates (force_fene + force_wca). */
int32_t tis_bdhi_forces(tis_ctx *ctx, double *force_out);
/* Fill D from the current coordinates and factor D = L L^T (FP64 diagonal
 * blocks and triangular solves, FP32 matrix-core trailing updates).
 * Returns TIS_ERR_STATE if a pivot is not positive. */
int32_t tis_bdhi_tensors(tis_ctx *ctx);
/* which: 0 D, 1 L (lower, upper part zero); buf (3n,3n) column-major. */
int32_t tis_bdhi_tensor_get(tis_ctx *ctx, int32_t which, double *buf);
/* ||L L^T - D||_F / ||D||_F of the current factor, on the device. */
int32_t tis_bdhi_factor_error(tis_ctx *ctx, double *rel_err);
/* nsteps Ermak-McCammon steps of size tstep; the tensors are refreshed
 * before the first step and then every refactor_every steps (1 = every
 * step, as the reference does); step s = step0 + 1 .. step0 + nsteps draws
 * the noise of (seed, key, s, bead). */
int32_t tis_bdhi_steps(tis_ctx *ctx, double tstep, int64_t nsteps, int32_t refactor_every, uint64_t step0);
/* Device time of the last tis_bdhi_tensors / tis_bdhi_steps call (ms) and
 * the share of the factorisation in it. */
int32_t tis_bdhi_timing(tis_ctx *ctx, double *total_ms, double *fill_ms, double *factor_ms);

/* The factorisation of tis_bdhi_tensors without any allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  D and L are the caller's device buffers, (n3,n3) column-major
 * FP64; L gets the lower Cholesky factor of D with a zero upper part, from
 * the same kernels (and the same bits) as tis_bdhi_tensors.  The FP32 copy
 * of one outer block's panels (256 columns x n3 rows), the pivot flag and
 * two 64x64 diagonal factors awaiting their store live in a static device
 * array of the library (9.5 MB per device, held until the process ends):
 * n3 <= 9216 (3,072 beads), TIS_ERR_ARG beyond.  Work is queued on `stream` (a hipStream_t;
 * NULL: the default stream).  With info != NULL the call waits for the
 * stream and stores the first failing pivot + 1 (0: none; TIS_ERR_STATE is
 * then returned); with info == NULL it returns at once.  Every call on a
 * device shares that scratch array: calls that overlap on one device race. */
int32_t tis_bdhi_factor_noalloc(int32_t device, const double *D, double *L, int32_t n3, void *stream,
                                int32_t *info);

#ifdef __cplusplus
}
#endif
#endif /* TIS_BDHI_H */
