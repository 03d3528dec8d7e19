/*
 * tis_oracle.h -- FP64 CPU restatement of the reference's DT15 force-and-
 * integrate path.  TEST INFRASTRUCTURE ONLY: the oracle that the HIP
 * library (tis_amd/csrc) is compared against, and the CPU baseline timed
 * beside it.  No product code path may call it.
 *
 * Every function cites the reference routine it restates (file:line under
 * the reference's src/).  Indices are 0-based.  Coordinates are (3,N)
 * column-major FP64 like xyz_mp_rep(:,:,irep).
 */
#ifndef TIS_ORACLE_H
#define TIS_ORACLE_H

#include <stdint.h>
#include "../include/tis_system.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based random numbers (Philox4x32-10, Salmon et al. SC'11) ----
 * The build replaces the reference's Mersenne Twister streams with a counter-
 * based generator so that a bead's noise depends only on (seed, system, step,
 * bead) and never on launch order.  The oracle and the HIP kernels share the
 * definition below bit for bit. */
void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void oracle_gauss3(uint64_t seed, uint32_t sys, uint64_t step, uint32_t bead, double g[3]);
double oracle_uniform(uint64_t seed, uint32_t sys, uint64_t step, uint32_t stream, uint32_t idx);

/* ---- periodic helpers: util_periodic.F90:4-39, util_pbneighbor.F90:4-38,
 *      setp_periodic.F90:80-115 ---- */
void oracle_wrap(const tis_system *s, double *pxyz);
void oracle_mirror_shift(const tis_system *s, int32_t imirror, double d[3]);

/* ---- neighbour lists ----
 * exv: neighbor_list.F90:56-87 + neighbor_assign.F90:481-559 (DT15 rules)
 * pairs[3*k] = (i, j, imirror), i < j, sorted by (i, j).  Returns count or
 * -1 if cap is exceeded. */
int64_t oracle_list_exv(const tis_system *s, const double *pxyz, int32_t *pairs, int64_t cap);
/* ele: neighbor_list_ele.F90:123-210.  coef[k] = q_i q_j * ele_coef. */
int64_t oracle_list_ele(const tis_system *s, const double *pxyz, int32_t *pairs,
                        double *coef, int64_t cap);
/* hb: neighbor_list_hb.F90:59-78; radius mloop_dtrna.F90:149-151.
 * hb_idx: sorted indices of candidates in the list. */
int32_t oracle_list_hb(const tis_system *s, const double *xyz, const double *pxyz, int32_t *hb_idx);

/* ---- force terms (accumulate into f (3,N)) ---- */
void oracle_force_bond(const tis_system *s, const double *xyz, double *f);        /* force_bond.F90:5-63 */
void oracle_force_bangle(const tis_system *s, const double *xyz, double *f);      /* force_bangle.F90:4-87 */
int32_t oracle_force_exv(const tis_system *s, const double *pxyz, const int32_t *pairs,
                         int64_t n, double *f);                                   /* force_exv_dt15.F90:4-118; returns clamp count */
int32_t oracle_force_dh(const tis_system *s, const double *pxyz, const int32_t *pairs,
                        const double *coef, int64_t n, double *f);                /* force_ele_DH.F90:4-176 */
void oracle_force_tstack(const tis_system *s, const double *xyz, const double *pxyz,
                         uint8_t *st_status, double *f);                          /* force_dtrna_stack_nlocal.F90:6-266 */
void oracle_force_stack(const tis_system *s, const double *xyz, const uint8_t *st_status,
                        double *f);                                               /* force_dtrna_stack.F90:7-168 */
/* force_dtrna_hbond15.F90:91-353: geometry for every listed candidate.
 * for_hb [18*nhb] (3,6,nhb), ene_hb [nhb], status [nhb] (1 if |r-r0| <= cut). */
void oracle_hbond_eval(const tis_system *s, const double *xyz, const double *pxyz,
                       const int32_t *hb_idx, int32_t nlist,
                       double *for_hb, double *ene_hb, uint8_t *status);
/* force_dtrna_hbond15.F90:373-510 valence-limit deletion with the counter-
 * based generator; returns the number of deletions. */
int32_t oracle_hbond_valence(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                             double *ene_hb, uint8_t *status,
                             uint64_t seed, uint32_t sys, uint64_t step);
/* The same pass drawing the reference's own numbers: raw[] holds the 32-bit
 * outputs of the replica's Mersenne-Twister stream (mts(irep,0)) in order;
 * each draw takes two of them and forms genrand_double4 (site pick) or
 * genrand_double1 (shuffle, Metropolis) as mt_stream.F90:708-785 do.
 * *used gets the number of raw words consumed; returns -1 if nraw was too
 * short. */
int32_t oracle_hbond_valence_mt(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                                double *ene_hb, uint8_t *status,
                                const uint32_t *raw, int64_t nraw, int64_t *used);
void oracle_hbond_scatter(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                          const double *for_hb, const uint8_t *status, double *f);

/* ---- energy terms: energy_*.F90 ---- */
double oracle_energy_bond(const tis_system *s, const double *xyz);
double oracle_energy_bangle(const tis_system *s, const double *xyz);
double oracle_energy_exv(const tis_system *s, const double *pxyz, const int32_t *pairs, int64_t n);
double oracle_energy_dh(const tis_system *s, const double *pxyz, const int32_t *pairs,
                        const double *coef, int64_t n);
double oracle_energy_stack(const tis_system *s, const double *xyz, const uint8_t *st_status);
double oracle_energy_tstack(const tis_system *s, const double *xyz, const double *pxyz);
double oracle_energy_kinetic(const tis_system *s, const double *velo);

/* ---- plain Coulomb (force_ele_coulomb.F90, energy_ele_coulomb.F90) and
 *      Ewald (mloop_ewld.F90:52-154, force_ele_coulomb_ewld.F90:56-137,
 *      energy_ele_coulomb_ewld.F90:56-212 without the dipole term) ---- */
void oracle_force_coulomb(const tis_system *s, const double *pxyz, const int32_t *pairs,
                          const double *coef, int64_t n, double *f);
double oracle_energy_coulomb(const tis_system *s, const double *pxyz, const int32_t *pairs,
                             const double *coef, int64_t n);
int32_t oracle_ewald_kvec(const tis_system *s, int32_t *h, double *coef);
void oracle_force_ewald(const tis_system *s, const double *pxyz, const int32_t *pairs,
                        const double *coef, int64_t n, double *f);
/* parts (may be NULL): Fourier, real, self */
double oracle_energy_ewald(const tis_system *s, const double *pxyz, const int32_t *pairs,
                           const double *coef, int64_t n, double parts[3]);

/* ---- full force_sumup / energy_sumup for one system (DT15, DH or none) ----
 * Lists are rebuilt from xyz first (neighbor.F90), as the parity harness
 * does.  energy[TIS_E_MAX]. */
typedef struct oracle_ws oracle_ws;
oracle_ws *oracle_ws_new(const tis_system *s);
void oracle_ws_free(oracle_ws *w);
void oracle_neighbor(oracle_ws *w, const double *xyz);
int32_t oracle_force_sumup(oracle_ws *w, const double *xyz, double *f,
                           uint64_t seed, uint32_t sys, uint64_t step);
/* force_sumup with the valence pass drawing the Mersenne-Twister words raw[]
 * (see oracle_hbond_valence_mt); returns -1 if nraw was too short. */
int32_t oracle_force_sumup_mt(oracle_ws *w, const double *xyz, double *f,
                              const uint32_t *raw, int64_t nraw, int64_t *used);
void oracle_energy_sumup(oracle_ws *w, const double *xyz, const double *velo, double *energy);
void oracle_energy_unit(oracle_ws *w, const double *xyz, const double *velo, int32_t style, double *eu);
int64_t oracle_ws_lists(oracle_ws *w, int32_t which, const int32_t **pairs, const double **coef);

/* ---- Brownian dynamics with hydrodynamic interaction ----
 * simu_hydro_tensors.F90:17-324, force_fene.F90, force_wca.F90,
 * time_integral.F90:429-457.  Coordinates (3,n); D, L (3n,3n) column-major. */
void oracle_rpy_fill(int32_t n, const double *xyz, const double *radius, double tempk, double visc,
                     int32_t type, double *D);
int32_t oracle_cholesky(int32_t n3, const double *D, double *L);
void oracle_force_fene(int32_t nfene, const int32_t *mp, const double *nat, const double *coef,
                       const double *dist2, const double *xyz, double *f);
void oracle_force_wca(int32_t nwca, const int32_t *mp, const double *nat2, const double *coef,
                      double cutoff, const double *xyz, double *f);
void oracle_bdhi_step(int32_t n, const double *D, const double *L, const double *F, double tempk,
                      double tstep, uint64_t seed, uint32_t key, uint64_t step, double *xyz);

/* ---- Langevin (time_integral_pre.F90:160-180, time_integral.F90:149-189) ---- */
void oracle_langevin_consts(const tis_system *s, double tstep, double *rlan /* [4*nmp] */);
/* Runs nsteps; state arrays (3,N) are updated in place.  The neighbour list is
 * rebuilt every rebuild_every steps (>0) or on the reference's displacement
 * trigger (rebuild_every == 0, time_integral.F90:57-93). */
int32_t oracle_langevin_run(oracle_ws *w, double tstep, int64_t nsteps, int32_t rebuild_every,
                            double *xyz, double *velo, double *accel,
                            uint64_t seed, uint32_t sys, uint64_t step0, int32_t init_accel);

#ifdef __cplusplus
}
#endif
#endif
