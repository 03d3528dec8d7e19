/*
 * tis_oracle.c -- FP64 CPU restatement of the reference's DT15 per-step
 * force-and-integrate path.  TEST INFRASTRUCTURE ONLY: the oracle for the
 * HIP library and the CPU baseline timed beside it.  It is built with
 * -ffp-contract=off so that every a*b+c rounds twice, as the reference's
 * Fortran does on baseline x86-64.
 *
 * Each function restates the reference routine named in its comment
 * (paths relative to the reference's src/).  Operation order follows the
 * Fortran expressions where it matters for rounding (cutoff tests).
 */
#include "tis_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F_PI 3.14159265358979323846264338
#define F_2PI (2.0 * F_PI)
#define DE_MAX 50.0            /* const_physical.F90:35 */
#define ZERO_JUDGE 1.0e-6      /* const_physical.F90:42 */
#define HIGH_ENERGY 1.0e10     /* const_physical.F90:44 */

/* const_physical.F90:37-38: cos(0.57/180.0*F_PI) with 0.57/180.0 evaluated in
 * default (single) real before promotion. */
static double max_abscos_hb15(void) {
  float r = 0.57f / 180.0f;
  return cos((double)r * F_PI);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
 * as easy as 1, 2, 3").  Multipliers and Weyl constants from that paper. */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
    mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in the open interval (0,1) from two 32-bit words. */
static inline double u01_53(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)a << 21) ^ ((uint64_t)b >> 11);
  return ((double)m + 0.5) * 0x1.0p-53;
}

static void philox_at(uint64_t seed, uint32_t sys, uint64_t step, uint32_t lane,
                      uint32_t tag, uint32_t out[4]) {
  uint32_t ctr[4] = {lane, (uint32_t)step, sys,
                     ((uint32_t)(step >> 32) & 0xFFFFu) | (tag << 16)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  oracle_philox4x32(ctr, key, out);
}

/* Three N(0,1) numbers for (system, step, bead) by the Box-Muller transform
 * (Box & Muller 1958).  Replaces rfunc_boxmuller.F90:8-73 (Marsaglia polar on
 * an MT19937 stream); the distribution is the same, the sequence is not. */
void oracle_gauss3(uint64_t seed, uint32_t sys, uint64_t step, uint32_t bead, double g[3]) {
  uint32_t r[4];
  philox_at(seed, sys, step, bead, 0u, r);
  double u1 = u01_53(r[0], r[1]), u2 = u01_53(r[2], r[3]);
  double rad = sqrt(-2.0 * log(u1));
  double th = F_2PI * u2;
  g[0] = rad * cos(th);
  g[1] = rad * sin(th);
  philox_at(seed, sys, step, bead, 1u, r);
  u1 = u01_53(r[0], r[1]); u2 = u01_53(r[2], r[3]);
  rad = sqrt(-2.0 * log(u1));
  th = F_2PI * u2;
  g[2] = rad * cos(th);
}

double oracle_uniform(uint64_t seed, uint32_t sys, uint64_t step, uint32_t stream, uint32_t idx) {
  uint32_t r[4];
  philox_at(seed, sys, step, idx, 2u + stream, r);
  return u01_53(r[0], r[1]);
}

/* ------------------------------------------------------------------ */
/* util_pbneighbor.F90:22-35: single-shift minimum image; mirror index
 * 9*i1+3*i2+i3 (0-based here). */
static int pbneighbor(const tis_system *s, double v[3]) {
  int imi[3];
  for (int d = 0; d < 3; ++d) {
    double L = s->p.box[d], h = L / 2.0;
    if (v[d] > h) { v[d] = v[d] - L; imi[d] = 0; }
    else if (v[d] < -h) { v[d] = v[d] + L; imi[d] = 1; }
    else imi[d] = 2;
  }
  return 9 * imi[0] + 3 * imi[1] + imi[2];
}

/* setp_periodic.F90:80-115: d_mirror(:, imirror). Non-periodic runs read
 * entry 1, which is zero. */
void oracle_mirror_shift(const tis_system *s, int32_t imirror, double d[3]) {
  if (!s->p.periodic) { d[0] = d[1] = d[2] = 0.0; return; }
  int idx[3] = {imirror / 9, (imirror / 3) % 3, imirror % 3};
  for (int k = 0; k < 3; ++k)
    d[k] = idx[k] == 0 ? -s->p.box[k] : (idx[k] == 1 ? s->p.box[k] : 0.0);
}

/* util_periodic.F90:26-36 */
void oracle_wrap(const tis_system *s, double *pxyz) {
  if (!s->p.periodic) return;
  for (int i = 0; i < s->nmp; ++i)
    for (int d = 0; d < 3; ++d) {
      double pw = s->p.box[d], pmax = 0.5 * pw, pmin = -0.5 * pw;
      double tx = pxyz[3 * i + d];
      if (tx > pmax) tx = tx - pw * (double)((int)((tx - pmax) / pw) + 1);
      else if (tx < pmin) tx = tx + pw * (double)((int)((pmin - tx) / pw) + 1);
      pxyz[3 * i + d] = tx;
    }
}

static inline double sq3(const double v[3]) {
  /* v(1)**2 + v(2)**2 + v(3)**2, left to right, no contraction */
  double a = v[0] * v[0];
  double b = v[1] * v[1];
  double c = v[2] * v[2];
  return (a + b) + c;
}

/* pair vector j - i for list building: neighbor_list.F90:65-73 */
static int pair_vec(const tis_system *s, const double *x, int i, int j, double v[3]) {
  for (int d = 0; d < 3; ++d) v[d] = x[3 * j + d] - x[3 * i + d];
  if (s->p.periodic) return pbneighbor(s, v);
  return 0; /* "mirror 1" of the reference (0-based 0): zero shift */
}

/* neighbor_assign.F90:481-559: DT15 keep rule for i < j */
static int exv_keep(const tis_system *s, int i, int j) {
  int ui = s->unit[i], uj = s->unit[j];
  int ci = s->iclass[i];
  if (ui == uj) {
    if (ci == TIS_CLASS_RNA) {
      int sep;
      switch (s->mptype[i]) {
        case TIS_MP_RNA_PHOS: sep = s->p.n_sep_P; break;
        case TIS_MP_RNA_SUGAR: sep = s->p.n_sep_S; break;
        case TIS_MP_RNA_BASE: sep = s->p.n_sep_B; break;
        default: return 0;
      }
      return j >= i + sep;
    }
    return 1; /* ION or LIG unit */
  }
  return 1; /* RNA-RNA, RNA-ION, ION-LIG (model 2015) */
}

int64_t oracle_list_exv(const tis_system *s, const double *pxyz, int32_t *pairs, int64_t cap) {
  int64_t n = 0;
  for (int i = 0; i < s->nmp; ++i)
    for (int j = i + 1; j < s->nmp; ++j) {
      double v[3];
      int mir = pair_vec(s, pxyz, i, j, v);
      double d2 = sq3(v);
      double r2 = s->rlist2_unit[s->unit[i] * s->nunit + s->unit[j]];
      if (!(d2 < r2)) continue;
      if (!exv_keep(s, i, j)) continue;
      if (n >= cap) return -1;
      pairs[3 * n] = i; pairs[3 * n + 1] = j; pairs[3 * n + 2] = mir;
      ++n;
    }
  return n;
}

/* neighbor_list_ele.F90:73-90 (neighbour-dynamic radius) */
static double ele_rlist2(const tis_system *s) {
  double m = s->p.neigh_margin;
  if (s->p.ele_form == TIS_ELE_DH) {
    if (s->p.dh_cutoff_type == 0) {
      double r = m + s->p.cutoff_ele * s->p.ele_cdist;
      return r * r;
    }
    double r = m + s->p.cutoff_ele;
    return r * r;
  }
  double r = m + s->p.cutoff_ele;
  return r * r;
}

int64_t oracle_list_ele(const tis_system *s, const double *pxyz, int32_t *pairs,
                        double *coef, int64_t cap) {
  if (s->p.ele_form == TIS_ELE_NONE) return 0;
  double r2 = ele_rlist2(s);
  int64_t n = 0;
  for (int i = 0; i < s->nmp; ++i) {
    if (fabs(s->charge[i]) < ZERO_JUDGE) continue;
    for (int j = i + 1; j < s->nmp; ++j) {
      if (fabs(s->charge[j]) < ZERO_JUDGE) continue;
      double v[3];
      int mir = pair_vec(s, pxyz, i, j, v);
      double d2 = sq3(v);
      if (!(d2 < r2)) continue;
      if (n >= cap) return -1;
      pairs[3 * n] = i; pairs[3 * n + 1] = j; pairs[3 * n + 2] = mir;
      coef[n] = (s->charge[i] * s->charge[j]) * s->p.ele_coef;
      ++n;
    }
  }
  return n;
}

int32_t oracle_list_hb(const tis_system *s, const double *xyz, const double *pxyz, int32_t *hb_idx) {
  int32_t n = 0;
  for (int h = 0; h < s->nhb; ++h) {
    int a = s->hb_mp[6 * h], b = s->hb_mp[6 * h + 1];
    double v[3];
    if (s->p.periodic) {
      for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * a + d] - pxyz[3 * b + d];
      pbneighbor(s, v);
    } else {
      for (int d = 0; d < 3; ++d) v[d] = xyz[3 * a + d] - xyz[3 * b + d];
    }
    double d2 = sq3(v);
    double r = (s->hb_nat[6 * h] + s->p.hb_cutoff_dist) + s->p.neigh_margin;
    if (d2 < r * r) hb_idx[n++] = h;
  }
  return n;
}

/* ------------------------------------------------------------------ */
void oracle_force_bond(const tis_system *s, const double *xyz, double *f) {
  for (int b = 0; b < s->nbd; ++b) {
    int i = s->bd_mp[2 * b], j = s->bd_mp[2 * b + 1];
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = xyz[3 * j + d] - xyz[3 * i + d];
    double dist = sqrt(sq3(v));
    double dd = dist - s->bd_nat[b];
    double fr = s->bd_coef[b] * (-2.0 * dd / dist);
    for (int d = 0; d < 3; ++d) {
      double fd = fr * v[d];
      f[3 * i + d] -= fd;
      f[3 * j + d] += fd;
    }
  }
}

void oracle_force_bangle(const tis_system *s, const double *xyz, double *f) {
  for (int a = 0; a < s->nba; ++a) {
    if (s->ba_coef[a] < ZERO_JUDGE) continue;
    int i1 = s->ba_mp[3 * a], i2 = s->ba_mp[3 * a + 1], i3 = s->ba_mp[3 * a + 2];
    double v21[3], v32[3];
    for (int d = 0; d < 3; ++d) {
      v21[d] = xyz[3 * i2 + d] - xyz[3 * i1 + d];
      v32[d] = xyz[3 * i3 + d] - xyz[3 * i2 + d];
    }
    double c11 = (v21[0] * v21[0] + v21[1] * v21[1]) + v21[2] * v21[2];
    double c22 = (v32[0] * v32[0] + v32[1] * v32[1]) + v32[2] * v32[2];
    double c21 = (v32[0] * v21[0] + v32[1] * v21[1]) + v32[2] * v21[2];
    double ct = -c21 / sqrt(c11 * c22);
    if (ct > 1.0) ct = 1.0; else if (ct < -1.0) ct = -1.0;
    double dba = acos(ct) - s->ba_nat[a];
    double t3 = c11 * c22 - c21 * c21;
    if (t3 <= 1.0) t3 = 1.0;
    double fr = 2.0 * s->ba_coef[a] * dba / sqrt(t3);
    for (int d = 0; d < 3; ++d) {
      double f21 = fr * (v21[d] * (c21 / c11) - v32[d]);
      double f32 = fr * (v32[d] * (c21 / c22) - v21[d]);
      f[3 * i1 + d] -= f21;
      f[3 * i2 + d] += f21 - f32;
      f[3 * i3 + d] += f32;
    }
  }
}

static double exv_dij(const tis_system *s, int i, int j) {
  /* force_exv_dt15.F90:74-80 (model 2015) */
  int ti = s->mptype[i], tj = s->mptype[j];
  if ((ti == TIS_MP_RNA_PHOS && tj == TIS_MP_RNA_SUGAR) ||
      (ti == TIS_MP_RNA_SUGAR && tj == TIS_MP_RNA_PHOS))
    return s->p.exv_dist_PS;
  if (s->iclass[i] == TIS_CLASS_RNA && s->iclass[j] == TIS_CLASS_RNA) return s->p.exv_dist;
  return s->exv_rad[i] + s->exv_rad[j];
}

int32_t oracle_force_exv(const tis_system *s, const double *pxyz, const int32_t *pairs,
                         int64_t n, double *f) {
  double a = s->p.exv_adjust, a2 = a * a;
  int32_t nclamp = 0;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double dij = exv_dij(s, i, j);
    double dist = sqrt(sq3(v));
    if (dist > dij) continue;
    double coef = s->exv_eps[i] * s->exv_eps[j];
    double dr = dist + a - dij;
    double dr2 = dr * dr;
    double r2 = a2 / dr2, r4 = r2 * r2, r8 = r4 * r4, r14 = r2 * r4 * r8;
    double dv = fabs(12.0 * coef * (r14 - r8) * dr / a2 / dist);
    if (dv > DE_MAX) { dv = DE_MAX; ++nclamp; }
    for (int d = 0; d < 3; ++d) {
      double fd = dv * v[d];
      f[3 * j + d] += fd;
      f[3 * i + d] -= fd;
    }
  }
  return nclamp;
}

int32_t oracle_force_dh(const tis_system *s, const double *pxyz, const int32_t *pairs,
                        const double *coef, int64_t n, double *f) {
  double c = s->p.cutoff_ele * s->p.ele_cdist;
  double cutoff2 = c * c;
  double rcdist = 1.0 / s->p.ele_cdist;
  int32_t nclamp = 0;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double d2 = sq3(v);
    if (d2 > cutoff2) continue;
    double d1 = sqrt(d2);
    double rd = 1.0 / d1;
    double dv = coef[k] * rd * rd * (rd + rcdist) * exp(-d1 * rcdist);
    if (dv > DE_MAX) { dv = DE_MAX; ++nclamp; }
    for (int d = 0; d < 3; ++d) {
      double fd = dv * v[d];
      f[3 * i + d] -= fd;
      f[3 * j + d] += fd;
    }
  }
  return nclamp;
}

static inline double dot3(const double a[3], const double b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* ---- plain Coulomb and Ewald ---- */

/* force_ele_coulomb.F90:63-84: dv = coef_ij / r^3 inside cutoff_ele, no clamp */
void oracle_force_coulomb(const tis_system *s, const double *pxyz, const int32_t *pairs,
                          const double *coef, int64_t n, double *f) {
  const double cutoff2 = s->p.cutoff_ele * s->p.cutoff_ele;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (d2 > cutoff2) continue;
    double d1 = sqrt(d2);
    double dv = coef[k] / (d1 * d2);
    for (int d = 0; d < 3; ++d) {
      double fd = dv * v[d];
      f[3 * i + d] -= fd;
      f[3 * j + d] += fd;
    }
  }
}

/* Per-unit-pair energies (energy_unit of energy_sumup.F90:17-36): while
 * eu_acc is set, every energy term also adds its energy to
 * eu_acc[(unit(i), unit(j), type)] (column-major nunit x nunit x TIS_E_MAX)
 * with the beads each term of the reference uses for (iunit, junit). */
static __thread double *eu_acc = NULL;
static __thread const tis_system *eu_sys = NULL;
static void eu_add(int type, int i, int j, double e) {
  if (!eu_acc) return;
  const int nu = eu_sys->nunit;
  eu_acc[(size_t)eu_sys->unit[i] + (size_t)nu * eu_sys->unit[j] + (size_t)nu * nu * type] += e;
}

/* energy_ele_coulomb.F90:47-80: coef_ij (1/r - 1/r_c) */
double oracle_energy_coulomb(const tis_system *s, const double *pxyz, const int32_t *pairs,
                             const double *coef, int64_t n) {
  const double cutoff2 = s->p.cutoff_ele * s->p.cutoff_ele, cut_inv = 1.0 / s->p.cutoff_ele;
  double e = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (d2 > cutoff2) continue;
    e += coef[k] * (1.0 / sqrt(d2) - cut_inv);
    eu_add(TIS_E_ELE, i, j, coef[k] * (1.0 / sqrt(d2) - cut_inv));
  }
  return e;
}

/* mloop_ewld.F90:52-136: integer vectors h != 0 with |h|^2 <= hmax^2, ordered
 * by |h|^2 (ties in generation order: the reference's comb sort orders ties
 * differently, which changes only the FP64 summation order), coefficient
 * f(h) = 2 pi / L^3 exp(-k^2 / 4 alpha^2) / k^2 with k = 2 pi h / L and L the
 * first box edge (cubic box).  Returns the count (h, coef may be NULL). */
int32_t oracle_ewald_kvec(const tis_system *s, int32_t *h, double *coef) {
  const int hmax = s->p.ewld_hmax, lmax = hmax;
  const double hmax2 = (double)hmax * (double)hmax;
  int32_t nk = 0;
  for (int r2 = 1; r2 <= hmax * hmax; ++r2)
    for (int ix = -lmax; ix <= lmax; ++ix)
      for (int iy = -lmax; iy <= lmax; ++iy)
        for (int iz = -lmax; iz <= lmax; ++iz) {
          if (ix * ix + iy * iy + iz * iz != r2 || (double)r2 > hmax2) continue;
          if (h) { h[3 * nk] = ix; h[3 * nk + 1] = iy; h[3 * nk + 2] = iz; }
          if (coef) {
            const double L = s->p.box[0], a = s->p.ewld_alpha;
            const double factor = 2.0 * F_PI / L;
            coef[nk] = 2.0 * F_PI / (L * L * L) * exp(-factor * factor * (double)r2 / (4.0 * a * a)) /
                       (factor * factor * (double)r2);
          }
          ++nk;
        }
  return nk;
}

/* force_ele_coulomb_ewld.F90:56-137.  Reciprocal part, per h:
 * S = sum_q q (cos, sin)(k.r_q); F_q += C 2 f(h) (q sin_q Sc - q cos_q Ss) k.
 * Real part over the list: dv = coef_ij (erfc(a r)/r + 2a/sqrt(pi)
 * exp(-a^2 r^2)) / r^2 inside cutoff_ele, no clamp. */
void oracle_force_ewald(const tis_system *s, const double *pxyz, const int32_t *pairs,
                        const double *coef, int64_t n, double *f) {
  const int32_t nk = oracle_ewald_kvec(s, NULL, NULL);
  int32_t *h = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(nk + 1));
  double *fk = (double *)malloc(sizeof(double) * (size_t)(nk + 1));
  oracle_ewald_kvec(s, h, fk);
  const double factor = 2.0 * F_PI / s->p.box[0], C = s->p.ele_coef;
  int nq = 0;
  for (int i = 0; i < s->nmp; ++i) if (fabs(s->charge[i]) >= ZERO_JUDGE) ++nq;
  int *qi = (int *)malloc(sizeof(int) * (size_t)(nq + 1));
  double *qc = (double *)malloc(sizeof(double) * (size_t)(nq + 1));
  double *qs = (double *)malloc(sizeof(double) * (size_t)(nq + 1));
  nq = 0;
  for (int i = 0; i < s->nmp; ++i) if (fabs(s->charge[i]) >= ZERO_JUDGE) qi[nq++] = i;
  for (int g = 0; g < nk; ++g) {
    const double kv[3] = {factor * h[3 * g], factor * h[3 * g + 1], factor * h[3 * g + 2]};
    double scos = 0.0, ssin = 0.0;
    for (int q = 0; q < nq; ++q) {
      const double *r = pxyz + 3 * qi[q];
      double dp = (kv[0] * r[0] + kv[1] * r[1]) + kv[2] * r[2];
      double qq = s->charge[qi[q]];
      qc[q] = qq * cos(dp); qs[q] = qq * sin(dp);
    }
    for (int q = 0; q < nq; ++q) { scos += qc[q]; ssin += qs[q]; }
    for (int q = 0; q < nq; ++q) {
      double dv = C * 2.0 * fk[g] * (qs[q] * scos - qc[q] * ssin);
      for (int d = 0; d < 3; ++d) f[3 * qi[q] + d] += dv * kv[d];
    }
  }
  free(h); free(fk); free(qi); free(qc); free(qs);
  const double a = s->p.ewld_alpha, a2 = a * a, beta = 2.0 * a / sqrt(F_PI);
  const double cutoff2 = s->p.cutoff_ele * s->p.cutoff_ele;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (d2 > cutoff2) continue;
    double d1 = sqrt(d2);
    double dv = coef[k] * (erfc(a * d1) / d1 + beta * exp(-a2 * d2)) / d2;
    for (int d = 0; d < 3; ++d) {
      double fd = dv * v[d];
      f[3 * i + d] -= fd;
      f[3 * j + d] += fd;
    }
  }
}

/* energy_ele_coulomb_ewld.F90:56-212 (no dipole term): Fourier
 * sum_h C f(h) (Sc^2 + Ss^2) + real sum coef_ij erfc(a r)/r + self
 * C (-a/sqrt(pi)) sum_q q^2 (mloop_ewld.F90:146-154). */
double oracle_energy_ewald(const tis_system *s, const double *pxyz, const int32_t *pairs,
                           const double *coef, int64_t n, double parts[3]) {
  const int32_t nk = oracle_ewald_kvec(s, NULL, NULL);
  int32_t *h = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(nk + 1));
  double *fk = (double *)malloc(sizeof(double) * (size_t)(nk + 1));
  oracle_ewald_kvec(s, h, fk);
  const double factor = 2.0 * F_PI / s->p.box[0], C = s->p.ele_coef;
  double efour = 0.0, ereal = 0.0, eself = 0.0, q2 = 0.0;
  for (int g = 0; g < nk; ++g) {
    const double kv[3] = {factor * h[3 * g], factor * h[3 * g + 1], factor * h[3 * g + 2]};
    double scos = 0.0, ssin = 0.0;
    for (int i = 0; i < s->nmp; ++i) {
      if (fabs(s->charge[i]) < ZERO_JUDGE) continue;
      const double *r = pxyz + 3 * i;
      double dp = (kv[0] * r[0] + kv[1] * r[1]) + kv[2] * r[2];
      scos += s->charge[i] * cos(dp);
      ssin += s->charge[i] * sin(dp);
    }
    efour += C * fk[g] * (scos * scos + ssin * ssin);
  }
  free(h); free(fk);
  const double a = s->p.ewld_alpha, cutoff2 = s->p.cutoff_ele * s->p.cutoff_ele;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (d2 > cutoff2) continue;
    double d1 = sqrt(d2);
    ereal += coef[k] * erfc(a * d1) / d1;
  }
  for (int i = 0; i < s->nmp; ++i)
    if (fabs(s->charge[i]) >= ZERO_JUDGE) q2 += s->charge[i] * s->charge[i];
  eself = C * (-a / sqrt(F_PI)) * q2;
  if (parts) { parts[0] = efour; parts[1] = ereal; parts[2] = eself; }
  return (efour + ereal) + eself;
}
static inline void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double wrap_pi(double d) {
  if (d > F_PI) return d - F_2PI;
  if (d < -F_PI) return d + F_2PI;
  return d;
}
static inline void sub3(const double *x, int a, int b, double v[3]) {
  for (int d = 0; d < 3; ++d) v[d] = x[3 * a + d] - x[3 * b + d];
}

/* force_dtrna_stack.F90:53-165 */
void oracle_force_stack(const tis_system *s, const double *xyz, const uint8_t *st_status, double *f) {
  for (int st = 0; st < s->nst; ++st) {
    if (st_status && !st_status[st]) continue;
    const int32_t *m7 = s->st_mp + 7 * st;
    const double *nat = s->st_nat + 3 * st, *cf = s->st_coef + 4 * st;
    double v21[3], v34[3], v54[3], v56[3], v76[3], m[3], n[3], fi[3], fl[3];
    double fo[7][3];
    memset(fo, 0, sizeof(fo));
    sub3(xyz, m7[1], m7[0], v21);
    sub3(xyz, m7[2], m7[3], v34);
    sub3(xyz, m7[4], m7[3], v54);
    sub3(xyz, m7[4], m7[5], v56);
    sub3(xyz, m7[6], m7[5], v76);
    double ediv = 1.0;
    double dist = sqrt(sq3(v21));
    double dd = dist - nat[0];
    ediv += cf[1] * dd * dd;
    for (int d = 0; d < 3; ++d) {
      fi[d] = 2.0 * cf[1] * dd / dist * v21[d];
      fo[0][d] = -fi[d]; fo[1][d] = fi[d];
    }
    /* dihedral 3-4-5-6 */
    cross3(v34, v54, m);
    cross3(v54, v56, n);
    double dmm = dot3(m, m), dnn = dot3(n, n), d5454 = dot3(v54, v54);
    double abs54 = sqrt(d5454), d5654 = dot3(v56, v54);
    double d3454o = dot3(v34, v54) / d5454, d5654o = d5654 / d5454;
    double dih = atan2(dot3(v34, n) * sqrt(d5454), dot3(m, n));
    double dph = wrap_pi(dih - nat[1]);
    ediv += cf[2] * dph * dph;
    for (int d = 0; d < 3; ++d) {
      fi[d] = 2.0 * cf[2] * dph * abs54 / dmm * m[d];
      fl[d] = -2.0 * cf[2] * dph * abs54 / dnn * n[d];
      fo[2][d] = fi[d];
      fo[3][d] = (-1.0 + d3454o) * fi[d] - d5654o * fl[d];
      fo[4][d] = (-1.0 + d5654o) * fl[d] - d3454o * fi[d];
      fo[5][d] = fl[d];
    }
    /* dihedral 7-6-5-4 */
    cross3(v76, v56, m);
    for (int d = 0; d < 3; ++d) n[d] = -n[d];
    dmm = dot3(m, m);
    double d5656 = dot3(v56, v56), abs56 = sqrt(d5656);
    double d7656o = dot3(v76, v56) / d5656, d5456o = d5654 / d5656;
    dih = atan2(dot3(v76, n) * sqrt(d5656), dot3(m, n));
    dph = wrap_pi(dih - nat[2]);
    ediv += cf[3] * dph * dph;
    for (int d = 0; d < 3; ++d) {
      fi[d] = 2.0 * cf[3] * dph * abs56 / dmm * m[d];
      fl[d] = -2.0 * cf[3] * dph * abs56 / dnn * n[d];
      fo[6][d] += fi[d];
      fo[5][d] += (-1.0 + d7656o) * fi[d] - d5456o * fl[d];
      fo[4][d] += -d7656o * fi[d] + (-1.0 + d5456o) * fl[d];
      fo[3][d] += fl[d];
    }
    double sc = cf[0] / (ediv * ediv);
    for (int k = 0; k < 7; ++k)
      for (int d = 0; d < 3; ++d) f[3 * m7[k] + d] += sc * fo[k][d];
  }
}

/* force_dtrna_stack_nlocal.F90:48-263 */
void oracle_force_tstack(const tis_system *s, const double *xyz, const double *pxyz,
                         uint8_t *st_status, double *f) {
  for (int k = 0; k < s->nst; ++k) if (st_status) st_status[k] = 1;
  for (int t = 0; t < s->ntst; ++t) {
    const int32_t *m6 = s->tst_mp + 6 * t;
    const double *nat = s->tst_nat + 6 * t, *cf = s->tst_coef + 7 * t;
    const double *x = s->p.periodic ? pxyz : xyz;
    double v12[3], v13[3], v53[3], v42[3], v46[3];
    sub3(x, m6[0], m6[1], v12);
    if (s->p.periodic) pbneighbor(s, v12);
    double d1212 = dot3(v12, v12), a12 = sqrt(d1212);
    double dd = a12 - nat[0];
    if (dd < 10.0 && st_status) {
      if (s->tst_excl[2 * t] && s->tst2st[2 * t] >= 0) st_status[s->tst2st[2 * t]] = 0;
      if (s->tst_excl[2 * t + 1] && s->tst2st[2 * t + 1] >= 0) st_status[s->tst2st[2 * t + 1]] = 0;
    }
    sub3(x, m6[0], m6[2], v13);
    sub3(x, m6[4], m6[2], v53);
    sub3(x, m6[3], m6[1], v42);
    sub3(x, m6[3], m6[5], v46);
    if (s->p.periodic) { pbneighbor(s, v13); pbneighbor(s, v53); pbneighbor(s, v42); pbneighbor(s, v46); }
    double fo[6][3], fi[3], fk[3], fl[3];
    memset(fo, 0, sizeof(fo));
    double d1313 = dot3(v13, v13), d4242 = dot3(v42, v42);
    double a13 = sqrt(d1313), a42 = sqrt(d4242);
    double d1213 = dot3(v13, v12), d1242 = dot3(v12, v42), d4246 = dot3(v42, v46), d1353 = dot3(v13, v53);
    double d1213o1212 = d1213 / d1212, d1213o1313 = d1213 / d1313, d1242o1212 = d1242 / d1212;
    double d4246o4242 = d4246 / d4242, d1242o4242 = d1242 / d4242, d1353o1313 = d1353 / d1313;
    double ediv = 1.0;
    ediv += cf[1] * dd * dd;
    for (int d = 0; d < 3; ++d) {
      fi[d] = 2.0 * cf[1] * dd * v12[d] / a12;
      fo[0][d] = fi[d]; fo[1][d] = -fi[d];
    }
    double ct = d1213 / (a13 * a12);
    double dang = acos(ct) - nat[1];
    double pre = 2.0 * cf[2] * dang / sqrt(d1313 * d1212 - d1213 * d1213);
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre * (v12[d] - d1213o1313 * v13[d]);
      fk[d] = pre * (v13[d] - d1213o1212 * v12[d]);
      fo[2][d] += fi[d]; fo[1][d] += fk[d]; fo[0][d] = fo[0][d] - fi[d] - fk[d];
    }
    ediv += cf[2] * dang * dang;
    ct = d1242 / (a12 * a42);
    dang = acos(ct) - nat[2];
    pre = 2.0 * cf[3] * dang / sqrt(d1212 * d4242 - d1242 * d1242);
    for (int d = 0; d < 3; ++d) {
      fi[d] = -pre * (v42[d] - d1242o1212 * v12[d]);
      fk[d] = -pre * (v12[d] - d1242o4242 * v42[d]);
      fo[0][d] += fi[d]; fo[3][d] += fk[d]; fo[1][d] = fo[1][d] - fi[d] - fk[d];
    }
    ediv += cf[3] * dang * dang;
    double c4212[3], c1213[3], m[3], n[3];
    cross3(v42, v12, c4212);
    cross3(v12, v13, c1213);
    double c4212a = dot3(c4212, c4212), c1213a = dot3(c1213, c1213);
    double dih = atan2(dot3(v42, c1213) * sqrt(d1212), dot3(c4212, c1213));
    double dph = wrap_pi(dih - nat[3]);
    ediv += cf[4] * dph * dph;
    pre = 2.0 * cf[4] * dph * a12;
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre / c4212a * c4212[d];
      fl[d] = -pre / c1213a * c1213[d];
      fo[3][d] += fi[d];
      fo[1][d] += (-1.0 + d1242o1212) * fi[d] - d1213o1212 * fl[d];
      fo[0][d] += (-1.0 + d1213o1212) * fl[d] - d1242o1212 * fi[d];
      fo[2][d] += fl[d];
    }
    cross3(v53, v13, m);
    for (int d = 0; d < 3; ++d) n[d] = -c1213[d];
    double dmm = dot3(m, m), dnn = c1213a;
    dih = atan2(dot3(v53, n) * a13, dot3(m, n));
    dph = wrap_pi(dih - nat[4]);
    ediv += cf[5] * dph * dph;
    pre = 2.0 * cf[5] * dph * a13;
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre / dmm * m[d];
      fl[d] = -pre / dnn * n[d];
      fo[4][d] += fi[d];
      fo[2][d] += (-1.0 + d1353o1313) * fi[d] - d1213o1313 * fl[d];
      fo[0][d] += (-1.0 + d1213o1313) * fl[d] - d1353o1313 * fi[d];
      fo[1][d] += fl[d];
    }
    for (int d = 0; d < 3; ++d) m[d] = -c4212[d];
    cross3(v42, v46, n);
    dmm = c4212a; dnn = dot3(n, n);
    dih = atan2(dot3(v12, n) * a42, dot3(m, n));
    dph = wrap_pi(dih - nat[5]);
    ediv += cf[6] * dph * dph;
    pre = 2.0 * cf[6] * dph * a42;
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre / dmm * m[d];
      fl[d] = -pre / dnn * n[d];
      fo[0][d] += fi[d];
      fo[1][d] += (-1.0 + d1242o4242) * fi[d] - d4246o4242 * fl[d];
      fo[3][d] += (-1.0 + d4246o4242) * fl[d] - d1242o4242 * fi[d];
      fo[5][d] += fl[d];
    }
    double sc = cf[0] / (ediv * ediv);
    for (int k = 0; k < 6; ++k)
      for (int d = 0; d < 3; ++d) f[3 * m6[k] + d] += fo[k][d] * sc;
  }
}

/* force_dtrna_hbond15.F90:91-353 (default branch, no _HTN_CONSISTENT) */
void oracle_hbond_eval(const tis_system *s, const double *xyz, const double *pxyz,
                       const int32_t *hb_idx, int32_t nlist,
                       double *for_hb, double *ene_hb, uint8_t *status) {
  const double MAXC = max_abscos_hb15();
  memset(for_hb, 0, sizeof(double) * 18 * (size_t)s->nhb);
  memset(ene_hb, 0, sizeof(double) * (size_t)s->nhb);
  memset(status, 0, (size_t)s->nhb);
  for (int q = 0; q < nlist; ++q) {
    int h = hb_idx[q];
    const int32_t *m6 = s->hb_mp + 6 * h;
    const double *nat = s->hb_nat + 6 * h, *cf = s->hb_coef + 7 * h;
    double *fo = for_hb + 18 * h;  /* fo[3*k + d] */
    double v12[3], v13[3], v53[3], v42[3], v46[3];
    if (s->p.periodic) { sub3(pxyz, m6[0], m6[1], v12); pbneighbor(s, v12); }
    else sub3(xyz, m6[0], m6[1], v12);
    double d1212 = dot3(v12, v12), a12 = sqrt(d1212);
    double dd = a12 - nat[0];
    if (fabs(dd) > s->p.hb_cutoff_dist) continue;
    status[h] = 1;
    double ex = -cf[1] * dd * dd;
    double fi[3], fk[3], fl[3];
    for (int d = 0; d < 3; ++d) {
      fi[d] = (2.0 * cf[1] * dd / a12) * v12[d];
      fo[0 + d] = fi[d]; fo[3 + d] = -fi[d];
    }
    /* v13..v46 use xyz_mp_rep also in periodic runs (force_dtrna_hbond15.F90:152-164) */
    sub3(xyz, m6[0], m6[2], v13);
    sub3(xyz, m6[4], m6[2], v53);
    sub3(xyz, m6[3], m6[1], v42);
    sub3(xyz, m6[3], m6[5], v46);
    if (s->p.periodic) { pbneighbor(s, v13); pbneighbor(s, v53); pbneighbor(s, v42); pbneighbor(s, v46); }
    double d1313 = dot3(v13, v13), d4242 = dot3(v42, v42);
    double a13 = sqrt(d1313), a42 = sqrt(d4242);
    double d1213 = dot3(v13, v12), d1242 = dot3(v12, v42), d4246 = dot3(v42, v46), d1353 = dot3(v13, v53);
    double d1213o1212 = d1213 / d1212, d1213o1313 = d1213 / d1313, d1242o1212 = d1242 / d1212;
    double d4246o4242 = d4246 / d4242, d1242o4242 = d1242 / d4242, d1353o1313 = d1353 / d1313;
    /* angle 3-1=2 */
    double c312 = d1213 / (a13 * a12);
    if (fabs(c312) > MAXC) d1213 = copysign(a12 * a13 * MAXC, c312);
    double da = acos(c312) - nat[1];
    double pre = 2.0 * cf[2] * da / sqrt(d1313 * d1212 - d1213 * d1213);
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre * (v12[d] - d1213o1313 * v13[d]);
      fk[d] = pre * (v13[d] - d1213o1212 * v12[d]);
      fo[6 + d] += fi[d]; fo[3 + d] += fk[d]; fo[0 + d] = fo[0 + d] - fi[d] - fk[d];
    }
    ex = ex - cf[2] * da * da;
    /* angle 1=2-4 */
    double c124 = d1242 / (a12 * a42);
    if (fabs(c124) > MAXC) d1242 = copysign(a12 * a42 * MAXC, c124);
    da = acos(c124) - nat[2];
    pre = 2.0 * cf[3] * da / sqrt(d1212 * d4242 - d1242 * d1242);
    for (int d = 0; d < 3; ++d) {
      fi[d] = -pre * (v42[d] - d1242o1212 * v12[d]);
      fk[d] = -pre * (v12[d] - d1242o4242 * v42[d]);
      fo[0 + d] += fi[d]; fo[9 + d] += fk[d]; fo[3 + d] = fo[3 + d] - fi[d] - fk[d];
    }
    ex = ex - cf[3] * da * da;
    /* dihedral 4-2=1=3 */
    double c4212[3], c1213[3], c5313[3], c4246[3], n[3];
    cross3(v42, v12, c4212);
    cross3(v12, v13, c1213);
    double c4212a = dot3(c4212, c4212), c1213a = dot3(c1213, c1213);
    if (fabs(c124) > MAXC) c4212a = d4242 * d1212 * (1.0 - MAXC * MAXC);
    if (fabs(c312) > MAXC) c1213a = d1212 * d1313 * (1.0 - MAXC * MAXC);
    double dih = atan2(dot3(v42, c1213) * a12, dot3(c4212, c1213));
    double dph = wrap_pi(dih - nat[3]);
    ex = ex - cf[4] * dph * dph;
    pre = 2.0 * cf[4] * dph * a12;
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre / c4212a * c4212[d];
      fl[d] = -pre / c1213a * c1213[d];
      fo[9 + d] += fi[d];
      fo[3 + d] += (-1.0 + d1242o1212) * fi[d] - d1213o1212 * fl[d];
      fo[0 + d] += (-1.0 + d1213o1212) * fl[d] - d1242o1212 * fi[d];
      fo[6 + d] += fl[d];
    }
    /* dihedral 5-3-1=2 */
    cross3(v53, v13, c5313);
    for (int d = 0; d < 3; ++d) n[d] = -c1213[d];
    double c5313a = dot3(c5313, c5313);
    double c531 = d1353 / (a13 * sqrt(sq3(v53)));
    if (fabs(c531) > MAXC) c5313a = dot3(v53, v53) * d1313 * (1.0 - MAXC * MAXC);
    dih = atan2(dot3(v53, n) * a13, dot3(c5313, n));
    dph = wrap_pi(dih - nat[4]);
    ex = ex - cf[5] * dph * dph;
    pre = 2.0 * cf[5] * dph * a13;
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre / c5313a * c5313[d];
      fl[d] = -pre / c1213a * n[d];
      fo[12 + d] += fi[d];
      fo[6 + d] += (-1.0 + d1353o1313) * fi[d] - d1213o1313 * fl[d];
      fo[0 + d] += (-1.0 + d1213o1313) * fl[d] - d1353o1313 * fi[d];
      fo[3 + d] += fl[d];
    }
    /* dihedral 1=2-4-6 */
    for (int d = 0; d < 3; ++d) n[d] = -c4212[d];
    cross3(v42, v46, c4246);
    double c4246a = dot3(c4246, c4246);
    double c246 = d4246 / (a42 * sqrt(sq3(v46)));
    if (fabs(c246) > MAXC) c4246a = dot3(v46, v46) * d4242 * (1.0 - MAXC * MAXC);
    dih = atan2(dot3(v12, c4246) * a42, dot3(n, c4246));
    dph = wrap_pi(dih - nat[5]);
    ex = ex - cf[6] * dph * dph;
    pre = 2.0 * cf[6] * dph * a42;
    for (int d = 0; d < 3; ++d) {
      fi[d] = pre / c4212a * n[d];
      fl[d] = -pre / c4246a * c4246[d];
      fo[0 + d] += fi[d];
      fo[3 + d] += (-1.0 + d1242o4242) * fi[d] - d4246o4242 * fl[d];
      fo[9 + d] += (-1.0 + d4246o4242) * fl[d] - d1242o4242 * fi[d];
      fo[15 + d] += fl[d];
    }
    double ene = cf[0] * exp(ex);
    ene_hb[h] = ene;
    for (int k = 0; k < 18; ++k) fo[k] = ene * fo[k];
  }
}

/* Uniform draws of the valence pass.  kind 0: the site pick of
 * force_dtrna_hbond15.F90:392 (genrand_double4, (0,1]); kind 1: the shuffle
 * and Metropolis draws of :413 and :426 (genrand_double1, [0,1]). */
typedef double (*valence_draw_fn)(void *ctx, int kind);

typedef struct { uint64_t seed; uint32_t sys; uint64_t step; uint32_t next; } philox_src;

static double philox_draw(void *ctx, int kind) {
  (void)kind;
  philox_src *p = (philox_src *)ctx;
  return oracle_uniform(p->seed, p->sys, p->step, 0, p->next++);
}

/* mt_genrand_double1 / _double2 / _double4 of mt_stream.F90:708-785: one
 * draw takes two 32-bit outputs a, b of the stream; with A = a >> 5 and
 * B = b >> 6 (logical shifts), double1 = (A*2^26 + B) / (2^53 - 1) in [0,1],
 * double2 = (A*2^26 + B) / 2^53 in [0,1), double4 = 1 - double2 in (0,1]. */
typedef struct { const uint32_t *raw; int64_t n, pos; int overrun; } mt_src;

static double mt_draw(void *ctx, int kind) {
  mt_src *m = (mt_src *)ctx;
  if (m->pos + 2 > m->n) { m->overrun = 1; return 0.5; }
  uint32_t a = m->raw[m->pos] >> 5, b = m->raw[m->pos + 1] >> 6;
  m->pos += 2;
  double v = (double)a * 67108864.0 + (double)b;
  if (kind == 0) return 1.0 - v * (1.0 / 9007199254740992.0);
  return v * (1.0 / 9007199254740991.0);
}

/* force_dtrna_hbond15.F90:373-510, drawing from `draw`; the per-site
 * membership lists follow neighbor_list_hb.F90:106-134. */
static int32_t hbond_valence_core(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                                  double *ene_hb, uint8_t *status,
                                  valence_draw_fn draw, void *dctx) {
  int ns = s->nhbsite;
  if (ns <= 0) return 0;
  int32_t *excess = (int32_t *)calloc((size_t)ns, sizeof(int32_t));
  int32_t *cnt = (int32_t *)calloc((size_t)ns, sizeof(int32_t));
  int32_t *lst = (int32_t *)malloc(sizeof(int32_t) * (size_t)ns);
  for (int q = 0; q < nlist; ++q) {
    int h = hb_idx[q];
    for (int k = 0; k < 6; ++k) {
      int site = s->hb_site[6 * h + k];
      if (site >= 0) ++cnt[site];
    }
  }
  int32_t *off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ns + 1));
  off[0] = 0;
  for (int i = 0; i < ns; ++i) off[i + 1] = off[i] + cnt[i];
  int32_t *mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(off[ns] + 1));
  memset(cnt, 0, sizeof(int32_t) * (size_t)ns);
  for (int q = 0; q < nlist; ++q) {
    int h = hb_idx[q];
    /* order of appends: site(i,1), site(i,2) for i = 1..3 */
    for (int i = 0; i < 3; ++i)
      for (int side = 0; side < 2; ++side) {
        int site = s->hb_site[6 * h + 2 * i + side];
        if (site >= 0) mem[off[site] + cnt[site]++] = h;
      }
    if (status[h])
      for (int k = 0; k < 6; ++k) {
        int site = s->hb_site[6 * h + k];
        if (site >= 0) ++excess[site];
      }
  }
  for (int i = 0; i < ns; ++i) excess[i] -= s->hbsite_valence[i];
  int32_t ndel = 0;
  int32_t *seq = (int32_t *)malloc(sizeof(int32_t) * (size_t)(off[ns] + 1));
  for (;;) {
    int nex = 0;
    for (int i = 0; i < ns; ++i) if (excess[i] > 0) lst[nex++] = i;
    if (nex == 0) break;
    double rnd = draw(dctx, 0);
    int site = lst[(int)ceil(rnd * nex) - 1];
    int nseq = 0;
    for (int k = off[site]; k < off[site] + cnt[site]; ++k)
      if (status[mem[k]]) seq[nseq++] = mem[k];
    for (int i = 0; i < nseq; ++i) {
      rnd = draw(dctx, 1);
      /* ceiling(rnd * nhb_seq) (:414): a draw of exactly 0 gives index 0,
       * which the reference would read as hb_seq(0); it stays at 1 here */
      int isw = (int)ceil(rnd * nseq) - 1;
      if (isw < 0) isw = 0;
      int t = seq[i]; seq[i] = seq[isw]; seq[isw] = t;
    }
    int del = seq[0];
    for (int i = 1; i < nseq; ++i) {
      int j = seq[i];
      double ratio = exp((ene_hb[j] - ene_hb[del]) * s->p.beta_hb);
      rnd = draw(dctx, 1);
      if (rnd < ratio) del = j;
    }
    status[del] = 0;
    ene_hb[del] = 0.0;
    ++ndel;
    for (int i = 0; i < 3; ++i)
      for (int side = 0; side < 2; ++side) {
        int st = s->hb_site[6 * del + 2 * i + side];
        if (st >= 0 && excess[st] > 0) --excess[st];
      }
  }
  free(seq); free(mem); free(off); free(lst); free(cnt); free(excess);
  return ndel;
}

int32_t oracle_hbond_valence(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                             double *ene_hb, uint8_t *status,
                             uint64_t seed, uint32_t sys, uint64_t step) {
  philox_src p = {seed, sys, step, 0u};
  return hbond_valence_core(s, hb_idx, nlist, ene_hb, status, philox_draw, &p);
}

int32_t oracle_hbond_valence_mt(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                                double *ene_hb, uint8_t *status,
                                const uint32_t *raw, int64_t nraw, int64_t *used) {
  mt_src m = {raw, nraw, 0, 0};
  int32_t nd = hbond_valence_core(s, hb_idx, nlist, ene_hb, status, mt_draw, &m);
  if (used) *used = m.pos;
  return m.overrun ? -1 : nd;
}

void oracle_hbond_scatter(const tis_system *s, const int32_t *hb_idx, int32_t nlist,
                          const double *for_hb, const uint8_t *status, double *f) {
  for (int q = 0; q < nlist; ++q) {
    int h = hb_idx[q];
    if (!status[h]) continue;
    for (int k = 0; k < 6; ++k)
      for (int d = 0; d < 3; ++d) f[3 * s->hb_mp[6 * h + k] + d] += for_hb[18 * h + 3 * k + d];
  }
}

/* ------------------------------------------------------------------ */
/* energy_bond.F90:39-57 */
double oracle_energy_bond(const tis_system *s, const double *xyz) {
  double e = 0.0;
  for (int b = 0; b < s->nbd; ++b) {
    double v[3];
    sub3(xyz, s->bd_mp[2 * b + 1], s->bd_mp[2 * b], v);
    double dd = sqrt(sq3(v)) - s->bd_nat[b];
    e += s->bd_coef[b] * (dd * dd);
    eu_add(TIS_E_BOND, s->bd_mp[2 * b], s->bd_mp[2 * b + 1], s->bd_coef[b] * (dd * dd));
  }
  return e;
}

/* energy_bangle.F90:36-52 + util_bondangle.F90 */
double oracle_energy_bangle(const tis_system *s, const double *xyz) {
  double e = 0.0;
  for (int a = 0; a < s->nba; ++a) {
    int i1 = s->ba_mp[3 * a], i2 = s->ba_mp[3 * a + 1], i3 = s->ba_mp[3 * a + 2];
    double v21[3], v32[3];
    sub3(xyz, i2, i1, v21);
    sub3(xyz, i3, i2, v32);
    double c11 = (v21[0] * v21[0] + v21[1] * v21[1]) + v21[2] * v21[2];
    double c22 = (v32[0] * v32[0] + v32[1] * v32[1]) + v32[2] * v32[2];
    double c21 = (v32[0] * v21[0] + v32[1] * v21[1]) + v32[2] * v21[2];
    double ct = -c21 / sqrt(c11 * c22);
    if (ct > 1.0) ct = 1.0; else if (ct < -1.0) ct = -1.0;
    double dba = acos(ct) - s->ba_nat[a];
    e += s->ba_coef[a] * (dba * dba);
    eu_add(TIS_E_BANGLE, i1, i2, s->ba_coef[a] * (dba * dba));
  }
  return e;
}

/* energy_exv_dt15.F90:60-116 */
double oracle_energy_exv(const tis_system *s, const double *pxyz, const int32_t *pairs, int64_t n) {
  double a = s->p.exv_adjust, e = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double dij = exv_dij(s, i, j);
    double dist = sqrt(sq3(v));
    if (dist > dij) continue;
    dist = dist + a - dij;
    double ene;
    if (dist > s->p.exv_inf) {
      double r = a / dist;
      double r6 = pow(r, 6.0), r12 = pow(r, 12.0);
      ene = s->exv_eps[i] * s->exv_eps[j] * (r12 - 2.0 * r6 + 1.0);
    } else {
      ene = HIGH_ENERGY;
    }
    e += ene;
    eu_add(TIS_E_EXV, i, j, ene);
  }
  return e;
}

/* energy_ele_DH.F90:33-100 (i_temp_independent = 0) */
double oracle_energy_dh(const tis_system *s, const double *pxyz, const int32_t *pairs,
                        const double *coef, int64_t n) {
  double cutoff2;
  if (s->p.dh_cutoff_type == 0) { double c = s->p.cutoff_ele * s->p.ele_cdist; cutoff2 = c * c; }
  else cutoff2 = s->p.cutoff_ele * s->p.cutoff_ele;
  double rcdist = 1.0 / s->p.ele_cdist, e = 0.0;
  for (int64_t k = 0; k < n; ++k) {
    int i = pairs[3 * k], j = pairs[3 * k + 1];
    double sh[3], v[3];
    oracle_mirror_shift(s, pairs[3 * k + 2], sh);
    for (int d = 0; d < 3; ++d) v[d] = pxyz[3 * j + d] - pxyz[3 * i + d] + sh[d];
    double d2 = sq3(v);
    if (d2 > cutoff2) continue;
    double d1 = sqrt(d2);
    e += coef[k] / d1 * exp(-d1 * rcdist);
    eu_add(TIS_E_ELE, i, j, coef[k] / d1 * exp(-d1 * rcdist));
  }
  return e;
}

/* energy_dtrna_stack.F90:85-160 */
double oracle_energy_stack(const tis_system *s, const double *xyz, const uint8_t *st_status) {
  double e = 0.0;
  for (int st = 0; st < s->nst; ++st) {
    if (st_status && !st_status[st]) continue;
    const int32_t *m7 = s->st_mp + 7 * st;
    const double *nat = s->st_nat + 3 * st, *cf = s->st_coef + 4 * st;
    double v21[3], v34[3], v54[3], v56[3], v76[3], m[3], n[3];
    sub3(xyz, m7[1], m7[0], v21);
    sub3(xyz, m7[2], m7[3], v34);
    sub3(xyz, m7[4], m7[3], v54);
    sub3(xyz, m7[4], m7[5], v56);
    sub3(xyz, m7[6], m7[5], v76);
    double ediv = 1.0;
    double dd = sqrt(sq3(v21)) - nat[0];
    ediv += cf[1] * dd * dd;
    cross3(v34, v54, m);
    cross3(v54, v56, n);
    double dih = atan2(dot3(v34, n) * sqrt(sq3(v54)), dot3(m, n));
    double dph = wrap_pi(dih - nat[1]);
    ediv += cf[2] * dph * dph;
    cross3(v76, v56, m);
    for (int d = 0; d < 3; ++d) n[d] = -n[d];
    dih = atan2(dot3(v76, n) * sqrt(sq3(v56)), dot3(m, n));
    dph = wrap_pi(dih - nat[2]);
    ediv += cf[3] * dph * dph;
    e += cf[0] / ediv;
    eu_add(TIS_E_STACK, m7[0], m7[0], cf[0] / ediv);
  }
  return e;
}

/* energy_dtrna_stack_nlocal.F90:55-230 */
double oracle_energy_tstack(const tis_system *s, const double *xyz, const double *pxyz) {
  double e = 0.0;
  for (int t = 0; t < s->ntst; ++t) {
    const int32_t *m6 = s->tst_mp + 6 * t;
    const double *nat = s->tst_nat + 6 * t, *cf = s->tst_coef + 7 * t;
    const double *x = s->p.periodic ? pxyz : xyz;
    double v12[3], v13[3], v53[3], v42[3], v46[3], c4212[3], c1213[3], m[3], n[3];
    sub3(x, m6[0], m6[1], v12);
    if (s->p.periodic) pbneighbor(s, v12);
    double a12 = sqrt(sq3(v12));
    double dd = a12 - nat[0];
    sub3(x, m6[0], m6[2], v13);
    sub3(x, m6[4], m6[2], v53);
    sub3(x, m6[3], m6[1], v42);
    sub3(x, m6[3], m6[5], v46);
    if (s->p.periodic) { pbneighbor(s, v13); pbneighbor(s, v53); pbneighbor(s, v42); pbneighbor(s, v46); }
    double ediv = 1.0 + cf[1] * dd * dd;
    double a13 = sqrt(sq3(v13));
    double da = acos(dot3(v13, v12) / (a13 * a12)) - nat[1];
    ediv += cf[2] * da * da;
    double a42 = sqrt(sq3(v42));
    da = acos(dot3(v12, v42) / (a12 * a42)) - nat[2];
    ediv += cf[3] * da * da;
    cross3(v42, v12, c4212);
    cross3(v12, v13, c1213);
    double dph = wrap_pi(atan2(dot3(v42, c1213) * a12, dot3(c4212, c1213)) - nat[3]);
    ediv += cf[4] * dph * dph;
    cross3(v53, v13, m);
    for (int d = 0; d < 3; ++d) n[d] = -c1213[d];
    dph = wrap_pi(atan2(dot3(v53, n) * a13, dot3(m, n)) - nat[4]);
    ediv += cf[5] * dph * dph;
    for (int d = 0; d < 3; ++d) m[d] = -c4212[d];
    cross3(v42, v46, n);
    dph = wrap_pi(atan2(dot3(v12, n) * a42, dot3(m, n)) - nat[5]);
    ediv += cf[6] * dph * dph;
    e += cf[0] / ediv;
    eu_add(TIS_E_TSTACK, m6[0], m6[1], cf[0] / ediv);
  }
  return e;
}

/* energy_kinetic.F90:33-44 */
double oracle_energy_kinetic(const tis_system *s, const double *velo) {
  double e = 0.0;
  for (int i = 0; i < s->nmp; ++i) {
    const double *v = velo + 3 * i;
    e += 0.5 * s->mass[i] * dot3(v, v);
    eu_add(TIS_E_VELO, i, i, 0.5 * s->mass[i] * dot3(v, v));
  }
  return e;
}

/* ------------------------------------------------------------------ */
struct oracle_ws {
  const tis_system *s;
  double *pxyz;
  int64_t cap_exv, cap_ele;
  int32_t *exv; int64_t nexv;
  int32_t *ele; double *ele_coef; int64_t nele;
  int32_t *hb; int32_t nhbl;
  double *for_hb, *ene_hb;
  uint8_t *hb_status, *st_status;
};

oracle_ws *oracle_ws_new(const tis_system *s) {
  oracle_ws *w = (oracle_ws *)calloc(1, sizeof(oracle_ws));
  w->s = s;
  int64_t n = s->nmp;
  w->pxyz = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  w->cap_exv = n * (n - 1) / 2 + 1;
  w->cap_ele = w->cap_exv;
  w->exv = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)w->cap_exv);
  w->ele = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)w->cap_ele);
  w->ele_coef = (double *)malloc(sizeof(double) * (size_t)w->cap_ele);
  w->hb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->nhb + 1));
  w->for_hb = (double *)malloc(sizeof(double) * 18 * (size_t)(s->nhb + 1));
  w->ene_hb = (double *)malloc(sizeof(double) * (size_t)(s->nhb + 1));
  w->hb_status = (uint8_t *)malloc((size_t)(s->nhb + 1));
  w->st_status = (uint8_t *)malloc((size_t)(s->nst + 1));
  return w;
}

void oracle_ws_free(oracle_ws *w) {
  if (!w) return;
  free(w->pxyz); free(w->exv); free(w->ele); free(w->ele_coef); free(w->hb);
  free(w->for_hb); free(w->ene_hb); free(w->hb_status); free(w->st_status);
  free(w);
}

/* neighbor.F90:35-76: wrap, exv list, ele list, hb list */
void oracle_neighbor(oracle_ws *w, const double *xyz) {
  const tis_system *s = w->s;
  memcpy(w->pxyz, xyz, sizeof(double) * 3 * (size_t)s->nmp);
  oracle_wrap(s, w->pxyz);
  w->nexv = oracle_list_exv(s, w->pxyz, w->exv, w->cap_exv);
  w->nele = oracle_list_ele(s, w->pxyz, w->ele, w->ele_coef, w->cap_ele);
  w->nhbl = oracle_list_hb(s, xyz, w->pxyz, w->hb);
}

int64_t oracle_ws_lists(oracle_ws *w, int32_t which, const int32_t **pairs, const double **coef) {
  if (which == 0) { *pairs = w->exv; if (coef) *coef = NULL; return w->nexv; }
  if (which == 1) { *pairs = w->ele; if (coef) *coef = w->ele_coef; return w->nele; }
  *pairs = w->hb; if (coef) *coef = NULL; return w->nhbl;
}

/* force_sumup.F90:66-289 for a DT15 system.  Non-periodic pair terms read
 * xyz itself (pxyz_mp_rep equals xyz_mp_rep there); periodic ones read the
 * wrapped copy, which the Langevin loop advances by the same dxyz as xyz
 * (time_integral.F90:156-160). */
static int32_t force_sumup_core(oracle_ws *w, const double *xyz, double *f,
                                 valence_draw_fn draw, void *dctx) {
  const tis_system *s = w->s;
  memset(f, 0, sizeof(double) * 3 * (size_t)s->nmp);
  int32_t nclamp = 0;
  oracle_force_bond(s, xyz, f);
  oracle_force_bangle(s, xyz, f);
  oracle_force_tstack(s, xyz, w->pxyz, w->st_status, f);
  oracle_force_stack(s, xyz, w->st_status, f);
  oracle_hbond_eval(s, xyz, w->pxyz, w->hb, w->nhbl, w->for_hb, w->ene_hb, w->hb_status);
  hbond_valence_core(s, w->hb, w->nhbl, w->ene_hb, w->hb_status, draw, dctx);
  oracle_hbond_scatter(s, w->hb, w->nhbl, w->for_hb, w->hb_status, f);
  const double *px = s->p.periodic ? w->pxyz : xyz;
  nclamp += oracle_force_exv(s, px, w->exv, w->nexv, f);
  if (s->p.ele_form == TIS_ELE_DH)
    nclamp += oracle_force_dh(s, px, w->ele, w->ele_coef, w->nele, f);
  else if (s->p.ele_form == TIS_ELE_COULOMB)
    oracle_force_coulomb(s, px, w->ele, w->ele_coef, w->nele, f);
  else if (s->p.ele_form == TIS_ELE_EWALD)
    oracle_force_ewald(s, px, w->ele, w->ele_coef, w->nele, f);
  return nclamp;
}

int32_t oracle_force_sumup(oracle_ws *w, const double *xyz, double *f,
                           uint64_t seed, uint32_t sys, uint64_t step) {
  philox_src p = {seed, sys, step, 0u};
  return force_sumup_core(w, xyz, f, philox_draw, &p);
}

int32_t oracle_force_sumup_mt(oracle_ws *w, const double *xyz, double *f,
                              const uint32_t *raw, int64_t nraw, int64_t *used) {
  mt_src m = {raw, nraw, 0, 0};
  int32_t nc = force_sumup_core(w, xyz, f, mt_draw, &m);
  if (used) *used = m.pos;
  return m.overrun ? -1 : nc;
}

/* energy_sumup.F90:149-529 for a DT15 system (hbond energies reuse the
 * force-time ene_hb, energy_dtrna_hbond15.F90:119-148). */
void oracle_energy_sumup(oracle_ws *w, const double *xyz, const double *velo, double *energy) {
  const tis_system *s = w->s;
  for (int k = 0; k < TIS_E_MAX; ++k) energy[k] = 0.0;
  if (velo) energy[TIS_E_VELO] = oracle_energy_kinetic(s, velo);
  energy[TIS_E_BOND] = oracle_energy_bond(s, xyz);
  energy[TIS_E_BANGLE] = oracle_energy_bangle(s, xyz);
  energy[TIS_E_TSTACK] = oracle_energy_tstack(s, xyz, w->pxyz);
  energy[TIS_E_STACK] = oracle_energy_stack(s, xyz, w->st_status);
  double ehb = 0.0, ethb = 0.0;
  for (int q = 0; q < w->nhbl; ++q) {
    int h = w->hb[q];
    if (!w->hb_status[h]) continue;
    if (s->hb_tertiary[h]) ethb += w->ene_hb[h]; else ehb += w->ene_hb[h];
    eu_add(s->hb_tertiary[h] ? TIS_E_THBOND : TIS_E_HBOND, s->hb_mp[6 * h], s->hb_mp[6 * h + 1], w->ene_hb[h]);
  }
  energy[TIS_E_HBOND] = ehb;
  energy[TIS_E_THBOND] = ethb;
  const double *px = s->p.periodic ? w->pxyz : xyz;
  energy[TIS_E_EXV] = oracle_energy_exv(s, px, w->exv, w->nexv);
  if (s->p.ele_form == TIS_ELE_DH)
    energy[TIS_E_ELE] = oracle_energy_dh(s, px, w->ele, w->ele_coef, w->nele);
  else if (s->p.ele_form == TIS_ELE_COULOMB)
    energy[TIS_E_ELE] = oracle_energy_coulomb(s, px, w->ele, w->ele_coef, w->nele);
  else if (s->p.ele_form == TIS_ELE_EWALD)
    energy[TIS_E_ELE] = oracle_energy_ewald(s, px, w->ele, w->ele_coef, w->nele, NULL);
  double tot = 0.0;
  for (int k = 0; k < TIS_E_MAX; ++k)
    if (k != TIS_E_TOTAL && k != TIS_E_VELO) tot += energy[k];
  energy[TIS_E_TOTAL] = tot;
}

/* energy_unit of energy_sumup.F90:76-507: the per-unit-pair sums of every
 * term (the Ewald terms add none: energy_ele_coulomb_ewld.F90:168-170 leaves
 * them out), then for style 0 each diagonal entry takes half of the upper
 * off-diagonal entries of its row and column (:480-492), for style 1 the
 * off-diagonal pairs are symmetrised (:494-503), and TOTAL is the sum over
 * the types other than TOTAL and VELO (:504-508). */
void oracle_energy_unit(oracle_ws *w, const double *xyz, const double *velo, int32_t style, double *eu) {
  const tis_system *s = w->s;
  const int nu = s->nunit;
  const size_t nn = (size_t)nu * nu;
  for (size_t k = 0; k < nn * TIS_E_MAX; ++k) eu[k] = 0.0;
  double energy[TIS_E_MAX];
  eu_acc = eu; eu_sys = s;
  oracle_energy_sumup(w, xyz, velo, energy);
  eu_acc = NULL; eu_sys = NULL;
#define EU(a, b, t) eu[(size_t)(a) + (size_t)nu * (b) + nn * (t)]
  if (style == 0) {
    for (int iu = 0; iu < nu; ++iu) {
      for (int ju = 0; ju < iu; ++ju)
        for (int t = 0; t < TIS_E_MAX; ++t) EU(iu, iu, t) = EU(iu, iu, t) + 0.5 * EU(ju, iu, t);
      for (int ju = iu + 1; ju < nu; ++ju)
        for (int t = 0; t < TIS_E_MAX; ++t) EU(iu, iu, t) = EU(iu, iu, t) + 0.5 * EU(iu, ju, t);
    }
  } else {
    for (int iu = 0; iu < nu; ++iu)
      for (int ju = iu + 1; ju < nu; ++ju)
        for (int t = 0; t < TIS_E_MAX; ++t) {
          double sume = EU(iu, ju, t) + EU(ju, iu, t);
          EU(iu, ju, t) = sume;
          EU(ju, iu, t) = sume;
        }
  }
  for (int t = 0; t < TIS_E_MAX; ++t) {
    if (t == TIS_E_TOTAL || t == TIS_E_VELO) continue;
    for (size_t k = 0; k < nn; ++k) eu[k + nn * TIS_E_TOTAL] += eu[k + nn * t];
  }
#undef EU
}

/* time_integral_pre.F90:160-170 */
#define BOLTZ_KCAL_MOL (1.38064852e-23 * ((1.0 / 4184.0) * 6.022140857e23))
void oracle_langevin_consts(const tis_system *s, double tstep, double *rlan) {
  for (int i = 0; i < s->nmp; ++i) {
    double fr = s->fric[i];
    double h = 0.5 * tstep * fr;
    double u1 = 1.0 - h, u2 = 1.0 - h + h * h;
    rlan[4 * i + 0] = sqrt(2.0 * fr * BOLTZ_KCAL_MOL * s->tempk / (tstep * s->mass[i]));
    rlan[4 * i + 1] = u1 * u2;
    rlan[4 * i + 2] = (0.5 * tstep) * u1;
    rlan[4 * i + 3] = tstep * u1;
  }
}

/* time_integral.F90:57-93 (trigger), :149-189 (Langevin), with the initial
 * acceleration of time_integral_pre.F90:176-179. */
int32_t oracle_langevin_run(oracle_ws *w, double tstep, int64_t nsteps, int32_t rebuild_every,
                            double *xyz, double *velo, double *accel,
                            uint64_t seed, uint32_t sys, uint64_t step0, int32_t init_accel) {
  const tis_system *s = w->s;
  int n = s->nmp;
  double *rlan = (double *)malloc(sizeof(double) * 4 * (size_t)n);
  double *f = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *dxyz = (double *)calloc(3 * (size_t)n, sizeof(double));
  oracle_langevin_consts(s, tstep, rlan);
  double tstep2 = 0.5 * tstep * tstep;
  int32_t nrebuild = 0;
  oracle_neighbor(w, xyz);
  ++nrebuild;
  if (init_accel) {
    oracle_force_sumup(w, xyz, f, seed, sys, step0);
    for (int i = 0; i < n; ++i) {
      double g[3];
      oracle_gauss3(seed, sys, step0, (uint32_t)i, g);
      for (int d = 0; d < 3; ++d)
        accel[3 * i + d] = f[3 * i + d] * (1.0 / s->mass[i]) + rlan[4 * i] * g[d];
    }
  }
  for (int64_t t = 0; t < nsteps; ++t) {
    uint64_t step = step0 + 1 + (uint64_t)t;
    int rebuild = 0;
    if (rebuild_every > 0) {
      rebuild = (t > 0) && (t % rebuild_every == 0);
    } else {
      double m1 = 0.0, m2 = 0.0;
      for (int i = 0; i < n; ++i) {
        double d2 = dot3(dxyz + 3 * i, dxyz + 3 * i);
        if (d2 > m1) { m2 = m1; m1 = d2; }
        else if (d2 > m2) m2 = d2;
      }
      rebuild = sqrt(m1) + sqrt(m2) > s->p.neigh_margin;
    }
    if (rebuild) {
      /* pxyz carries the last wrap; rebuild re-wraps the current xyz */
      oracle_neighbor(w, xyz);
      memset(dxyz, 0, sizeof(double) * 3 * (size_t)n);
      ++nrebuild;
    }
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        double dx = rlan[4 * i + 3] * velo[3 * i + d] + tstep2 * accel[3 * i + d];
        xyz[3 * i + d] += dx;
        w->pxyz[3 * i + d] += dx;
        dxyz[3 * i + d] += dx;
      }
    oracle_force_sumup(w, xyz, f, seed, sys, step);
    for (int i = 0; i < n; ++i) {
      double g[3];
      oracle_gauss3(seed, sys, step, (uint32_t)i, g);
      double rcm = 1.0 / s->mass[i];
      for (int d = 0; d < 3; ++d) {
        double af = f[3 * i + d] * rcm + rlan[4 * i] * g[d];
        velo[3 * i + d] = rlan[4 * i + 1] * velo[3 * i + d] + rlan[4 * i + 2] * (accel[3 * i + d] + af);
        accel[3 * i + d] = af;
      }
    }
  }
  free(rlan); free(f); free(dxyz);
  return nrebuild;
}

/* ------------------------------------------------------------------ */
/* Brownian dynamics with hydrodynamic interaction (Ermak-McCammon)    */

/* simu_hydro_tensors.F90:17-302: D (3n x 3n, column-major, FP64).
 * i_hydro_tensor (const_index.F90:39-45): 4 Zuk et al. RPY with overlap and
 * unequal radii (:50-134), 1 RPY (:139-182), 2 RPY with overlap (:187-241),
 * 3 Ermak with overlap (:246-302); types 1-3 use radius[0] for every bead.  c6 = kT/(6 pi eta),
 * c8 = kT/(8 pi eta). */
static void rpy_block(double *D, int n3, int i, int j, const double r[3], double pre1, double pre2) {
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double v = pre2 * r[a] * r[b] + (a == b ? pre1 : 0.0);
      D[(3 * i + a) + (size_t)(3 * j + b) * n3] = v;
      D[(3 * j + b) + (size_t)(3 * i + a) * n3] = v;
    }
}

void oracle_rpy_fill(int32_t n, const double *xyz, const double *radius, double tempk, double visc,
                     int32_t type, double *D) {
  const int n3 = 3 * n;
  const double c6 = BOLTZ_KCAL_MOL * tempk / (6.0 * F_PI * visc);
  const double c8 = BOLTZ_KCAL_MOL * tempk / (8.0 * F_PI * visc);
  memset(D, 0, sizeof(double) * (size_t)n3 * n3);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double r[3];
      for (int d = 0; d < 3; ++d) r[d] = xyz[3 * j + d] - xyz[3 * i + d];
      const double r_sq = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
      const double r_abs = sqrt(r_sq);
      double pre1, pre2;
      if (type == 4) {
        const double ri = radius[i], rj = radius[j];
        if (ri + rj < r_abs) {
          pre1 = c8 / r_abs * (1.0 + (ri * ri + rj * rj) / (3.0 * r_sq));
          pre2 = c8 / r_abs * (1.0 - (ri * ri + rj * rj) / r_sq) / r_sq;
        } else {
          const double rs = ri < rj ? ri : rj, rl = ri < rj ? rj : ri;
          if (rl - rs < r_abs) {
            const double r_cb = r_sq * r_abs, dd = (ri - rj) * (ri - rj);
            pre1 = c6 / ri / rj * (16.0 * r_cb * (ri + rj) - (dd + 3.0 * r_sq) * (dd + 3.0 * r_sq)) / 32.0 / r_cb;
            pre2 = c6 / ri / rj * 3.0 * (dd - r_sq) * (dd - r_sq) / 32.0 / r_cb / r_sq;
          } else {
            pre1 = c6 / rl;
            pre2 = 0.0;
          }
        }
      } else {
        const double a = radius[0], a2 = a * a;
        if (type == 1 || 2.0 * a < r_abs) {
          pre1 = c8 / r_abs * (1.0 + 2.0 * a2 / (3.0 * r_sq));
          pre2 = c8 / r_abs * (1.0 - 2.0 * a2 / r_sq) / r_sq;
        } else if (type == 2) {
          pre1 = c6 / a * (1.0 - 9.0 / 32.0 * r_abs / a);
          pre2 = c6 / a * 3.0 / 32.0 / a / r_abs;
        } else {
          pre1 = c8 / 2.0 / a * 7.0 / 6.0;
          pre2 = c8 / 2.0 / 2.0 / r_sq;
        }
      }
      rpy_block(D, n3, i, j, r, pre1, pre2);
    }
  for (int k = 0; k < n3; ++k) D[k + (size_t)k * n3] = c6 / (type == 4 ? radius[k / 3] : radius[0]);
}

/* simu_hydro_tensors.F90:304-324: left-looking Cholesky, L lower
 * (column-major, upper part zero).  Returns 0, or k+1 if the pivot of
 * row k is not positive. */
int32_t oracle_cholesky(int32_t n3, const double *D, double *L) {
  memset(L, 0, sizeof(double) * (size_t)n3 * n3);
  for (int i = 0; i < n3; ++i) {
    double x = D[i + (size_t)i * n3];
    for (int k = 0; k < i; ++k) x -= L[i + (size_t)k * n3] * L[i + (size_t)k * n3];
    if (x <= 0.0) return i + 1;
    x = sqrt(x);
    L[i + (size_t)i * n3] = x;
    for (int r = i + 1; r < n3; ++r) {
      double y = D[i + (size_t)r * n3];
      for (int k = 0; k < i; ++k) y -= L[r + (size_t)k * n3] * L[i + (size_t)k * n3];
      L[r + (size_t)i * n3] = y / x;
    }
  }
  return 0;
}

/* force_fene.F90:40-57: F = -k (r - r0) / (1 - (r - r0)^2 / R0^2) / r along v21 */
void oracle_force_fene(int32_t nfene, const int32_t *mp, const double *nat, const double *coef,
                       const double *dist2, const double *xyz, double *f) {
  for (int b = 0; b < nfene; ++b) {
    int i = mp[2 * b], j = mp[2 * b + 1];
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = xyz[3 * j + d] - xyz[3 * i + d];
    double dist = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    double dd = dist - nat[b];
    double fr = -coef[b] * dd / (1.0 - dd * dd / dist2[b]) / dist;
    for (int d = 0; d < 3; ++d) { f[3 * i + d] -= fr * v[d]; f[3 * j + d] += fr * v[d]; }
  }
}

/* force_wca.F90:64-96 (non-periodic): (s/r)^2 >= 1/cutoff^2 pairs;
 * coef rep (r < s) or att; dv = 12 coef / s^2 ((s/r)^14 - (s/r)^8), clamped at DE_MAX */
void oracle_force_wca(int32_t nwca, const int32_t *mp, const double *nat2, const double *coef,
                      double cutoff, const double *xyz, double *f) {
  const double rc2 = 1.0 / (cutoff * cutoff);
  for (int k = 0; k < nwca; ++k) {
    int i = mp[2 * k], j = mp[2 * k + 1];
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = xyz[3 * j + d] - xyz[3 * i + d];
    double d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    double ro2 = nat2[k] / d2;
    if (ro2 < rc2) continue;
    double ro14 = ro2 * ro2 * ro2 * ro2 * ro2 * ro2 * ro2, ro8 = ro2 * ro2 * ro2 * ro2;
    double c = ro2 > 1.0 ? coef[2 * k] : coef[2 * k + 1];
    double dv = 12.0 * c / nat2[k] * (ro14 - ro8);
    if (dv > DE_MAX) dv = DE_MAX;
    for (int d = 0; d < 3; ++d) { f[3 * i + d] -= dv * v[d]; f[3 * j + d] += dv * v[d]; }
  }
}

/* time_integral.F90:429-457 with time_integral_pre.F90:313-317: one
 * Ermak-McCammon step dx = (dt/kT) D F + sqrt(2 dt) L xi, xi ~ N(0,1)
 * from the counter-based stream of (seed, key, step, bead). */
void oracle_bdhi_step(int32_t n, const double *D, const double *L, const double *F, double tempk,
                      double tstep, uint64_t seed, uint32_t key, uint64_t step, double *xyz) {
  const int n3 = 3 * n;
  const double c1 = tstep / (BOLTZ_KCAL_MOL * tempk), c2 = sqrt(2.0 * tstep);
  double *xi = (double *)malloc(sizeof(double) * (size_t)n3);
  double *dx = (double *)malloc(sizeof(double) * (size_t)n3);
  for (int i = 0; i < n; ++i) oracle_gauss3(seed, key, step, (uint32_t)i, xi + 3 * i);
  for (int k = 0; k < n3; ++k) {
    double a = 0.0, b = 0.0;
    for (int l = 0; l < n3; ++l) a += D[l + (size_t)k * n3] * F[l];
    for (int l = 0; l <= k; ++l) b += L[k + (size_t)l * n3] * xi[l];
    dx[k] = c1 * a + c2 * b;
  }
  for (int k = 0; k < n3; ++k) xyz[k] += dx[k];
  free(xi); free(dx);
}
