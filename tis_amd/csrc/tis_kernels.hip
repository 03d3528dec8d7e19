// tis_kernels.hip -- CDNA4 (gfx950) kernels of the DT15 TIS force-and-
// integrate path.  One wave (64 lanes) per work tile; a tile is 64 beads (or
// 64 charges) of one system.  Each lane owns one bead and gathers every force
// on it in a fixed order (bonded terms via the incidence list, excluded
// volume via its own neighbour list, Debye-Hueckel via the per-charge sums of
// the pre-pass), accumulating in FP64.  Pair arithmetic (mWCA, DH) is FP32 on
// FP64 coordinate differences; cutoff and list tests are FP64 with the
// reference's rounding.  Reference routines restated (src/ of the reference):
//   force_bond.F90:40-58, force_bangle.F90:40-84, force_dtrna_stack.F90:53-165,
//   force_dtrna_stack_nlocal.F90:48-263, force_dtrna_hbond15.F90:91-535,
//   force_exv_dt15.F90:53-115, force_ele_DH.F90:68-118,
//   neighbor_list.F90:56-87 + neighbor_assign.F90:481-559, neighbor_list_ele.F90:123-210,
//   neighbor_list_hb.F90:59-78, util_periodic.F90:26-36, time_integral.F90:149-183,
//   time_integral_pre.F90:176-179, energy_*.F90.
#include "tis_dev.h"

namespace tis {

__device__ inline void ld3(const double *X, int n, int g, double r[3]) {
  r[0] = X[g]; r[1] = X[n + g]; r[2] = X[2 * n + g];
}
__device__ inline double dot3(const double a[3], const double b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__device__ inline void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double wrap_pi(double d) {
  if (d > kPi) return d - 2.0 * kPi;
  if (d < -kPi) return d + 2.0 * kPi;
  return d;
}
__device__ inline void vsub(const double *X, int n, int a, int b, double v[3]) {
  v[0] = X[a] - X[b]; v[1] = X[n + a] - X[n + b]; v[2] = X[2 * n + a] - X[2 * n + b];
}

// ------------------------------------------------------------------ bonded
// force on slot `slot` of bond b (force_bond.F90:40-58)
__device__ inline void bond_slot(const DevPtrs &P, const double *X, int b, int slot, double acc[3]) {
  int i = P.bd_mp[2 * b], j = P.bd_mp[2 * b + 1];
  double v[3];
  vsub(X, P.ntot, j, i, v);
  double dist = sqrt(dot3(v, v));
  double dd = dist - P.bd_nat[b];
  double fr = P.bd_coef[b] * (-2.0 * dd / dist);
  double sg = slot == 0 ? -1.0 : 1.0;
  acc[0] += sg * (fr * v[0]); acc[1] += sg * (fr * v[1]); acc[2] += sg * (fr * v[2]);
}

// force_bangle.F90:40-84
__device__ inline void angle_slot(const DevPtrs &P, const double *X, int a, int slot, double acc[3]) {
  double coef = P.ba_coef[a];
  if (coef < 1.0e-6) return;
  int i1 = P.ba_mp[3 * a], i2 = P.ba_mp[3 * a + 1], i3 = P.ba_mp[3 * a + 2];
  double v21[3], v32[3];
  vsub(X, P.ntot, i2, i1, v21);
  vsub(X, P.ntot, i3, i2, v32);
  double c11 = dot3(v21, v21), c22 = dot3(v32, v32), c21 = dot3(v32, v21);
  double ct = -c21 / sqrt(c11 * c22);
  ct = ct > 1.0 ? 1.0 : (ct < -1.0 ? -1.0 : ct);
  double dba = acos(ct) - P.ba_nat[a];
  double t3 = c11 * c22 - c21 * c21;
  if (t3 <= 1.0) t3 = 1.0;
  double fr = 2.0 * coef * dba / sqrt(t3);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double f21 = fr * (v21[d] * (c21 / c11) - v32[d]);
    double f32 = fr * (v32[d] * (c21 / c22) - v21[d]);
    acc[d] += slot == 0 ? -f21 : (slot == 1 ? f21 - f32 : f32);
  }
}

// force_dtrna_stack.F90:53-165, force on one of its seven beads
__device__ inline void stack_slot(const DevPtrs &P, const double *X, int st, int slot, double acc[3]) {
  if (!P.st_status[st]) return;
  const int n = P.ntot;
  const int32_t *m7 = P.st_mp + 7 * st;
  const double *nat = P.st_nat + 3 * st, *cf = P.st_coef + 4 * st;
  double v21[3], v34[3], v54[3], v56[3], v76[3], m[3], nn[3], fi[3], fl[3], out[3] = {0, 0, 0};
  vsub(X, n, m7[1], m7[0], v21);
  vsub(X, n, m7[2], m7[3], v34);
  vsub(X, n, m7[4], m7[3], v54);
  vsub(X, n, m7[4], m7[5], v56);
  vsub(X, n, m7[6], m7[5], v76);
  double ediv = 1.0;
  double dist = sqrt(dot3(v21, v21));
  double dd = dist - nat[0];
  ediv += cf[1] * dd * dd;
  if (slot <= 1) {
    double sg = slot == 0 ? -1.0 : 1.0;
    for (int d = 0; d < 3; ++d) out[d] = sg * (2.0 * cf[1] * dd / dist * v21[d]);
  }
  cross3(v34, v54, m);
  cross3(v54, v56, nn);
  double dmm = dot3(m, m), dnn = dot3(nn, nn), d5454 = dot3(v54, v54);
  double abs54 = sqrt(d5454), d5654 = dot3(v56, v54);
  double d3454o = dot3(v34, v54) / d5454, d5654o = d5654 / d5454;
  double dph = wrap_pi(atan2(dot3(v34, nn) * sqrt(d5454), dot3(m, nn)) - nat[1]);
  ediv += cf[2] * dph * dph;
  for (int d = 0; d < 3; ++d) {
    fi[d] = 2.0 * cf[2] * dph * abs54 / dmm * m[d];
    fl[d] = -2.0 * cf[2] * dph * abs54 / dnn * nn[d];
  }
  if (slot == 2) for (int d = 0; d < 3; ++d) out[d] += fi[d];
  if (slot == 3) for (int d = 0; d < 3; ++d) out[d] += (-1.0 + d3454o) * fi[d] - d5654o * fl[d];
  if (slot == 4) for (int d = 0; d < 3; ++d) out[d] += (-1.0 + d5654o) * fl[d] - d3454o * fi[d];
  if (slot == 5) for (int d = 0; d < 3; ++d) out[d] += fl[d];
  cross3(v76, v56, m);
  for (int d = 0; d < 3; ++d) nn[d] = -nn[d];
  dmm = dot3(m, m);
  double d5656 = dot3(v56, v56), abs56 = sqrt(d5656);
  double d7656o = dot3(v76, v56) / d5656, d5456o = d5654 / d5656;
  dph = wrap_pi(atan2(dot3(v76, nn) * sqrt(d5656), dot3(m, nn)) - nat[2]);
  ediv += cf[3] * dph * dph;
  for (int d = 0; d < 3; ++d) {
    fi[d] = 2.0 * cf[3] * dph * abs56 / dmm * m[d];
    fl[d] = -2.0 * cf[3] * dph * abs56 / dnn * nn[d];
  }
  if (slot == 6) for (int d = 0; d < 3; ++d) out[d] += fi[d];
  if (slot == 5) for (int d = 0; d < 3; ++d) out[d] += (-1.0 + d7656o) * fi[d] - d5456o * fl[d];
  if (slot == 4) for (int d = 0; d < 3; ++d) out[d] += -d7656o * fi[d] + (-1.0 + d5456o) * fl[d];
  if (slot == 3) for (int d = 0; d < 3; ++d) out[d] += fl[d];
  double sc = cf[0] / (ediv * ediv);
  for (int d = 0; d < 3; ++d) acc[d] += sc * out[d];
}

// ------------------------------------------------------------------ tertiary stack
// force_dtrna_stack_nlocal.F90:48-263 for one tertiary stack: all six slots
// into fo[18]; sets the exclusive secondary stacks off when d < 10.
__device__ void tstack_eval(const DevPtrs &P, const DSys &s, const double *X, const double *PX,
                            int t, double fo[18]) {
  const int n = P.ntot;
  const int32_t *m6 = P.tst_mp + 6 * t;
  const double *nat = P.tst_nat + 6 * t, *cf = P.tst_coef + 7 * t;
  const double *Y = s.periodic ? PX : X;
  double v12[3], v13[3], v53[3], v42[3], v46[3], fi[3], fk[3], fl[3];
  vsub(Y, n, m6[0], m6[1], v12);
  if (s.periodic) pbneighbor(s, v12);
  double d1212 = dot3(v12, v12), a12 = sqrt(d1212);
  double dd = a12 - nat[0];
  if (dd < 10.0) {
    if (P.tst_excl[2 * t] && P.tst2st[2 * t] >= 0) P.st_status[P.tst2st[2 * t]] = 0;
    if (P.tst_excl[2 * t + 1] && P.tst2st[2 * t + 1] >= 0) P.st_status[P.tst2st[2 * t + 1]] = 0;
  }
  vsub(Y, n, m6[0], m6[2], v13);
  vsub(Y, n, m6[4], m6[2], v53);
  vsub(Y, n, m6[3], m6[1], v42);
  vsub(Y, n, m6[3], m6[5], v46);
  if (s.periodic) { pbneighbor(s, v13); pbneighbor(s, v53); pbneighbor(s, v42); pbneighbor(s, v46); }
  for (int k = 0; k < 18; ++k) fo[k] = 0.0;
  double d1313 = dot3(v13, v13), d4242 = dot3(v42, v42);
  double a13 = sqrt(d1313), a42 = sqrt(d4242);
  double d1213 = dot3(v13, v12), d1242 = dot3(v12, v42), d4246 = dot3(v42, v46), d1353 = dot3(v13, v53);
  double o1213_1212 = d1213 / d1212, o1213_1313 = d1213 / d1313, o1242_1212 = d1242 / d1212;
  double o4246_4242 = d4246 / d4242, o1242_4242 = d1242 / d4242, o1353_1313 = d1353 / d1313;
  double ediv = 1.0 + cf[1] * dd * dd;
  for (int d = 0; d < 3; ++d) {
    fi[d] = 2.0 * cf[1] * dd * v12[d] / a12;
    fo[d] = fi[d]; fo[3 + d] = -fi[d];
  }
  double da = acos(d1213 / (a13 * a12)) - nat[1];
  double pre = 2.0 * cf[2] * da / sqrt(d1313 * d1212 - d1213 * d1213);
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre * (v12[d] - o1213_1313 * v13[d]);
    fk[d] = pre * (v13[d] - o1213_1212 * v12[d]);
    fo[6 + d] += fi[d]; fo[3 + d] += fk[d]; fo[d] = fo[d] - fi[d] - fk[d];
  }
  ediv += cf[2] * da * da;
  da = acos(d1242 / (a12 * a42)) - nat[2];
  pre = 2.0 * cf[3] * da / sqrt(d1212 * d4242 - d1242 * d1242);
  for (int d = 0; d < 3; ++d) {
    fi[d] = -pre * (v42[d] - o1242_1212 * v12[d]);
    fk[d] = -pre * (v12[d] - o1242_4242 * v42[d]);
    fo[d] += fi[d]; fo[9 + d] += fk[d]; fo[3 + d] = fo[3 + d] - fi[d] - fk[d];
  }
  ediv += cf[3] * da * da;
  double c4212[3], c1213[3], m[3], nn[3];
  cross3(v42, v12, c4212);
  cross3(v12, v13, c1213);
  double c4212a = dot3(c4212, c4212), c1213a = dot3(c1213, c1213);
  double dph = wrap_pi(atan2(dot3(v42, c1213) * sqrt(d1212), dot3(c4212, c1213)) - nat[3]);
  ediv += cf[4] * dph * dph;
  pre = 2.0 * cf[4] * dph * a12;
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre / c4212a * c4212[d];
    fl[d] = -pre / c1213a * c1213[d];
    fo[9 + d] += fi[d];
    fo[3 + d] += (-1.0 + o1242_1212) * fi[d] - o1213_1212 * fl[d];
    fo[d] += (-1.0 + o1213_1212) * fl[d] - o1242_1212 * fi[d];
    fo[6 + d] += fl[d];
  }
  cross3(v53, v13, m);
  for (int d = 0; d < 3; ++d) nn[d] = -c1213[d];
  double dmm = dot3(m, m), dnn = c1213a;
  dph = wrap_pi(atan2(dot3(v53, nn) * a13, dot3(m, nn)) - nat[4]);
  ediv += cf[5] * dph * dph;
  pre = 2.0 * cf[5] * dph * a13;
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre / dmm * m[d];
    fl[d] = -pre / dnn * nn[d];
    fo[12 + d] += fi[d];
    fo[6 + d] += (-1.0 + o1353_1313) * fi[d] - o1213_1313 * fl[d];
    fo[d] += (-1.0 + o1213_1313) * fl[d] - o1353_1313 * fi[d];
    fo[3 + d] += fl[d];
  }
  for (int d = 0; d < 3; ++d) m[d] = -c4212[d];
  cross3(v42, v46, nn);
  dmm = c4212a; dnn = dot3(nn, nn);
  dph = wrap_pi(atan2(dot3(v12, nn) * a42, dot3(m, nn)) - nat[5]);
  ediv += cf[6] * dph * dph;
  pre = 2.0 * cf[6] * dph * a42;
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre / dmm * m[d];
    fl[d] = -pre / dnn * nn[d];
    fo[d] += fi[d];
    fo[3 + d] += (-1.0 + o1242_4242) * fi[d] - o4246_4242 * fl[d];
    fo[9 + d] += (-1.0 + o4246_4242) * fl[d] - o1242_4242 * fi[d];
    fo[15 + d] += fl[d];
  }
  double sc = cf[0] / (ediv * ediv);
  for (int k = 0; k < 18; ++k) fo[k] *= sc;
}

// ------------------------------------------------------------------ hbond15
// force_dtrna_hbond15.F90:91-353 (default branch).  Returns 1 and fills
// fo[18], *ene when |r12 - r0| <= hb_cutoff_dist.
__device__ int hbond_eval(const DevPtrs &P, const DSys &s, const double *X, const double *PX,
                          int h, double fo[18], double *ene) {
  const int n = P.ntot;
  const double MAXC = cos((double)(0.57f / 180.0f) * kPi);  // const_physical.F90:37-38
  const int32_t *m6 = P.hb_mp + 6 * h;
  const double *nat = P.hb_nat + 6 * h, *cf = P.hb_coef + 7 * h;
  double v12[3], v13[3], v53[3], v42[3], v46[3], fi[3], fk[3], fl[3];
  if (s.periodic) { vsub(PX, n, m6[0], m6[1], v12); pbneighbor(s, v12); }
  else vsub(X, n, m6[0], m6[1], v12);
  double d1212 = dot3(v12, v12), a12 = sqrt(d1212);
  double dd = a12 - nat[0];
  if (fabs(dd) > s.hb_cut) return 0;
  double ex = -cf[1] * dd * dd;
  for (int d = 0; d < 3; ++d) {
    fi[d] = (2.0 * cf[1] * dd / a12) * v12[d];
    fo[d] = fi[d]; fo[3 + d] = -fi[d];
  }
  for (int k = 6; k < 18; ++k) fo[k] = 0.0;
  vsub(X, n, m6[0], m6[2], v13);
  vsub(X, n, m6[4], m6[2], v53);
  vsub(X, n, m6[3], m6[1], v42);
  vsub(X, n, m6[3], m6[5], v46);
  if (s.periodic) { pbneighbor(s, v13); pbneighbor(s, v53); pbneighbor(s, v42); pbneighbor(s, v46); }
  double d1313 = dot3(v13, v13), d4242 = dot3(v42, v42);
  double a13 = sqrt(d1313), a42 = sqrt(d4242);
  double d1213 = dot3(v13, v12), d1242 = dot3(v12, v42), d4246 = dot3(v42, v46), d1353 = dot3(v13, v53);
  double o1213_1212 = d1213 / d1212, o1213_1313 = d1213 / d1313, o1242_1212 = d1242 / d1212;
  double o4246_4242 = d4246 / d4242, o1242_4242 = d1242 / d4242, o1353_1313 = d1353 / d1313;
  double c312 = d1213 / (a13 * a12);
  if (fabs(c312) > MAXC) d1213 = copysign(a12 * a13 * MAXC, c312);
  double da = acos(c312) - nat[1];
  double pre = 2.0 * cf[2] * da / sqrt(d1313 * d1212 - d1213 * d1213);
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre * (v12[d] - o1213_1313 * v13[d]);
    fk[d] = pre * (v13[d] - o1213_1212 * v12[d]);
    fo[6 + d] += fi[d]; fo[3 + d] += fk[d]; fo[d] = fo[d] - fi[d] - fk[d];
  }
  ex = ex - cf[2] * da * da;
  double c124 = d1242 / (a12 * a42);
  if (fabs(c124) > MAXC) d1242 = copysign(a12 * a42 * MAXC, c124);
  da = acos(c124) - nat[2];
  pre = 2.0 * cf[3] * da / sqrt(d1212 * d4242 - d1242 * d1242);
  for (int d = 0; d < 3; ++d) {
    fi[d] = -pre * (v42[d] - o1242_1212 * v12[d]);
    fk[d] = -pre * (v12[d] - o1242_4242 * v42[d]);
    fo[d] += fi[d]; fo[9 + d] += fk[d]; fo[3 + d] = fo[3 + d] - fi[d] - fk[d];
  }
  ex = ex - cf[3] * da * da;
  double c4212[3], c1213[3], c5313[3], c4246[3], nn[3];
  cross3(v42, v12, c4212);
  cross3(v12, v13, c1213);
  double c4212a = dot3(c4212, c4212), c1213a = dot3(c1213, c1213);
  if (fabs(c124) > MAXC) c4212a = d4242 * d1212 * (1.0 - MAXC * MAXC);
  if (fabs(c312) > MAXC) c1213a = d1212 * d1313 * (1.0 - MAXC * MAXC);
  double dph = wrap_pi(atan2(dot3(v42, c1213) * a12, dot3(c4212, c1213)) - nat[3]);
  ex = ex - cf[4] * dph * dph;
  pre = 2.0 * cf[4] * dph * a12;
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre / c4212a * c4212[d];
    fl[d] = -pre / c1213a * c1213[d];
    fo[9 + d] += fi[d];
    fo[3 + d] += (-1.0 + o1242_1212) * fi[d] - o1213_1212 * fl[d];
    fo[d] += (-1.0 + o1213_1212) * fl[d] - o1242_1212 * fi[d];
    fo[6 + d] += fl[d];
  }
  cross3(v53, v13, c5313);
  for (int d = 0; d < 3; ++d) nn[d] = -c1213[d];
  double c5313a = dot3(c5313, c5313);
  double c531 = d1353 / (a13 * sqrt(dot3(v53, v53)));
  if (fabs(c531) > MAXC) c5313a = dot3(v53, v53) * d1313 * (1.0 - MAXC * MAXC);
  dph = wrap_pi(atan2(dot3(v53, nn) * a13, dot3(c5313, nn)) - nat[4]);
  ex = ex - cf[5] * dph * dph;
  pre = 2.0 * cf[5] * dph * a13;
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre / c5313a * c5313[d];
    fl[d] = -pre / c1213a * nn[d];
    fo[12 + d] += fi[d];
    fo[6 + d] += (-1.0 + o1353_1313) * fi[d] - o1213_1313 * fl[d];
    fo[d] += (-1.0 + o1213_1313) * fl[d] - o1353_1313 * fi[d];
    fo[3 + d] += fl[d];
  }
  for (int d = 0; d < 3; ++d) nn[d] = -c4212[d];
  cross3(v42, v46, c4246);
  double c4246a = dot3(c4246, c4246);
  double c246 = d4246 / (a42 * sqrt(dot3(v46, v46)));
  if (fabs(c246) > MAXC) c4246a = dot3(v46, v46) * d4242 * (1.0 - MAXC * MAXC);
  dph = wrap_pi(atan2(dot3(v12, c4246) * a42, dot3(nn, c4246)) - nat[5]);
  ex = ex - cf[6] * dph * dph;
  pre = 2.0 * cf[6] * dph * a42;
  for (int d = 0; d < 3; ++d) {
    fi[d] = pre / c4212a * nn[d];
    fl[d] = -pre / c4246a * c4246[d];
    fo[d] += fi[d];
    fo[3 + d] += (-1.0 + o1242_4242) * fi[d] - o4246_4242 * fl[d];
    fo[9 + d] += (-1.0 + o4246_4242) * fl[d] - o1242_4242 * fi[d];
    fo[15 + d] += fl[d];
  }
  double e = cf[0] * exp(ex);
  *ene = e;
  for (int k = 0; k < 18; ++k) fo[k] = e * fo[k];
  return 1;
}

__device__ inline int site_count(const DevPtrs &P, int site) {
  return __hip_atomic_load(&P.site_cnt[site], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Valence-limit deletion of force_dtrna_hbond15.F90:373-510 for one system,
// run by one lane; the random draws come from the counter-based generator so
// that the result equals the oracle's bit for bit.
__device__ int hbond_valence_serial(const DevPtrs &P, const DSys &s, uint64_t step) {
  int ndel = 0;
  uint32_t draw = 0;
  for (int iter = 0; iter < (1 << 20); ++iter) {
    int nex = 0;
    for (int k = 0; k < s.nsite; ++k)
      if (site_count(P, s.sitoff + k) - P.site_val[s.sitoff + k] > 0) ++nex;
    if (nex == 0) break;
    double rnd = uniform01(P.seed, s.key, step, 0, draw++);
    int pick = (int)ceil(rnd * nex) - 1, site = -1;
    for (int k = 0, c = 0; k < s.nsite; ++k)
      if (site_count(P, s.sitoff + k) - P.site_val[s.sitoff + k] > 0) {
        if (c == pick) { site = s.sitoff + k; break; }
        ++c;
      }
    // hbonds of the site in list order, shuffled in place (at most 64 kept)
    int seq[64];
    int nseq = 0;
    for (int q = 0; q < s.nhb && nseq < 64; ++q) {
      int h = s.hoff + q;
      if (!P.hb_inlist[h] || !P.hb_status[h]) continue;
      for (int k = 0; k < 6; ++k)
        if (P.hb_site[6 * h + k] == site) { seq[nseq++] = h; break; }
    }
    if (nseq == 0) break;
    for (int i = 0; i < nseq; ++i) {
      rnd = uniform01(P.seed, s.key, step, 0, draw++);
      int isw = (int)ceil(rnd * nseq) - 1;
      int t = seq[i]; seq[i] = seq[isw]; seq[isw] = t;
    }
    int del = seq[0];
    for (int i = 1; i < nseq; ++i) {
      int j = seq[i];
      double ratio = exp((P.hb_ene[j] - P.hb_ene[del]) * s.beta_hb);
      rnd = uniform01(P.seed, s.key, step, 0, draw++);
      if (rnd < ratio) del = j;
    }
    P.hb_status[del] = 0;
    P.hb_ene[del] = 0.0;
    ++ndel;
    // excess -= 1 where positive; count kept as valence + excess
    for (int k = 0; k < 6; ++k) {
      int st = P.hb_site[6 * del + k];
      if (st >= 0 && site_count(P, st) - P.site_val[st] > 0) atomicSub(&P.site_cnt[st], 1);
    }
  }
  return ndel;
}

// ------------------------------------------------------------------ pre-pass
// Work items: [0, nsystask) = one system each (tertiary stacks, hbonds and
// their valence pass); [nsystask, nsystask+nqtile) = DH charge tiles.
__global__ __launch_bounds__(kTile) void k_pre(DevPtrs P, const double *X, const double *PX,
                                               const int32_t *systask, int nsystask,
                                               const Tile *qtiles, int nqtile, int32_t t_local,
                                               int do_dh) {
  const int lane = threadIdx.x;
  const int w = blockIdx.x;
  const uint64_t step = *P.stepbase + (uint64_t)t_local + 1u;
  if (w < nsystask) {
    const DSys s = P.sys[systask[w]];
    // tertiary stacks
    for (int k = lane; k < s.nst; k += kTile) P.st_status[s.soff + k] = 1;
    __syncthreads();
    for (int k = lane; k < s.ntst; k += kTile) {
      double fo[18];
      tstack_eval(P, s, X, PX, s.toff + k, fo);
      for (int q = 0; q < 18; ++q) P.for_tst[(size_t)q * P.ntst + s.toff + k] = fo[q];
    }
    // hbond candidates in the list
    if (s.hb_excess_possible)
      for (int k = lane; k < s.nsite; k += kTile) P.site_cnt[s.sitoff + k] = 0;
    __syncthreads();
    for (int k = lane; k < s.nhb; k += kTile) {
      int h = s.hoff + k;
      int st = 0;
      double e = 0.0, fo[18];
      if (P.hb_inlist[h]) st = hbond_eval(P, s, X, PX, h, fo, &e);
      P.hb_status[h] = (uint8_t)st;
      P.hb_ene[h] = e;
      if (st) {
        for (int q = 0; q < 18; ++q) P.for_hb[(size_t)q * P.nhb + h] = fo[q];
        if (s.hb_excess_possible)
          for (int q = 0; q < 6; ++q) {
            int site = P.hb_site[6 * h + q];
            if (site >= 0) atomicAdd(&P.site_cnt[site], 1);
          }
      }
    }
    __syncthreads();
    if (s.hb_excess_possible) {
      int any = 0;
      for (int k = lane; k < s.nsite; k += kTile)
        if (__hip_atomic_load(&P.site_cnt[s.sitoff + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                P.site_val[s.sitoff + k] > 0)
          any = 1;
      any = __syncthreads_or(any);
      if (any && lane == 0) {
        int nd = hbond_valence_serial(P, s, step);
        if (nd) atomicAdd(&P.counters[2], (unsigned long long)nd);
      }
    }
    return;
  }
  if (!do_dh) return;
  // ---- Debye-Hueckel charge tile (force_ele_DH.F90:68-118), owner-computes
  __shared__ double sx[kTile], sy[kTile], sz[kTile], sq[kTile];
  const Tile tl = qtiles[w - nsystask];
  const DSys s = P.sys[tl.sys];
  if (s.ele_form != 0) return;
  const int qi = tl.start + lane;
  const bool valid = qi < s.nq;
  const double *Y = s.periodic ? PX : X;
  const int n = P.ntot;
  double xi[3] = {0, 0, 0}, chi = 0.0;
  if (valid) {
    int gi = P.qbead[s.qoff + qi];
    ld3(Y, n, gi, xi);
    chi = P.charge[gi];
  }
  double F[3] = {0.0, 0.0, 0.0};
  const float rcd = (float)s.rcdist;
  unsigned nclamp = 0;
  const uint32_t *row = P.qmask + s.moff + (int64_t)(valid ? qi : 0) * s.mwords;
  for (int jt = 0; jt < s.nq; jt += kTile) {
    int qj = jt + lane;
    if (qj < s.nq) {
      int gj = P.qbead[s.qoff + qj];
      sx[lane] = Y[gj]; sy[lane] = Y[n + gj]; sz[lane] = Y[2 * n + gj];
      sq[lane] = P.charge[gj];
    }
    __syncthreads();
    if (valid) {
      const int wbase = jt >> 5;
      uint32_t w0 = row[wbase];
      uint32_t w1 = (wbase + 1 < s.mwords) ? row[wbase + 1] : 0u;
      uint64_t bits = (uint64_t)w0 | ((uint64_t)w1 << 32);
      while (bits) {
        int k = __ffsll((unsigned long long)bits) - 1;
        bits &= bits - 1;
        double v0 = sx[k] - xi[0], v1 = sy[k] - xi[1], v2 = sz[k] - xi[2];
        if (s.periodic) {
          int code = P.qmirror[s.mioff + (int64_t)qi * s.nq + jt + k];
          v0 += mirror_shift(s, code, 0); v1 += mirror_shift(s, code, 1); v2 += mirror_shift(s, code, 2);
        }
        double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (d2 > s.dh_cut2) continue;
        float coef = (float)((chi * sq[k]) * s.ele_coef);
        float f0 = (float)v0, f1 = (float)v1, f2 = (float)v2;
        float d1 = sqrtf((float)d2);
        float rd = 1.0f / d1;
        float dv = coef * rd * rd * (rd + rcd) * __expf(-d1 * rcd);
        if (dv > (float)kDeMax) { dv = (float)kDeMax; if (qi < jt + k) ++nclamp; }
        F[0] -= (double)(dv * f0); F[1] -= (double)(dv * f1); F[2] -= (double)(dv * f2);
      }
    }
    __syncthreads();
  }
  if (valid) {
    const int q = s.qoff + qi;
    P.fdh[q] = F[0];
    P.fdh[(size_t)P.nqtot + q] = F[1];
    P.fdh[2 * (size_t)P.nqtot + q] = F[2];
  }
  if (nclamp) atomicAdd(&P.counters[1], (unsigned long long)nclamp);
}


// ------------------------------------------------------------------ main pass
// mode 0: store F; mode 1: Langevin step; mode 2: initial acceleration;
// mode 3: store F and predict positions from the stored v, a.
// Langevin (time_integral.F90:169-183, then the next step's position update
// :152-160): a' = F/m + c1 xi; v = c2 v + c3 (a + a'); a = a';
// x_next = x + c4 v + (dt^2/2) a.
__device__ inline void integrate(const DevPtrs &P, const DSys &s, int g, int li, const double F[3],
                                 const double *X, const double *PX, double *XN, double *PXN,
                                 uint64_t step, int mode) {
#pragma clang fp contract(off)
  const int n = P.ntot;
  double gs[3];
  gauss3(P.seed, s.key, step, (uint32_t)li, gs);
  const double r0 = P.rlan[g], r1 = P.rlan[n + g], r2 = P.rlan[2 * n + g], r3 = P.rlan[3 * n + g];
  const double rcm = P.rcm[g];
  for (int d = 0; d < 3; ++d) {
    double af = F[d] * rcm + r0 * gs[d];
    double vv = P.v[d * n + g];
    double a0 = P.a[d * n + g];
    if (mode == 3) af = a0;
    if (mode == 1) vv = r1 * vv + r2 * (a0 + af);
    P.v[d * n + g] = vv;
    P.a[d * n + g] = af;
    double dx = r3 * vv + P.tstep2 * af;
    XN[d * n + g] = X[d * n + g] + dx;
    if (s.periodic) PXN[d * n + g] = PX[d * n + g] + dx;
  }
}

__global__ __launch_bounds__(kTile) void k_main(DevPtrs P, const double *X, const double *PX,
                                                double *XN, double *PXN, const Tile *tiles,
                                                int ntile, int32_t t_local, int mode, int do_dh) {
  const int lane = threadIdx.x;
  const Tile tl = tiles[blockIdx.x];
  const DSys s = P.sys[tl.sys];
  const int li = tl.start + lane;
  if (li >= s.nmp) return;
  const int g = s.off + li;
  const int n = P.ntot;
  const uint64_t step = *P.stepbase + (uint64_t)t_local + (mode == 2 ? 0u : 1u);
  double F[3] = {0.0, 0.0, 0.0};
  // bonded terms in incidence order
  const int ib = P.inc_off[g], ie = P.inc_off[g + 1];
  for (int k = ib; k < ie; ++k) {
    uint32_t e = P.inc[k];
    uint32_t type = e >> 29, slot = (e >> 26) & 7u, term = e & ((1u << 26) - 1u);
    switch (type) {
      case INC_BOND: bond_slot(P, X, (int)term, (int)slot, F); break;
      case INC_ANGLE: angle_slot(P, X, (int)term, (int)slot, F); break;
      case INC_STACK: stack_slot(P, X, (int)term, (int)slot, F); break;
      case INC_TSTACK:
        for (int d = 0; d < 3; ++d) F[d] += P.for_tst[(size_t)(slot * 3 + d) * P.ntst + term];
        break;
      default:
        if (P.hb_status[term])
          for (int d = 0; d < 3; ++d) F[d] += P.for_hb[(size_t)(slot * 3 + d) * P.nhb + term];
        break;
    }
  }
  // excluded volume (force_exv_dt15.F90:53-115), owner view of each pair
  {
    const double *Y = s.periodic ? PX : X;
    double xi[3];
    ld3(Y, n, g, xi);
    const int ti = P.mptype[g], ci = P.iclass[g];
    const double ri = P.exv_rad[g];
    const float epi = P.exv_eps[g];
    const float a = (float)s.exv_a, a2 = a * a;
    const int cnt = min(P.nb_cnt[g], P.nb_cap);
    unsigned nclamp = 0;
    for (int m = 0; m < cnt; ++m) {
      uint32_t e = P.nb[(size_t)m * n + g];
      int j = (int)(e & kIdxMask);
      double v0 = Y[j] - xi[0], v1 = Y[n + j] - xi[1], v2 = Y[2 * n + j] - xi[2];
      if (s.periodic) {
        int code = (int)(e >> kMirrorShift);
        v0 += mirror_shift(s, code, 0); v1 += mirror_shift(s, code, 1); v2 += mirror_shift(s, code, 2);
      }
      const int tj = P.mptype[j];
      double dij;
      if ((ti == 2 && tj == 4) || (ti == 4 && tj == 2)) dij = s.exv_dist_PS;
      else if (ci == 2 && P.iclass[j] == 2) dij = s.exv_dist;
      else dij = ri + P.exv_rad[j];
      double d2 = v0 * v0 + v1 * v1 + v2 * v2;
      if (d2 > dij * dij) continue;
      float dist = sqrtf((float)d2);
      float coef = epi * P.exv_eps[j];
      float dr = dist + a - (float)dij;
      float r2 = a2 / (dr * dr), r4 = r2 * r2, r8 = r4 * r4, r14 = r2 * r4 * r8;
      float dv = fabsf(12.0f * coef * (r14 - r8) * dr / a2 / dist);
      if (dv > (float)kDeMax) { dv = (float)kDeMax; if (g < j) ++nclamp; }
      F[0] -= (double)(dv * (float)v0); F[1] -= (double)(dv * (float)v1); F[2] -= (double)(dv * (float)v2);
    }
    if (nclamp) atomicAdd(&P.counters[0], (unsigned long long)nclamp);
  }
  // Debye-Hueckel sums of the pre-pass
  if (do_dh && s.ele_form == 0) {
    const int q = P.qidx[g];
    if (q >= 0) { F[0] += P.fdh[q]; F[1] += P.fdh[(size_t)P.nqtot + q]; F[2] += P.fdh[2 * (size_t)P.nqtot + q]; }
  }
  if (mode == 0) {
    P.f[g] = F[0]; P.f[n + g] = F[1]; P.f[2 * n + g] = F[2];
    return;
  }
  P.f[g] = F[0]; P.f[n + g] = F[1]; P.f[2 * n + g] = F[2];
  integrate(P, s, g, li, F, X, PX, XN, PXN, step, mode);
}

// ------------------------------------------------------------------ lists
// util_periodic.F90:26-36 on the current coordinates; with with_pred the
// predicted wrapped positions px + (c4 v + dt^2/2 a) are refreshed too.
__global__ __launch_bounds__(kTile) void k_wrap(DevPtrs P, const double *X, double *PXc, double *PXp,
                                                const Tile *tiles, int with_pred) {
#pragma clang fp contract(off)
  const Tile tl = tiles[blockIdx.x];
  const DSys s = P.sys[tl.sys];
  const int li = tl.start + threadIdx.x;
  if (!s.periodic || li >= s.nmp) return;
  const int g = s.off + li, n = P.ntot;
  for (int d = 0; d < 3; ++d) {
    double pw = s.box[d], pmax = 0.5 * pw, pmin = -0.5 * pw;
    double tx = X[d * n + g];
    if (tx > pmax) tx = tx - pw * (double)((int)((tx - pmax) / pw) + 1);
    else if (tx < pmin) tx = tx + pw * (double)((int)((pmin - tx) / pw) + 1);
    PXc[d * n + g] = tx;
    if (with_pred) {
      double dx = P.rlan[3 * n + g] * P.v[d * n + g] + P.tstep2 * P.a[d * n + g];
      PXp[d * n + g] = tx + dx;
    }
  }
}

// DT15 keep rule for lo < hi (neighbor_assign.F90:481-559)
__device__ inline bool exv_keep(const DSys &s, int ulo, int uhi, int clo, int tlo, int lo, int hi) {
  if (ulo != uhi) return true;
  if (clo != 2) return true;  // ion / ligand unit
  int sep = tlo == 2 ? s.nsep[0] : (tlo == 4 ? s.nsep[1] : (tlo == 3 ? s.nsep[2] : -1));
  if (sep < 0) return false;
  return hi >= lo + sep;
}

// exv lists (neighbor_list.F90:56-87 + neighbor_assign.F90): every bead of
// the tile scans the system's beads staged through LDS, FP64 test
// d^2 < (cutoff + margin)^2 exactly as the reference rounds it.
__global__ __launch_bounds__(kTile) void k_list_exv(DevPtrs P, const double *X, const double *PX, const Tile *tiles) {
  __shared__ double sx[kTile], sy[kTile], sz[kTile];
  __shared__ int su[kTile], stc[kTile];
  const int lane = threadIdx.x;
  const Tile tl = tiles[blockIdx.x];
  const DSys s = P.sys[tl.sys];
  const int li = tl.start + lane;
  const bool valid = li < s.nmp;
  const int n = P.ntot;
  const double *Y = s.periodic ? PX : X;
  const int g = s.off + (valid ? li : 0);
  double xi[3] = {0, 0, 0};
  int ui = 0, ti = 0, ci = 0;
  if (valid) { ld3(Y, n, g, xi); ui = P.unit[g]; ti = P.mptype[g]; ci = P.iclass[g]; }
  int cnt = 0;
  for (int jt = 0; jt < s.nmp; jt += kTile) {
    int lj = jt + lane;
    if (lj < s.nmp) {
      int gj = s.off + lj;
      sx[lane] = Y[gj]; sy[lane] = Y[n + gj]; sz[lane] = Y[2 * n + gj];
      su[lane] = P.unit[gj];
      stc[lane] = (int)P.mptype[gj] | ((int)P.iclass[gj] << 8);
    }
    __syncthreads();
    const int kend = min(kTile, s.nmp - jt);
    if (valid)
      for (int k = 0; k < kend; ++k) {
        int lj2 = jt + k;
        if (lj2 == li) continue;
        const bool own_lo = li < lj2;
        double v[3];
        if (own_lo) { v[0] = __dadd_rn(sx[k], -xi[0]); v[1] = __dadd_rn(sy[k], -xi[1]); v[2] = __dadd_rn(sz[k], -xi[2]); }
        else { v[0] = __dadd_rn(xi[0], -sx[k]); v[1] = __dadd_rn(xi[1], -sy[k]); v[2] = __dadd_rn(xi[2], -sz[k]); }
        int code = s.periodic ? pbneighbor(s, v) : 0;
        double d2 = sq3_exact(v[0], v[1], v[2]);
        int uj = su[k];
        int ulo = own_lo ? ui : uj, uhi = own_lo ? uj : ui;
        double r2 = P.rl2[s.uoff + ulo * s.nunit + uhi];
        if (!(d2 < r2)) continue;
        int tj = stc[k] & 0xFF, cj = stc[k] >> 8;
        bool keep = own_lo ? exv_keep(s, ui, uj, ci, ti, li, lj2) : exv_keep(s, uj, ui, cj, tj, lj2, li);
        if (!keep) continue;
        int oc = own_lo ? code : mirror_flip(code);
        if (cnt < P.nb_cap) P.nb[(size_t)cnt * n + g] = ((uint32_t)oc << kMirrorShift) | (uint32_t)(s.off + lj2);
        ++cnt;
      }
    __syncthreads();
  }
  if (valid) {
    P.nb_cnt[g] = cnt;
    if (cnt > P.nb_cap) atomicMax(&P.counters[5], (unsigned long long)cnt);
  }
}

// ele bit matrix (neighbor_list_ele.F90:123-210): bit (qi, qj) set when the
// pair is in the list; radius (margin + cutoff)^2 per i_DH_cutoff_type.
__global__ __launch_bounds__(kTile) void k_list_ele(DevPtrs P, const double *X, const double *PX, const Tile *qtiles) {
  __shared__ double sx[kTile], sy[kTile], sz[kTile];
  const int lane = threadIdx.x;
  const Tile tl = qtiles[blockIdx.x];
  const DSys s = P.sys[tl.sys];
  const int qi = tl.start + lane;
  const bool valid = qi < s.nq;
  const int n = P.ntot;
  const double *Y = s.periodic ? PX : X;
  double xi[3] = {0, 0, 0};
  if (valid) ld3(Y, n, P.qbead[s.qoff + qi], xi);
  uint32_t *row = P.qmask + s.moff + (int64_t)(valid ? qi : 0) * s.mwords;
  for (int jt = 0; jt < s.nq; jt += kTile) {
    int qj = jt + lane;
    if (qj < s.nq) {
      int gj = P.qbead[s.qoff + qj];
      sx[lane] = Y[gj]; sy[lane] = Y[n + gj]; sz[lane] = Y[2 * n + gj];
    }
    __syncthreads();
    uint32_t w[2] = {0u, 0u};
    const int kend = min(kTile, s.nq - jt);
    if (valid)
      for (int k = 0; k < kend; ++k) {
        int qj2 = jt + k;
        if (qj2 == qi) continue;
        const bool own_lo = qi < qj2;
        double v[3];
        if (own_lo) { v[0] = __dadd_rn(sx[k], -xi[0]); v[1] = __dadd_rn(sy[k], -xi[1]); v[2] = __dadd_rn(sz[k], -xi[2]); }
        else { v[0] = __dadd_rn(xi[0], -sx[k]); v[1] = __dadd_rn(xi[1], -sy[k]); v[2] = __dadd_rn(xi[2], -sz[k]); }
        int code = s.periodic ? pbneighbor(s, v) : 0;
        double d2 = sq3_exact(v[0], v[1], v[2]);
        if (!(d2 < s.ele_rl2)) continue;
        w[k >> 5] |= 1u << (k & 31);
        if (s.periodic) P.qmirror[s.mioff + (int64_t)qi * s.nq + qj2] = (uint8_t)(own_lo ? code : mirror_flip(code));
      }
    if (valid) {
      row[jt >> 5] = w[0];
      if ((jt >> 5) + 1 < s.mwords) row[(jt >> 5) + 1] = w[1];
    }
    __syncthreads();
  }
}

// hb candidate list (neighbor_list_hb.F90:59-78; radius mloop_dtrna.F90:149)
__global__ void k_list_hb(DevPtrs P, const double *X, const double *PX, int h0, int nh) {
#pragma clang fp contract(off)
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nh) return;
  int h = h0 + k;
  const DSys s = P.sys[P.hb_sys[h]];
  const int n = P.ntot;
  int a = P.hb_mp[6 * h], b = P.hb_mp[6 * h + 1];
  double v[3];
  if (s.periodic) {
    v[0] = __dadd_rn(PX[a], -PX[b]); v[1] = __dadd_rn(PX[n + a], -PX[n + b]); v[2] = __dadd_rn(PX[2 * n + a], -PX[2 * n + b]);
    pbneighbor(s, v);
  } else {
    v[0] = __dadd_rn(X[a], -X[b]); v[1] = __dadd_rn(X[n + a], -X[n + b]); v[2] = __dadd_rn(X[2 * n + a], -X[2 * n + b]);
  }
  double d2 = sq3_exact(v[0], v[1], v[2]);
  double r = __dadd_rn(__dadd_rn(P.hb_nat[6 * h], s.hb_cut), s.margin);
  P.hb_inlist[h] = (uint8_t)(d2 < __dmul_rn(r, r));
}

__global__ void k_advance(uint64_t *stepbase, uint64_t by, unsigned long long *counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) { *stepbase += by; counters[4] += by; }
}

__global__ void k_count_rebuild(unsigned long long *counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counters[3] += 1;
}

// ------------------------------------------------------------------ energies
// energy_sumup.F90:149-529 for one system per block; deterministic tree sum.
constexpr int kEBlock = 256;
constexpr int kNE = 10;  // bond angle stack tstack hbond thbond exv ele velo spare
__device__ inline double block_sum(double v, double *red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = kEBlock / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kEBlock) void k_energy(DevPtrs P, const double *X, const double *PX,
                                                    const int32_t *sysidx, double *out) {
  __shared__ double red[kEBlock];
  const int t = threadIdx.x;
  const int is = sysidx[blockIdx.x];
  const DSys s = P.sys[is];
  const int n = P.ntot;
  double e[kNE];
  for (int k = 0; k < kNE; ++k) e[k] = 0.0;
  // bonds (energy_bond.F90:39-57)
  for (int b = s.boff + t; b < s.boff + s.nbd; b += kEBlock) {
    double v[3];
    vsub(X, n, P.bd_mp[2 * b + 1], P.bd_mp[2 * b], v);
    double dd = sqrt(dot3(v, v)) - P.bd_nat[b];
    e[0] += P.bd_coef[b] * (dd * dd);
  }
  // angles (energy_bangle.F90:36-52)
  for (int a = s.aoff + t; a < s.aoff + s.nba; a += kEBlock) {
    double v21[3], v32[3];
    vsub(X, n, P.ba_mp[3 * a + 1], P.ba_mp[3 * a], v21);
    vsub(X, n, P.ba_mp[3 * a + 2], P.ba_mp[3 * a + 1], v32);
    double ct = -dot3(v32, v21) / sqrt(dot3(v21, v21) * dot3(v32, v32));
    ct = ct > 1.0 ? 1.0 : (ct < -1.0 ? -1.0 : ct);
    double dba = acos(ct) - P.ba_nat[a];
    e[1] += P.ba_coef[a] * (dba * dba);
  }
  // consecutive stacks (energy_dtrna_stack.F90:85-160)
  for (int st = s.soff + t; st < s.soff + s.nst; st += kEBlock) {
    if (!P.st_status[st]) continue;
    const int32_t *m7 = P.st_mp + 7 * st;
    const double *nat = P.st_nat + 3 * st, *cf = P.st_coef + 4 * st;
    double v21[3], v34[3], v54[3], v56[3], v76[3], m[3], nn[3];
    vsub(X, n, m7[1], m7[0], v21); vsub(X, n, m7[2], m7[3], v34); vsub(X, n, m7[4], m7[3], v54);
    vsub(X, n, m7[4], m7[5], v56); vsub(X, n, m7[6], m7[5], v76);
    double dd = sqrt(dot3(v21, v21)) - nat[0];
    double ediv = 1.0 + cf[1] * dd * dd;
    cross3(v34, v54, m); cross3(v54, v56, nn);
    double dph = wrap_pi(atan2(dot3(v34, nn) * sqrt(dot3(v54, v54)), dot3(m, nn)) - nat[1]);
    ediv += cf[2] * dph * dph;
    cross3(v76, v56, m);
    for (int d = 0; d < 3; ++d) nn[d] = -nn[d];
    dph = wrap_pi(atan2(dot3(v76, nn) * sqrt(dot3(v56, v56)), dot3(m, nn)) - nat[2]);
    ediv += cf[3] * dph * dph;
    e[2] += cf[0] / ediv;
  }
  // tertiary stacks (energy_dtrna_stack_nlocal.F90:55-230)
  for (int k = s.toff + t; k < s.toff + s.ntst; k += kEBlock) {
    const int32_t *m6 = P.tst_mp + 6 * k;
    const double *nat = P.tst_nat + 6 * k, *cf = P.tst_coef + 7 * k;
    const double *Y = s.periodic ? PX : X;
    double v12[3], v13[3], v53[3], v42[3], v46[3], c4212[3], c1213[3], m[3], nn[3];
    vsub(Y, n, m6[0], m6[1], v12); vsub(Y, n, m6[0], m6[2], v13); vsub(Y, n, m6[4], m6[2], v53);
    vsub(Y, n, m6[3], m6[1], v42); vsub(Y, n, m6[3], m6[5], v46);
    if (s.periodic) { pbneighbor(s, v12); pbneighbor(s, v13); pbneighbor(s, v53); pbneighbor(s, v42); pbneighbor(s, v46); }
    double a12 = sqrt(dot3(v12, v12)), dd = a12 - nat[0];
    double ediv = 1.0 + cf[1] * dd * dd;
    double a13 = sqrt(dot3(v13, v13));
    double da = acos(dot3(v13, v12) / (a13 * a12)) - nat[1];
    ediv += cf[2] * da * da;
    double a42 = sqrt(dot3(v42, v42));
    da = acos(dot3(v12, v42) / (a12 * a42)) - nat[2];
    ediv += cf[3] * da * da;
    cross3(v42, v12, c4212); cross3(v12, v13, c1213);
    double dph = wrap_pi(atan2(dot3(v42, c1213) * a12, dot3(c4212, c1213)) - nat[3]);
    ediv += cf[4] * dph * dph;
    cross3(v53, v13, m);
    for (int d = 0; d < 3; ++d) nn[d] = -c1213[d];
    dph = wrap_pi(atan2(dot3(v53, nn) * a13, dot3(m, nn)) - nat[4]);
    ediv += cf[5] * dph * dph;
    for (int d = 0; d < 3; ++d) m[d] = -c4212[d];
    cross3(v42, v46, nn);
    dph = wrap_pi(atan2(dot3(v12, nn) * a42, dot3(m, nn)) - nat[5]);
    ediv += cf[6] * dph * dph;
    e[3] += cf[0] / ediv;
  }
  // hbonds: energies of the last force evaluation (energy_dtrna_hbond15.F90:119-148)
  for (int h = s.hoff + t; h < s.hoff + s.nhb; h += kEBlock)
    if (P.hb_inlist[h] && P.hb_status[h]) e[P.hb_tert[h] ? 5 : 4] += P.hb_ene[h];
  // excluded volume, each pair once by its lower bead (energy_exv_dt15.F90:60-116)
  {
    const double *Y = s.periodic ? PX : X;
    const double a = s.exv_a;
    for (int li = t; li < s.nmp; li += kEBlock) {
      int g = s.off + li;
      double xi[3];
      ld3(Y, n, g, xi);
      int cnt = min(P.nb_cnt[g], P.nb_cap);
      for (int m = 0; m < cnt; ++m) {
        uint32_t ent = P.nb[(size_t)m * n + g];
        int j = (int)(ent & kIdxMask);
        if (j < g) continue;
        double v[3] = {Y[j] - xi[0], Y[n + j] - xi[1], Y[2 * n + j] - xi[2]};
        if (s.periodic) { int c = (int)(ent >> kMirrorShift); for (int d = 0; d < 3; ++d) v[d] += mirror_shift(s, c, d); }
        int ti = P.mptype[g], tj = P.mptype[j];
        double dij;
        if ((ti == 2 && tj == 4) || (ti == 4 && tj == 2)) dij = s.exv_dist_PS;
        else if (P.iclass[g] == 2 && P.iclass[j] == 2) dij = s.exv_dist;
        else dij = P.exv_rad[g] + P.exv_rad[j];
        double dist = sqrt(dot3(v, v));
        if (dist > dij) continue;
        dist = dist + a - dij;
        if (dist > s.exv_inf) {
          double r = a / dist, r2 = r * r, r6 = r2 * r2 * r2;
          e[6] += ((double)P.exv_eps[g] * (double)P.exv_eps[j]) * (r6 * r6 - 2.0 * r6 + 1.0);
        } else {
          e[6] += kHighEnergy;
        }
      }
    }
  }
  // Debye-Hueckel, each pair once (energy_ele_DH.F90:33-100)
  if (s.ele_form == 0) {
    const double *Y = s.periodic ? PX : X;
    for (int qi = t; qi < s.nq; qi += kEBlock) {
      int gi = P.qbead[s.qoff + qi];
      double xi[3];
      ld3(Y, n, gi, xi);
      const uint32_t *row = P.qmask + s.moff + (int64_t)qi * s.mwords;
      for (int qj = qi + 1; qj < s.nq; ++qj) {
        if (!((row[qj >> 5] >> (qj & 31)) & 1u)) continue;
        int gj = P.qbead[s.qoff + qj];
        double v[3] = {Y[gj] - xi[0], Y[n + gj] - xi[1], Y[2 * n + gj] - xi[2]};
        if (s.periodic) { int c = P.qmirror[s.mioff + (int64_t)qi * s.nq + qj]; for (int d = 0; d < 3; ++d) v[d] += mirror_shift(s, c, d); }
        double d2 = dot3(v, v);
        if (d2 > s.dh_ecut2) continue;
        double d1 = sqrt(d2);
        e[7] += (P.charge[gi] * P.charge[gj]) * s.ele_coef / d1 * exp(-d1 * s.rcdist);
      }
    }
  }
  // kinetic (energy_kinetic.F90:33-44)
  for (int li = t; li < s.nmp; li += kEBlock) {
    int g = s.off + li;
    double v[3] = {P.v[g], P.v[n + g], P.v[2 * n + g]};
    e[8] += 0.5 * P.mass[g] * dot3(v, v);
  }
  double tot[kNE];
  for (int k = 0; k < kNE; ++k) tot[k] = block_sum(e[k], red);
  if (t == 0) {
    double *o = out + (size_t)blockIdx.x * 25;
    for (int k = 0; k < 25; ++k) o[k] = 0.0;
    o[2] = tot[0]; o[3] = tot[1]; o[15] = tot[2]; o[18] = tot[3]; o[16] = tot[4]; o[19] = tot[5];
    o[17] = tot[6]; o[7] = tot[7]; o[1] = tot[8];
    double sum = 0.0;
    for (int k = 2; k < 25; ++k) sum += o[k];
    o[0] = sum;
  }
}

}  // namespace tis
