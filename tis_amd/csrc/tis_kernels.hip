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
// Each bonded term is evaluated once (one lane per term) and its per-bead
// forces are written to the slots of the beads' incidence lists.

// force_bond.F90:40-58: fo[0..2] on bead 1, fo[3..5] on bead 2
__device__ inline void bond_eval(const DevPtrs &P, const double *X, int b, double fo[6]) {
  int i = P.bd_mp[2 * b], j = P.bd_mp[2 * b + 1];
  double v[3];
  vsub(X, P.ntot, j, i, v);
  double dist = sqrt(dot3(v, v));
  double dd = dist - P.bd_nat[b];
  double fr = P.bd_coef[b] * (-2.0 * dd / dist);
  for (int d = 0; d < 3; ++d) { fo[d] = -(fr * v[d]); fo[3 + d] = fr * v[d]; }
}

// force_bangle.F90:40-84
__device__ inline void angle_eval(const DevPtrs &P, const double *X, int a, double fo[9]) {
  for (int k = 0; k < 9; ++k) fo[k] = 0.0;
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
  for (int d = 0; d < 3; ++d) {
    double f21 = fr * (v21[d] * (c21 / c11) - v32[d]);
    double f32 = fr * (v32[d] * (c21 / c22) - v21[d]);
    fo[d] = -f21; fo[3 + d] = f21 - f32; fo[6 + d] = f32;
  }
}

// force_dtrna_stack.F90:53-165: fo[3*k + d] on bead k of B1 B2 P1 S1 P2 S2 P3
__device__ inline void stack_eval(const DevPtrs &P, const double *X, int st, double fo[21]) {
  const int n = P.ntot;
  const int32_t *m7 = P.st_mp + 7 * st;
  const double *nat = P.st_nat + 3 * st, *cf = P.st_coef + 4 * st;
  double v21[3], v34[3], v54[3], v56[3], v76[3], m[3], nn[3], fi[3], fl[3];
  vsub(X, n, m7[1], m7[0], v21);
  vsub(X, n, m7[2], m7[3], v34);
  vsub(X, n, m7[4], m7[3], v54);
  vsub(X, n, m7[4], m7[5], v56);
  vsub(X, n, m7[6], m7[5], v76);
  double ediv = 1.0;
  double dist = sqrt(dot3(v21, v21));
  double dd = dist - nat[0];
  ediv += cf[1] * dd * dd;
  for (int d = 0; d < 3; ++d) {
    fi[d] = 2.0 * cf[1] * dd / dist * v21[d];
    fo[d] = -fi[d]; fo[3 + d] = fi[d];
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
    fo[6 + d] = fi[d];
    fo[9 + d] = (-1.0 + d3454o) * fi[d] - d5654o * fl[d];
    fo[12 + d] = (-1.0 + d5654o) * fl[d] - d3454o * fi[d];
    fo[15 + d] = fl[d];
  }
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
    fo[18 + d] = fi[d];
    fo[15 + d] += (-1.0 + d7656o) * fi[d] - d5456o * fl[d];
    fo[12 + d] += -d7656o * fi[d] + (-1.0 + d5456o) * fl[d];
    fo[9 + d] += fl[d];
  }
  double sc = cf[0] / (ediv * ediv);
  for (int k = 0; k < 21; ++k) fo[k] *= sc;
}

__device__ inline void put_slots(const DevPtrs &P, int64_t base, int term, int nslot, const double *fo) {
  float *o = P.fslot + 3 * (base + (int64_t)nslot * term);   // the term's nslot 12-byte slots, contiguous
  for (int k = 0; k < 3 * nslot; ++k) o[k] = (float)fo[k];
}

// st_status of force_dtrna_stack_nlocal.F90:83-92 evaluated where it is
// read: consecutive stack st is off when one of the exclusive tertiary
// stacks that name it has d - d0 < 10 (the same FP64 operations as
// tstack_eval).
__device__ inline bool stack_on(const DevPtrs &P, const double *X, const double *PX, int st) {
  const int n = P.ntot;
  for (int e = P.st_exoff[st]; e < P.st_exoff[st + 1]; ++e) {
    const int t = P.st_ex[e];
    const DSys *sp = P.sys + P.tst_sys[t];
    const int32_t *m6 = P.tst_mp + 6 * t;
    double v12[3];
    if (sp->periodic) { vsub(PX, n, m6[0], m6[1], v12); pbneighbor(*sp, v12); }
    else vsub(X, n, m6[0], m6[1], v12);
    const double d1212 = dot3(v12, v12), a12 = sqrt(d1212);
    if (a12 - P.tst_nat[6 * t] < 10.0) return false;
  }
  return true;
}

// Bonded terms of the launch range: one lane per consecutive stack, angle or
// bond, in that order (the heaviest terms start first; the force pass runs
// these blocks last).

__device__ inline void bonded_lane(const DevPtrs &P, const double *X, const double *PX, const BondedRange &R,
                                   int t) {
  const int nb = R.nb, na = R.na, ns = R.ns, b0 = R.b0, a0 = R.a0, s0 = R.s0;
  if (t >= nb + na + ns) return;
  t = t < ns ? t + nb + na : (t < ns + na ? t - ns + nb : t - ns - na);   // stacks, angles, bonds
  if (t < nb) {
    double fo[6];
    bond_eval(P, X, b0 + t, fo);
    put_slots(P, P.sb_bd, b0 + t, 2, fo);
    return;
  }
  t -= nb;
  if (t < na) {
    double fo[9];
    angle_eval(P, X, a0 + t, fo);
    put_slots(P, P.sb_ba, a0 + t, 3, fo);
    return;
  }
  t -= na;
  if (t < ns) {
    const int st = s0 + t;
    double fo[21];
    if (stack_on(P, X, PX, st)) stack_eval(P, X, st, fo);
    else for (int k = 0; k < 21; ++k) fo[k] = 0.0;
    put_slots(P, P.sb_st, st, 7, fo);
  }
}

// ------------------------------------------------------------------ tertiary stack
// force_dtrna_stack_nlocal.F90:48-263 for one tertiary stack: all six slots
// into fo[18] (the exclusion of secondary stacks when d < 10 is stack_on).
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

// Orders the memory operations of a wave's lanes around it.
// A list build that overflowed the per-bead capacity (TIS_CNT_NB_OVERFLOW,
// set by the list passes) stops the run: every later kernel of the step
// sequence returns at once, so the state stays as it was when that
// rebuild's step began, and the host grows the capacity and resumes from
// that step.
__device__ inline bool run_stopped(const DevPtrs &P) {
  return __builtin_nontemporal_load(&P.counters[5]) != 0ull;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ inline bool site_in_excess(const DevPtrs &P, int site) {
  return site_count(P, site) - P.site_val[site] > 0;
}

// Valence-limit deletion of force_dtrna_hbond15.F90:373-510 for one system,
// run by the whole wave (every lane calls it).  The scans are wave-wide:
// sites in excess are counted and the pick-th one found by ballots over
// chunks of 64 sites, and the listed, formed hbonds of the picked site are
// gathered in list order by ballots over chunks of 64 candidates into the
// system's scratch sequence (P.hb_seq, one entry per candidate, so any
// length fits).  A sequence of at most 64 is moved into registers (lane i
// holds element i) and the shuffle and the Metropolis choice run on every
// lane alike; a longer one is shuffled and chosen from by lane 0 in the
// scratch copy.  Either way the draws of the counter-based generator are
// taken in the order of a serial pass, so that the result equals the
// oracle's bit for bit.  Lane 0 applies each deletion; the fence after it
// orders the next scan's reads.
__device__ int hbond_valence_wave(const DevPtrs &P, const DSys &s, uint64_t step, int lane) {
  const uint64_t below = (1ull << lane) - 1ull;
  int32_t *seq = P.hb_seq + s.hoff;
  int ndel = 0;
  uint32_t draw = 0;
  for (int iter = 0; iter < (1 << 20); ++iter) {
    int nex = 0;
    for (int k0 = 0; k0 < s.nsite; k0 += kTile) {
      const int k = k0 + lane;
      nex += __popcll(__ballot(k < s.nsite && site_in_excess(P, s.sitoff + k)));
    }
    if (nex == 0) break;
    double rnd = uniform01(P.seed, s.key, step, 0, draw++);
    const int pick = (int)ceil(rnd * nex) - 1;  // rnd in (0,1): 0 <= pick < nex
    int site = -1;
    for (int k0 = 0, c = 0; k0 < s.nsite; k0 += kTile) {
      const int k = k0 + lane;
      const bool ex = k < s.nsite && site_in_excess(P, s.sitoff + k);
      const uint64_t b = __ballot(ex);
      const int n = __popcll(b);
      if (pick < c + n) {
        const uint64_t w = __ballot(ex && c + __popcll(b & below) == pick);
        site = s.sitoff + k0 + __builtin_ctzll(w);
        break;
      }
      c += n;
    }
    // hbonds of the site in list order into the scratch sequence
    int nseq = 0;
    for (int q0 = 0; q0 < s.nhb; q0 += kTile) {
      const int q = q0 + lane;
      const int h = s.hoff + q;
      bool m = false;
      if (q < s.nhb && P.hb_inlist[h] && P.hb_status[h])
        for (int k = 0; k < 6; ++k) m = m || P.hb_site[6 * h + k] == site;
      const uint64_t b = __ballot(m);
      if (m) seq[nseq + __popcll(b & below)] = h;
      nseq += __popcll(b);
    }
    if (nseq == 0) break;
    wave_sync();
    int del;
    if (nseq <= kTile) {
      int myh = lane < nseq ? seq[lane] : -1;
      for (int i = 0; i < nseq; ++i) {
        rnd = uniform01(P.seed, s.key, step, 0, draw++);
        const int isw = (int)ceil(rnd * nseq) - 1;  // 0 <= isw < nseq
        const int a = __shfl(myh, i), c = __shfl(myh, isw);
        if (lane == i) myh = c;
        if (lane == isw) myh = a;
      }
      const double myene = lane < nseq ? P.hb_ene[myh] : 0.0;
      del = __shfl(myh, 0);
      double edel = __shfl(myene, 0);
      for (int i = 1; i < nseq; ++i) {
        const int j = __shfl(myh, i);
        const double ej = __shfl(myene, i);
        const double ratio = exp((ej - edel) * s.beta_hb);
        rnd = uniform01(P.seed, s.key, step, 0, draw++);
        if (rnd < ratio) { del = j; edel = ej; }
      }
    } else {
      del = -1;
      if (lane == 0) {
        uint32_t d = draw;
        for (int i = 0; i < nseq; ++i) {
          const double r = uniform01(P.seed, s.key, step, 0, d++);
          const int isw = (int)ceil(r * nseq) - 1;
          const int a = seq[i];
          seq[i] = seq[isw];
          seq[isw] = a;
        }
        del = seq[0];
        double edel = P.hb_ene[del];
        for (int i = 1; i < nseq; ++i) {
          const int j = seq[i];
          const double ej = P.hb_ene[j];
          const double ratio = exp((ej - edel) * s.beta_hb);
          if (uniform01(P.seed, s.key, step, 0, d++) < ratio) { del = j; edel = ej; }
        }
      }
      del = __shfl(del, 0);
      draw += 2u * (uint32_t)nseq - 1u;
    }
    if (lane == 0) {
      P.hb_status[del] = 0;
      P.hb_ene[del] = 0.0;
      // excess -= 1 where positive; count kept as valence + excess
      for (int k = 0; k < 6; ++k) {
        const int st = P.hb_site[6 * del + k];
        if (st >= 0 && site_in_excess(P, st)) atomicSub(&P.site_cnt[st], 1);
      }
    }
    ++ndel;
    wave_sync();
  }
  return ndel;
}

// One hbond of a system none of whose sites can exceed its valence: every
// listed hbond within the cutoff stays (force_dtrna_hbond15.F90:91-353 without
// the valence pass), so the hbonds of all such systems are independent lanes.
__device__ inline void hbond_lane(const DevPtrs &P, const double *X, const double *PX, int h) {
  const DSys &s = P.sys[P.hb_sys[h]];
  int st = 0;
  double e = 0.0, fo[18];
  if (P.hb_inlist[h]) st = hbond_eval(P, s, X, PX, h, fo, &e);
  if (!st) for (int q = 0; q < 18; ++q) fo[q] = 0.0;
  P.hb_status[h] = (uint8_t)st;
  P.hb_ene[h] = e;
  put_slots(P, P.sb_hb, h, 6, fo);
}

// ------------------------------------------------------------------ terms pass
// One wave per system that has tertiary stacks or hbond candidates some of
// whose sites can exceed their valence: tertiary stack forces, and for the
// latter the hbond geometry of listed candidates and the valence pass.  Forces go to per-term slots gathered by k_main.  The phases of a
// system are ordered by wave-level fences: the four waves of a block serve
// four systems.

__device__ void terms_wave(const DevPtrs &P, const double *X, const double *PX, int isys, int32_t t_local) {
  const int lane = threadIdx.x & (kTile - 1), nt = kTile;
  const DSys s = P.sys[isys];
  const uint64_t step = *P.stepbase + (uint64_t)t_local + 1u - s.step_origin;
  for (int k = lane; k < s.ntst; k += nt) {
    double fo[18];
    tstack_eval(P, s, X, PX, s.toff + k, fo);
    put_slots(P, P.sb_tst, s.toff + k, 6, fo);
  }
  // no site can exceed its valence: the hbonds run as lanes of their own
  // (hbond_lane)
  if (!s.hb_excess_possible) return;
  for (int k = lane; k < s.nsite; k += nt) P.site_cnt[s.sitoff + k] = 0;
  wave_sync();
  for (int k = lane; k < s.nhb; k += nt) {
    int h = s.hoff + k;
    int st = 0;
    double e = 0.0, fo[18];
    if (P.hb_inlist[h]) st = hbond_eval(P, s, X, PX, h, fo, &e);
    P.hb_status[h] = (uint8_t)st;
    P.hb_ene[h] = e;
    if (st) {
      for (int q = 0; q < 18; ++q) P.for_hb[(size_t)q * P.nhb + h] = fo[q];
      for (int q = 0; q < 6; ++q) {
        int site = P.hb_site[6 * h + q];
        if (site >= 0) atomicAdd(&P.site_cnt[site], 1);
      }
    }
  }
  wave_sync();
  int any = 0;
  for (int k = lane; k < s.nsite; k += nt)
    if (site_count(P, s.sitoff + k) - P.site_val[s.sitoff + k] > 0) any = 1;
  if (__any(any)) {
    const int nd = hbond_valence_wave(P, s, step, lane);
    if (nd && lane == 0) atomicAdd(&P.counters[2], (unsigned long long)nd);
  }
  wave_sync();
  // forces of the hbonds that survive into their beads' slots
  for (int k = lane; k < s.nhb; k += nt) {
    const int h = s.hoff + k;
    double fo[18];
    const bool on = P.hb_status[h] != 0;
    for (int q = 0; q < 18; ++q) fo[q] = on ? P.for_hb[(size_t)q * P.nhb + h] : 0.0;
    put_slots(P, P.sb_hb, h, 6, fo);
  }
}

// ------------------------------------------------------------------ electrostatics
// Pair magnitude dv (force on i = -dv * (x_j - x_i)) in FP32 for the three
// forms: Debye-Hueckel (force_ele_DH.F90:100-112, clamped at DE_MAX), plain
// Coulomb (force_ele_coulomb.F90:75-79) and the Ewald real-space term
// (force_ele_coulomb_ewld.F90:128-132).
// Hardware reciprocal square root and base-2 exponential (about 1 ulp);
// d2 is a normal positive number here and exp underflows to 0 gracefully.
__device__ inline float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline float hw_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

template <int FORM>
__device__ inline float ele_dv(const DSys &s, float qq, float d2, unsigned &nclamp, bool count) {
  const float rinv = hw_rsq(d2);
  if (FORM == 0) {
    const float rcd = (float)s.rcdist;
    const float d1 = d2 * rinv;
    const float dv = qq * rinv * rinv * (rinv + rcd) * hw_exp(-d1 * rcd);
    const bool cl = dv > (float)kDeMax;
    nclamp += (cl && count) ? 1u : 0u;
    return cl ? (float)kDeMax : dv;
  } else if (FORM == 1) {
    return qq * rinv * rinv * rinv;
  } else {
    const float d1 = d2 * rinv;
    return qq * (erfcf((float)s.ew_alpha * d1) * rinv + (float)s.ew_beta * hw_exp(-(float)s.ew_a2 * d2)) * rinv * rinv;
  }
}

// Staging of the J chunk of each wave's chunk pair.
// The J chunk of each wave's chunk pair: ele_pair_task uses the first 64
// entries of each array, ele_pair_unrolled keeps the chunk twice over (so
// that its reads need no wrap-around).  Arrays, not records, so that the 64
// lanes' reads of consecutive entries fall in distinct LDS banks.
struct EleWave { double x[2 * kTile], y[2 * kTile], z[2 * kTile]; float q[2 * kTile]; };
struct EleLds {
  EleWave w[kWaves];
};

// Whole-wave rotation by one lane (DPP wave_rol:1): each lane receives the
// value of a neighbouring lane.  The direction is read off at run time from
// the lane ids (wave_dir), so the pair schedule below does not depend on it.
__device__ inline uint32_t wave_rot(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x134, 0xF, 0xF, false);
}
__device__ inline double wave_rot(double v) {
  const uint64_t b = __double_as_longlong(v);
  const uint64_t lo = wave_rot((uint32_t)b), hi = wave_rot((uint32_t)(b >> 32));
  return __longlong_as_double((long long)(lo | (hi << 32)));
}
// +1 or 63 (= -1 mod 64) when wave_rot is a uniform rotation of the wave,
// else 0
__device__ inline int wave_dir() {
  const int lane = threadIdx.x & (kTile - 1);
  const int d = ((int)wave_rot((uint32_t)lane) - lane) & (kTile - 1);
  const int d0 = __shfl(d, 0);
  const bool ok = (d0 == 1 || d0 == kTile - 1) && __all(d == d0);
  return ok ? d0 : 0;
}

// force_ele_DH.F90:68-118, force_ele_coulomb.F90:63-84 and the real-space
// loop of force_ele_coulomb_ewld.F90:118-137.  One wave per chunk pair
// (I, J), I <= J, of a system's charges (chunks of 64), four pairs per
// block.  At iteration k lane l pairs its charge i of chunk I with
// j = (l + dir k) mod 64 of chunk J (staged in the wave's LDS slice); the
// force on i accumulates in the lane, the opposite force on j in a register
// that rotates with the schedule, so after 64 iterations lane l holds the
// sum for j = l.  The diagonal chunk (J = I) runs k = 1..32 (k = 32 for
// lanes < 32 only), which meets every unordered pair of the chunk once.
// List membership from the bit matrix, cutoff test in FP64, pair arithmetic
// in FP32, accumulation in FP64; periodic pairs take the mirror code the list
// stored for them (v = x_j - x_i + shift, util_pbneighbor.F90:22-35).  The
// forces on chunk I go to slot 0 of the task's partial buffer, those on
// chunk J (I < J) to slot 1, and k_main sums a charge's slots in task order:
// a fixed order, so results are bitwise reproducible.  Charges beyond nq are
// staged as zeros with their bits clear.  dir = 0 (no whole-wave DPP
// rotation) rotates with __shfl instead.
__host__ __device__ inline int64_t ele_task_index(int I, int J, int nch) {
  return (int64_t)I * (2 * nch - I + 1) / 2 + (J - I);
}

// the next value of the rotating sums: DPP when the whole-wave rotation is
// available (decided once per wave, so the step loop has no branch), else
// __shfl
template <bool DPP>
__device__ inline double rot_next(double v) {
  if (DPP) return wave_rot(v);
  const int lane = threadIdx.x & (kTile - 1);
  return __shfl(v, (lane + 1) & (kTile - 1));
}

template <int FORM, bool DPP>
__device__ inline void ele_pair_task(const DevPtrs &P, const DSys &s, const double *Y, int I, int J, int dir,
                                     double *SX, double *SY, double *SZ, float *SQ) {
  const int lane = threadIdx.x & (kTile - 1);
  const int rd = dir != 0 ? dir : 1;
  const int n = P.ntot;
  const int qi = I * kTile + lane;
  const bool valid = qi < s.nq;
  double xi0 = 0.0, xi1 = 0.0, xi2 = 0.0, chi = 0.0;
  if (valid) {
    const int gi = P.qbead[s.qoff + qi];
    xi0 = Y[gi]; xi1 = Y[n + gi]; xi2 = Y[2 * n + gi];
    chi = P.charge[gi];
  }
  const int jt = J * kTile;
  {
    const bool vj = jt + lane < s.nq;
    const int gj = vj ? P.qbead[s.qoff + jt + lane] : 0;
    SX[lane] = vj ? Y[gj] : 0.0; SY[lane] = vj ? Y[n + gj] : 0.0; SZ[lane] = vj ? Y[2 * n + gj] : 0.0;
    SQ[lane] = vj ? (float)P.charge[gj] : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const float qiC = (float)(chi * s.ele_coef);
  const double cut2 = FORM == 0 ? s.dh_cut2 : s.cl_cut2;
  const uint32_t *row = P.qmask + s.moff + (int64_t)(valid ? qi : 0) * s.mwords;
  const uint8_t *mrow = s.periodic ? P.qmirror + s.mioff + (int64_t)(valid ? qi : 0) * s.nq + jt : nullptr;
  uint64_t m64 = 0;
  if (valid) {
    const int wb = jt >> 5;
    m64 = row[wb];
    if (wb + 1 < s.mwords) m64 |= (uint64_t)row[wb + 1] << 32;
  }
  const bool diag = I == J;
  double F0 = 0.0, F1 = 0.0, F2 = 0.0, G0 = 0.0, G1 = 0.0, G2 = 0.0;
  unsigned nclamp = 0;
  auto pair = [&](int k, bool last_half) {
    const int jl = (lane + rd * k) & (kTile - 1);
    bool bit = (m64 >> jl) & 1u;
    if (last_half) bit = bit && lane < 32;
    // a step with no listed pair in the wave adds zeros: only the sums move
    if (__any(bit)) {
      double v0 = SX[jl] - xi0, v1 = SY[jl] - xi1, v2 = SZ[jl] - xi2;
      if (s.periodic) {
        const int code = bit ? (int)mrow[jl] : 0;
        v0 += mirror_shift(s, code, 0); v1 += mirror_shift(s, code, 1); v2 += mirror_shift(s, code, 2);
      }
      const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
      const bool on = bit && d2 <= cut2;
      float dv = ele_dv<FORM>(s, qiC * SQ[jl], on ? (float)d2 : 1.0f, nclamp, on);
      dv = on ? dv : 0.0f;
      const double dd = (double)dv;
      F0 -= dd * v0; F1 -= dd * v1; F2 -= dd * v2;
      G0 += dd * v0; G1 += dd * v1; G2 += dd * v2;
    }
    G0 = rot_next<DPP>(G0); G1 = rot_next<DPP>(G1); G2 = rot_next<DPP>(G2);
  };
  if (__any(m64 != 0)) {
    if (diag) {
#pragma unroll 1
      for (int k = 1; k < 32; ++k) pair(k, false);
      pair(32, true);
    } else {
#pragma unroll 1
      for (int k = 0; k < kTile; ++k) pair(k, false);
    }
  }
  double *pp = P.pp + s.pp_off + ele_task_index(I, J, s.nqt) * (6 * kTile);
  if (diag) {
    const int src = (lane - rd * 33) & (kTile - 1);
    F0 += __shfl(G0, src); F1 += __shfl(G1, src); F2 += __shfl(G2, src);
  } else {
    pp[3 * kTile + lane] = G0; pp[4 * kTile + lane] = G1; pp[5 * kTile + lane] = G2;
  }
  pp[lane] = F0; pp[kTile + lane] = F1; pp[2 * kTile + lane] = F2;
  if (nclamp) atomicAdd(&P.counters[1], (unsigned long long)nclamp);
}

// ele_pair_task for a non-periodic system with the whole-wave DPP rotation
// (dir = 1 or 63): the same pairs in the same order and the same FP64
// distances, cutoff tests and accumulation (the DH exponent takes its scale
// as one constant, -rcd log2(e), so its FP32 rounding can differ by an ulp),
// in fewer instructions per step: the J chunk is
// staged twice over in an order where lane l meets, at step k, entry c + k
// (c = l for dir = 1, 63 - l for dir = 63), so the reads of eight
// consecutive steps take immediate offsets from one address; the list bits
// are turned into the same order once (bit k of mr is the pair of step k); a
// pair that is not on gets charge product 0, so its force is exactly zero
// and adds nothing, with no branch; the DE_MAX clamps are counted per wave.
template <int FORM>
__device__ inline void ele_pair_unrolled(const DevPtrs &P, const DSys &s, const double *Y, int I, int J, int dir,
                                         EleWave *JE) {
  // the chunk pair is the wave's: uniform, so the step loops below are scalar
  I = __builtin_amdgcn_readfirstlane(I);
  J = __builtin_amdgcn_readfirstlane(J);
  const int lane = threadIdx.x & (kTile - 1);
  const int c = dir == 1 ? lane : kTile - 1 - lane;
  const int n = P.ntot;
  const int qi = I * kTile + lane;
  const bool valid = qi < s.nq;
  // absent charges sit far apart (+-1e15, charge 0): their pairs stay
  // finite and off
  constexpr double kFar = 1e15;
  double xi0 = -kFar, xi1 = -kFar, xi2 = -kFar, chi = 0.0;
  if (valid) {
    const int gi = P.qbead[s.qoff + qi];
    xi0 = Y[gi]; xi1 = Y[n + gi]; xi2 = Y[2 * n + gi];
    chi = P.charge[gi];
  }
  const int jt = J * kTile;
  {
    const bool vj = jt + lane < s.nq;
    const int gj = vj ? P.qbead[s.qoff + jt + lane] : 0;
    const double x = vj ? Y[gj] : kFar, y = vj ? Y[n + gj] : kFar, z = vj ? Y[2 * n + gj] : kFar;
    const float q = vj ? (float)P.charge[gj] : 0.0f;
    JE->x[c] = x; JE->y[c] = y; JE->z[c] = z; JE->q[c] = q;
    JE->x[c + kTile] = x; JE->y[c + kTile] = y; JE->z[c + kTile] = z; JE->q[c + kTile] = q;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const float qiC = (float)(chi * s.ele_coef);
  const double cut2 = FORM == 0 ? s.dh_cut2 : s.cl_cut2;
  const uint32_t *row = P.qmask + s.moff + (int64_t)(valid ? qi : 0) * s.mwords;
  uint64_t m64 = 0;
  if (valid) {
    const int wb = jt >> 5;
    m64 = row[wb];
    if (wb + 1 < s.mwords) m64 |= (uint64_t)row[wb + 1] << 32;
  }
  const bool diag = I == J;
  // bit k of mr: the pair of step k, j = (lane + dir k) mod 64
  uint64_t mr = __builtin_rotateright64(dir == 1 ? m64 : __builtin_bitreverse64(m64), (uint64_t)c);
  if (diag && lane >= 32) mr &= ~(1ull << 32);   // step 32 of the diagonal chunk: lanes < 32 only
  double F0 = 0.0, F1 = 0.0, F2 = 0.0, G0 = 0.0, G1 = 0.0, G2 = 0.0;
  uint32_t nclamp = 0;
  const float rcd = (float)s.rcdist, nrl = -rcd * 1.44269504088896341f;   // exp(-r rcd) = 2^(r nrl)
  // eight steps from k0; BITS false: every pair of the chunk pair is listed,
  // so only the cutoff decides (a charge beyond nq has charge 0 and adds 0)
  auto steps8 = [&](const int k0, const uint32_t b8, auto bits_t) {
    constexpr bool BITS = decltype(bits_t)::value;
    const double *px = JE->x + c + k0, *py = JE->y + c + k0, *pz = JE->z + c + k0;
    const float *pq = JE->q + c + k0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const double v0 = px[kk] - xi0, v1 = py[kk] - xi1, v2 = pz[kk] - xi2;
      const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
      const bool on = (!BITS || ((b8 >> kk) & 1u)) && d2 <= cut2;
      float sq = pq[kk];
      asm volatile("" : "+v"(sq));   // read for every lane: a select below, not a branch
      const float qq = on ? qiC * sq : 0.0f;
      // every pair listed: an off pair is beyond the cutoff (d2 > 0)
      const float d2f = BITS && !on ? 1.0f : (float)d2;
      const float rinv = hw_rsq(d2f);
      float dv;
      if (FORM == 0) {
        const float d1 = d2f * rinv;
        dv = qq * rinv * rinv * (rinv + rcd) * __builtin_amdgcn_exp2f(d1 * nrl);
        const bool cl = dv > (float)kDeMax;
        nclamp += (uint32_t)__popcll(__ballot(cl));
        dv = cl ? (float)kDeMax : dv;
      } else {
        dv = qq * rinv * rinv * rinv;
      }
      const double dd = (double)dv;
      F0 -= dd * v0; F1 -= dd * v1; F2 -= dd * v2;
      G0 += dd * v0; G1 += dd * v1; G2 += dd * v2;
      G0 = wave_rot(G0); G1 = wave_rot(G1); G2 = wave_rot(G2);
    }
  };
  if (__any(m64 != 0)) {
    // every valid j listed for every valid lane (off-diagonal chunk pairs)
    const int nj = s.nq - jt < kTile ? s.nq - jt : kTile;
    const uint64_t vjm = nj >= kTile ? ~0ull : (1ull << nj) - 1ull;
    const bool all_listed = !diag && __all(!valid || (m64 & vjm) == vjm);
    if (all_listed) {
#pragma unroll 1
      for (int k0 = 0; k0 < kTile; k0 += 8) steps8(k0, 0u, std::false_type{});
    } else {
      const int k_begin = diag ? 1 : 0, k_end = diag ? 33 : kTile;
#pragma unroll 1
      for (int k0 = k_begin; k0 < k_end; k0 += 8) steps8(k0, (uint32_t)(mr >> k0), std::true_type{});
    }
  }
  double *pp = P.pp + s.pp_off + ele_task_index(I, J, s.nqt) * (6 * kTile);
  if (diag) {
    const int src = (lane - dir * 33) & (kTile - 1);
    F0 += __shfl(G0, src); F1 += __shfl(G1, src); F2 += __shfl(G2, src);
  } else {
    pp[3 * kTile + lane] = G0; pp[4 * kTile + lane] = G1; pp[5 * kTile + lane] = G2;
  }
  pp[lane] = F0; pp[kTile + lane] = F1; pp[2 * kTile + lane] = F2;
  if (lane == 0 && nclamp) atomicAdd(&P.counters[1], (unsigned long long)nclamp);
}

// The force pass, one launch with four kinds of blocks: 256 hbond lanes per
// block (hbond_lane: the hbonds of systems without valence-limited sites, one
// lane each, first because each lane is a long FP64 chain), four systems'
// tertiary stacks and valence-limited hbonds per block (terms_wave, one wave
// each), four electrostatic chunk pairs per block (ele_pair_task, one wave
// each), then the bonded terms, 256 lanes per block (bonded_lane).  k_main
// gathers all of their outputs.
template <int PPW, int TPL>
__global__ __launch_bounds__(kTile * kWaves) __attribute__((amdgpu_waves_per_eu(4)))
void k_force(DevPtrs P, const double *X, const double *PX, const int32_t *systask, int ntask,
             const EwTask *ptasks, int npt, BondedRange br, int32_t t_local) {
  constexpr int ppw = PPW, tpl = TPL;
  __shared__ EleLds lds;
  if (run_stopped(P)) return;
  int b = (int)blockIdx.x;
  const int nhdb = (br.nhd + kTile * kWaves - 1) / (kTile * kWaves);   // blocks of 256 hbond lanes
  const int npb = (npt + kWaves * ppw - 1) / (kWaves * ppw);   // blocks of 4 ppw chunk pairs
  const int ntb = (ntask + kWaves - 1) / kWaves;   // blocks of four system terms passes
  if (b < nhdb) {
    const int t = b * (kTile * kWaves) + (int)threadIdx.x;
    if (t < br.nhd) hbond_lane(P, X, PX, P.hb_dense[br.hd0 + t]);
    return;
  }
  b -= nhdb;
  if (b < ntb) {
    const int task = b * kWaves + (int)threadIdx.x / kTile;
    if (task < ntask) terms_wave(P, X, PX, systask[task], t_local);
    return;
  }
  b -= ntb;
  if (b >= npb) {
    // tpl bonded terms per lane, a grid-wide stride apart
    const int nbb = (int)gridDim.x - nhdb - ntb - npb;
    for (int i = 0; i < tpl; ++i)
      bonded_lane(P, X, PX, br, (b - npb + i * nbb) * (kTile * kWaves) + (int)threadIdx.x);
    return;
  }
  const int w = (int)threadIdx.x / kTile;
  const int dir = wave_dir();
  EleWave *JE = &lds.w[w];
  double *SX = JE->x, *SY = JE->y, *SZ = JE->z;
  float *SQ = JE->q;
  // ppw consecutive chunk pairs per wave, one after the other through the
  // wave's LDS slice
#pragma unroll 1
  for (int i = 0; i < ppw; ++i) {
    const int task = (b * kWaves + w) * ppw + i;
    if (task >= npt) return;
    if (i) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const EwTask tk = ptasks[task];
    const DSys s = P.sys[tk.sys];
    const double *Y = s.periodic ? PX : X;
    if (dir != 0 && !s.periodic && s.ele_form == 0) {
      ele_pair_unrolled<0>(P, s, Y, tk.a, tk.b, dir, JE);
    } else if (dir != 0 && !s.periodic && s.ele_form == 1) {
      ele_pair_unrolled<1>(P, s, Y, tk.a, tk.b, dir, JE);
    } else if (dir != 0) {
      if (s.ele_form == 0) ele_pair_task<0, true>(P, s, Y, tk.a, tk.b, dir, SX, SY, SZ, SQ);
      else if (s.ele_form == 1) ele_pair_task<1, true>(P, s, Y, tk.a, tk.b, dir, SX, SY, SZ, SQ);
      else ele_pair_task<2, true>(P, s, Y, tk.a, tk.b, dir, SX, SY, SZ, SQ);
    } else {
      if (s.ele_form == 0) ele_pair_task<0, false>(P, s, Y, tk.a, tk.b, dir, SX, SY, SZ, SQ);
      else if (s.ele_form == 1) ele_pair_task<1, false>(P, s, Y, tk.a, tk.b, dir, SX, SY, SZ, SQ);
      else ele_pair_task<2, false>(P, s, Y, tk.a, tk.b, dir, SX, SY, SZ, SQ);
    }
  }
}

template __global__ void k_force<1, 1>(DevPtrs, const double *, const double *, const int32_t *, int, const EwTask *,
                                       int, BondedRange, int32_t);
template __global__ void k_force<2, 1>(DevPtrs, const double *, const double *, const int32_t *, int, const EwTask *,
                                       int, BondedRange, int32_t);
template __global__ void k_force<1, 2>(DevPtrs, const double *, const double *, const int32_t *, int, const EwTask *,
                                       int, BondedRange, int32_t);
template __global__ void k_force<2, 2>(DevPtrs, const double *, const double *, const int32_t *, int, const EwTask *,
                                       int, BondedRange, int32_t);

// ------------------------------------------------------------------ Ewald reciprocal space
// force_ele_coulomb_ewld.F90:62-84.  The phase k.r = 2 pi (h . r/L) is
// formed in FP64 from fractional coordinates and reduced to (-1/2, 1/2]
// turns before the FP32 sincospi, so its rounding does not grow with |h| or
// the box; structure factors and forces accumulate in FP64.
//
// Pass 1: partial structure factors.  One lane per reciprocal vector of a
// 64-vector tile, over one 64-charge chunk staged in LDS (broadcast reads).
// exact = 1 (energy evaluation) takes sin/cos in FP64.
__global__ __launch_bounds__(kTile) void k_ewald_sk(DevPtrs P, const double *PX, const EwTask *tasks, int exact) {
  __shared__ double fx[kTile], fy[kTile], fz[kTile], fq[kTile];
  if (run_stopped(P)) return;
  const EwTask tk = tasks[blockIdx.x];
  const DSys s = P.sys[tk.sys];
  const int lane = threadIdx.x, n = P.ntot;
  const int qi = tk.b * kTile + lane;
  if (qi < s.nq) {
    const int g = P.qbead[s.qoff + qi];
    fx[lane] = PX[g] * s.ew_invL; fy[lane] = PX[n + g] * s.ew_invL; fz[lane] = PX[2 * n + g] * s.ew_invL;
    fq[lane] = P.charge[g];
  } else {
    fx[lane] = fy[lane] = fz[lane] = 0.0; fq[lane] = 0.0;
  }
  __syncthreads();
  const int k = tk.a * kTile + lane;
  if (k >= s.nk) return;
  const int kg = s.ewk0 + k;
  const double hx = P.ewh[kg], hy = P.ewh[P.nktot + kg], hz = P.ewh[2 * P.nktot + kg];
  const int jn = min(kTile, s.nq - tk.b * kTile);
  double sc = 0.0, ss = 0.0;
  if (exact) {
    for (int j = 0; j < jn; ++j) {
      double t = hx * fx[j] + hy * fy[j] + hz * fz[j];
      t -= rint(t);
      double sn, cs;
      sincospi(2.0 * t, &sn, &cs);
      sc += fq[j] * cs;
      ss += fq[j] * sn;
    }
  } else {
    for (int j = 0; j < jn; ++j) {
      double t = hx * fx[j] + hy * fy[j] + hz * fz[j];
      t -= rint(t);
      float sn, cs;
      sincospif((float)(2.0 * t), &sn, &cs);
      sc += fq[j] * (double)cs;
      ss += fq[j] * (double)sn;
    }
  }
  double *o = P.psk + s.psk_off + 2 * ((int64_t)tk.b * s.nk + k);
  o[0] = sc; o[1] = ss;
}

// Pass 2: forces.  A block takes a 64-charge tile and a chunk of 256
// reciprocal vectors: it first sums the partial structure factors of its
// vectors (fixed chunk order) into LDS, then each wave sweeps 64 of them for
// its lane's charge; partial forces per chunk go to pf (summed by k_main).
__global__ __launch_bounds__(kTile * kWaves) void k_ewald_f(DevPtrs P, const double *PX, const EwTask *tasks) {
  __shared__ double shx[kEwKChunk], shy[kEwKChunk], shz[kEwKChunk], sfc[kEwKChunk], ssc[kEwKChunk], sss[kEwKChunk];
  __shared__ double red[3][kWaves][kTile];
  if (run_stopped(P)) return;
  const EwTask tk = tasks[blockIdx.x];
  const DSys s = P.sys[tk.sys];
  const int tid = threadIdx.x, n = P.ntot;
  {
    const int k = tk.b * kEwKChunk + tid;
    double sc = 0.0, ss = 0.0, fc = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
    if (k < s.nk) {
      const double *ps = P.psk + s.psk_off + 2 * (int64_t)k;
      for (int c = 0; c < s.nqc; ++c) { sc += ps[2 * (int64_t)c * s.nk]; ss += ps[2 * (int64_t)c * s.nk + 1]; }
      const int kg = s.ewk0 + k;
      fc = P.ewc[kg]; hx = P.ewh[kg]; hy = P.ewh[P.nktot + kg]; hz = P.ewh[2 * P.nktot + kg];
    }
    shx[tid] = hx; shy[tid] = hy; shz[tid] = hz; sfc[tid] = fc; ssc[tid] = sc; sss[tid] = ss;
  }
  __syncthreads();
  const int lane = tid & (kTile - 1), w = tid / kTile;
  const int qi = tk.a * kTile + lane;
  const bool valid = qi < s.nq;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  if (valid) {
    const int g = P.qbead[s.qoff + qi];
    const double x0 = PX[g] * s.ew_invL, x1 = PX[n + g] * s.ew_invL, x2 = PX[2 * n + g] * s.ew_invL;
    const int kb = w * kTile, ke = min(kb + kTile, s.nk - tk.b * kEwKChunk);
    for (int k = kb; k < ke; ++k) {
      double t = shx[k] * x0 + shy[k] * x1 + shz[k] * x2;
      t -= rint(t);
      float sn, cs;
      sincospif((float)(2.0 * t), &sn, &cs);
      const double term = sfc[k] * ((double)sn * ssc[k] - (double)cs * sss[k]);
      f0 += term * shx[k]; f1 += term * shy[k]; f2 += term * shz[k];
    }
  }
  red[0][w][lane] = f0; red[1][w][lane] = f1; red[2][w][lane] = f2;
  __syncthreads();
  if (w == 0 && valid) {
    const double scale = s.ew_fscale * P.charge[P.qbead[s.qoff + qi]];
    for (int d = 0; d < 3; ++d)
      P.pf[s.pf_off + ((int64_t)tk.b * 3 + d) * s.nq + qi] =
          scale * (((red[d][0][lane] + red[d][1][lane]) + red[d][2][lane]) + red[d][3][lane]);
  }
}

// ------------------------------------------------------------------ main pass
// mode 0: store F; mode 1: Langevin step; mode 2: initial acceleration;
// mode 3: store F and predict positions from the stored v, a; mode 4: mode
// 2 or 3 per system (DSys::init_accel).
// Langevin (time_integral.F90:169-183, then the next step's position update
// :152-160): a' = F/m + c1 xi; v = c2 v + c3 (a + a'); a = a';
// x_next = x + c4 v + (dt^2/2) a.
// A bead's integrator inputs (Langevin constants, 1/m, v, a, x, wrapped x,
// displacement since the last rebuild), loaded before the barrier of k_main
// so that their latency overlaps the other wave's work.
struct BeadState { double r[4], rcm, v[3], a[3], x[3], px[3], o[3]; };

__device__ inline void load_state(const DevPtrs &P, const DSys &s, int g, const double *X, const double *PX,
                                  BeadState &b) {
  const int n = P.ntot;
  const LangClass &lc = P.lcls[s.lc0 + P.lcl[g]];
  for (int k = 0; k < 4; ++k) b.r[k] = lc.r[k];
  b.rcm = lc.rcm;
  for (int d = 0; d < 3; ++d) {
    b.v[d] = P.v[d * n + g];
    b.a[d] = P.a[d * n + g];
    b.x[d] = X[d * n + g];
    b.px[d] = s.periodic ? PX[d * n + g] : 0.0;
    b.o[d] = P.dyn ? P.dxyz[d * n + g] : 0.0;
  }
}

__device__ inline void integrate(const DevPtrs &P, const DSys &s, int g, const BeadState &b, const double gs[3],
                                 const double F[3], double *XN, double *PXN, int mode, double dxo[3]) {
#pragma clang fp contract(off)
  const int n = P.ntot;
  for (int d = 0; d < 3; ++d) {
    double af = F[d] * b.rcm + b.r[0] * gs[d];
    double vv = b.v[d];
    const double a0 = b.a[d];
    if (mode == 3) af = a0;
    if (mode == 1) vv = b.r[1] * vv + b.r[2] * (a0 + af);
    P.v[d * n + g] = vv;
    P.a[d * n + g] = af;
    double dx = b.r[3] * vv + P.tstep2 * af;
    dxo[d] = dx;
    XN[d * n + g] = b.x[d] + dx;
    if (s.periodic) PXN[d * n + g] = b.px[d] + dx;
  }
}

// The boxes of a tile of 64 positions (one per lane of the wave): the tile's
// box to *tile and, for bead tiles (groups != nullptr), the boxes of its four
// 16-bead groups.
__device__ inline void tile_boxes(double p0, double p1, double p2, BBox *tile, BBox *groups) {
  const int lane = threadIdx.x & (kTile - 1);
  double lo[3] = {p0, p1, p2}, hi[3] = {p0, p1, p2};
#pragma unroll
  for (int off = 1; off < kTile; off <<= 1) {
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off));
    }
    // after the offsets 1..8 each group of 16 lanes holds its group's box
    if (off == 8 && groups && (lane & 15) == 0) {
      BBox b;
      for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.hi[d] = hi[d]; }
      groups[lane / 16] = b;
    }
  }
  if (lane == 0) {
    BBox b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.hi[d] = hi[d]; }
    *tile = b;
  }
}

// Two waves share a tile of 64 beads: wave 0 sums each bead's bonded slots,
// wave 1 draws its Langevin noise and adds its electrostatic sums, and the
// two waves take alternate groups of four excluded-volume neighbours; the
// partial sums are added in wave order (a fixed order: results are bitwise
// reproducible), then wave 0 integrates.
__device__ inline void exv_entries(const DevPtrs &P, int g, int m0, int cnt, uint32_t e[4]) {
  // entries past cnt re-read entry cnt - 1 (m0 < cnt) and are skipped
#pragma unroll
  for (int u = 0; u < 4; ++u) e[u] = P.nb[(size_t)min(m0 + u, cnt - 1) * P.ntot + g];
}

// e: the group's list entries (exv_entries); the entries of the wave's next
// group (m0 + 8) are loaded into en while this group's neighbours are read
__device__ inline void exv_group(const DevPtrs &P, const DSys &s, const double *Y, int g, int m0, int cnt,
                                 double xi0, double xi1, double xi2, const SpPair *tab, double F[3],
                                 unsigned &nclamp, const uint32_t e[4], uint32_t en[4]) {
  const int n = P.ntot;
  // every load of the group is issued before the arithmetic
  double y0[4], y1[4], y2[4];
  SpPair tp[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int j = (int)(e[u] & kIdxMask);
    y0[u] = Y[j]; y1[u] = Y[n + j]; y2[u] = Y[2 * n + j];
    tp[u] = tab[P.species[j]];
  }
  if (m0 + 8 < cnt) exv_entries(P, g, m0 + 8, cnt, en);
  const double a = s.exv_a;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (m0 + u >= cnt) continue;
    double v0 = y0[u] - xi0, v1 = y1[u] - xi1, v2 = y2[u] - xi2;
    if (s.periodic) {
      const int code = (int)(e[u] >> kMirrorShift);
      v0 += mirror_shift(s, code, 0); v1 += mirror_shift(s, code, 1); v2 += mirror_shift(s, code, 2);
    }
    const SpPair t = tp[u];
    const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
    if (d2 > t.dij2) continue;
    // contact pair: the mWCA factor (a/dr)^14 amplifies rounding 14-fold, so
    // this formula runs in FP64 (only pairs within D_ij reach it)
    const double dist = sqrt(d2);
    const double dr = dist + a - t.dij;
    const double id = 1.0 / dr, ar = a * id, r2 = ar * ar, r6 = r2 * r2 * r2, r12 = r6 * r6;
    double dv = fabs(12.0 * t.coef * (r12 - r6) * id / dist);
    if (dv > kDeMax) { dv = kDeMax; if (g < (int)(e[u] & kIdxMask)) ++nclamp; }
    F[0] -= dv * v0; F[1] -= dv * v1; F[2] -= dv * v2;
  }
}

// The two largest squared displacements of a bead tile since the last
// rebuild (wave merge of top-2 pairs; -1 marks lanes without a bead), for
// the rebuild decision of the next step (system_trigger).
__device__ inline void tile_top2(const DevPtrs &P, int tile, double d2) {
  double a = d2, b = -1.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double a2 = __shfl_xor(a, off), b2 = __shfl_xor(b, off);
    const double hi = fmax(a, a2), lo = fmax(fmin(a, a2), fmax(b, b2));
    a = hi; b = lo;
  }
  if ((threadIdx.x & (kTile - 1)) == 0) { P.ttop[2 * tile] = a; P.ttop[2 * tile + 1] = b; }
}

// dxyz_mp of time_integral.F90:160: |dxyz|^2 before the update (as
// :64 computes it), then dxyz += dx; the update is kept for a rebuild.
__device__ inline double accumulate_displacement(const DevPtrs &P, int g, const double o[3], const double dx[3]) {
#pragma clang fp contract(off)
  const int n = P.ntot;
  const double o0 = o[0], o1 = o[1], o2 = o[2];
  const double d2 = o0 * o0 + o1 * o1 + o2 * o2;
  P.dxyz[g] = o0 + dx[0]; P.dxyz[n + g] = o1 + dx[1]; P.dxyz[2 * n + g] = o2 + dx[2];
  P.dlast[g] = dx[0]; P.dlast[n + g] = dx[1]; P.dlast[2 * n + g] = dx[2];
  return d2;
}

// Electrostatic pair sums of the force pass (the chunk-pair task slots that
// hold the charge, in task order, four tasks' loads in flight at a time),
// then the Ewald reciprocal partial forces in chunk order, added to F.
__device__ inline void ele_sums(const DevPtrs &P, const DSys &s, int g, double F[3]) {
  const int q = P.qidx[g];
  if (q < 0) return;
  const int lq = q - s.qoff, C = lq / kTile, lc = lq % kTile;
  for (int I0 = 0; I0 < s.nqt; I0 += 4) {
    double f[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int I = min(I0 + u, s.nqt - 1);
      const int a = I <= C ? I : C, bb = I <= C ? C : I;
      const int slot = I < C ? 1 : 0;
      const double *pp = P.pp + s.pp_off + ele_task_index(a, bb, s.nqt) * (6 * kTile) + slot * 3 * kTile + lc;
      f[u][0] = pp[0]; f[u][1] = pp[kTile]; f[u][2] = pp[2 * kTile];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (I0 + u < s.nqt) { F[0] += f[u][0]; F[1] += f[u][1]; F[2] += f[u][2]; }
  }
  if (s.ele_form == 2) {
    const double *pf = P.pf + s.pf_off + lq;
    for (int c = 0; c < s.nkc; ++c)
      for (int d = 0; d < 3; ++d) F[d] += pf[((int64_t)c * 3 + d) * s.nq];
  }
}

// The step index of a bead's noise and valence draws in mode (after mode 4
// resolves to 2 or 3 per system).
__device__ inline int resolve_mode(const DSys &s, int mode) { return mode == 4 ? (s.init_accel ? 2 : 3) : mode; }
__device__ inline uint64_t noise_step(const DevPtrs &P, const DSys &s, int32_t t_local, int mode) {
  return *P.stepbase + (uint64_t)t_local + (mode == 2 ? 0u : 1u) - s.step_origin;
}

// Forces of one bead from the two waves' partial sums, then the integrator
// with the noise gs; returns the squared displacement since the last
// rebuild before this step's update (and accumulates the update) when
// rebuilds are displacement-triggered.
__device__ inline double main_tail(const DevPtrs &P, const DSys &s, int g, double F[3], double (*red)[2][kTile],
                                   const BeadState &b, const double gs[3], int lane, double *XN, double *PXN,
                                   int mode) {
  const int n = P.ntot;
  for (int d = 0; d < 3; ++d) F[d] = red[d][0][lane] + red[d][1][lane];
  P.f[g] = F[0]; P.f[n + g] = F[1]; P.f[2 * n + g] = F[2];
  if (mode == 0) return -1.0;
  double dx[3];
  integrate(P, s, g, b, gs, F, XN, PXN, resolve_mode(s, mode), dx);
  if (!P.dyn) return -1.0;
  return accumulate_displacement(P, g, b.o, dx);
}


__global__ __launch_bounds__(kTile * 2) void k_main(DevPtrs P, const double *X, const double *PX,
                                                    double *XN, double *PXN, const Tile *tiles,
                                                    int ntile, int32_t t_local, int mode, int do_dh,
                                                    int count_step) {
  __shared__ double red[3][2][kTile], gsh[3][kTile];
  if (run_stopped(P)) return;
  // steps completed in the current run (the host resumes a stopped run
  // after the last of them)
  if (count_step && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&P.counters[6], 1ull);
  const int lane = threadIdx.x & (kTile - 1), w = threadIdx.x / kTile;
  const Tile tl = tiles[blockIdx.x];
  const DSys s = P.sys[tl.sys];
  // a displacement-triggered rebuild of this step is counted here, in the
  // step's last kernel, so that a step the run resumes from counts once
  if (mode == 1 && P.dyn && tl.start == 0 && threadIdx.x == 0 && P.sys_flag[tl.sys])
    atomicAdd(&P.counters[3], 1ull);
  const int li = tl.start + lane;
  const bool valid = li < s.nmp;
  const int g = s.off + (valid ? li : 0);
  const int n = P.ntot;
  double F[3] = {0.0, 0.0, 0.0};
  if (valid) {
    if (w == 1) {
      // the Langevin noise of the bead (wave 0 integrates after the
      // barrier) and its electrostatic sums
      if (mode != 0) {
        const int m = resolve_mode(s, mode);
        double gs[3];
        gauss3(P.seed, s.key, noise_step(P, s, t_local, m), (uint32_t)li, gs);
        gsh[0][lane] = gs[0]; gsh[1][lane] = gs[1]; gsh[2][lane] = gs[2];
      }
      if (do_dh && s.ele_form >= 0) ele_sums(P, s, g, F);
    } else {
      // bonded slots in incidence order, eight loads in flight at a time
      // (slots past ie re-read slot ie - 1 and are not added)
      // (the next eight slot indices load while these slots are read)
      const int ib = P.inc_off[g], ie = P.inc_off[g + 1];
      int64_t at[8] = {}, atn[8] = {};
      if (ib < ie) {
#pragma unroll
        for (int u = 0; u < 8; ++u) at[u] = P.inc_slot[min(ib + u, ie - 1)];
      }
      for (int k0 = ib; k0 < ie; k0 += 8) {
        float sl[8][3];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const float *q = P.fslot + 3 * at[u];
          sl[u][0] = q[0]; sl[u][1] = q[1]; sl[u][2] = q[2];
        }
        if (k0 + 8 < ie) {
#pragma unroll
          for (int u = 0; u < 8; ++u) atn[u] = P.inc_slot[min(k0 + 8 + u, ie - 1)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (k0 + u < ie) { F[0] += (double)sl[u][0]; F[1] += (double)sl[u][1]; F[2] += (double)sl[u][2]; }
#pragma unroll
        for (int u = 0; u < 8; ++u) at[u] = atn[u];
      }
    }
    // excluded volume (force_exv_dt15.F90:53-115), owner view of each pair;
    // the two waves take alternate groups of four neighbours
    const double *Y = s.periodic ? PX : X;
    const double xi0 = Y[g], xi1 = Y[n + g], xi2 = Y[2 * n + g];
    const SpPair *tab = P.sptab + s.sptab + (int)P.species[g] * s.nspecies;
    const int cnt = min(P.nb_cnt[g], P.nb_cap);
    unsigned nclamp = 0;
    uint32_t e[4] = {0u, 0u, 0u, 0u}, en[4] = {0u, 0u, 0u, 0u};
    if (4 * w < cnt) exv_entries(P, g, 4 * w, cnt, e);
    for (int m0 = 4 * w; m0 < cnt; m0 += 8) {
      exv_group(P, s, Y, g, m0, cnt, xi0, xi1, xi2, tab, F, nclamp, e, en);
#pragma unroll
      for (int u = 0; u < 4; ++u) e[u] = en[u];
    }
    if (nclamp) atomicAdd(&P.counters[0], (unsigned long long)nclamp);
  }
  for (int d = 0; d < 3; ++d) red[d][w][lane] = F[d];
  BeadState bs;
  if (w == 0 && valid && mode != 0) load_state(P, s, g, X, PX, bs);
  __syncthreads();
  if (w != 0) return;
  double d2 = -1.0;   // squared displacement since the last rebuild, before this update
  if (valid) {
    const double gs[3] = {gsh[0][lane], gsh[1][lane], gsh[2][lane]};
    d2 = main_tail(P, s, g, F, red, bs, gs, lane, XN, PXN, mode);
  }
  if (P.dyn && mode != 0) {
    const int it = s.bt0 + tl.start / kTile;
    tile_top2(P, it, d2);
    // the tile's boxes at the new positions, which a displacement-triggered
    // rebuild of the next step culls by (k_lists then runs without k_prep;
    // lanes past the last bead take the tile's first bead, as bbox_tile)
    if (!s.periodic) {
      double p[3] = {0.0, 0.0, 0.0};
      if (valid) {
#pragma unroll
        for (int d = 0; d < 3; ++d) p[d] = XN[d * n + g];   // this lane's own store
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double p0 = __shfl(p[d], 0);
        p[d] = valid ? p[d] : p0;
      }
      tile_boxes(p[0], p[1], p[2], &P.bbox[it], &P.sbox[4 * it]);
    }
  }
}

// ------------------------------------------------------------------ lists
// util_periodic.F90:26-36 on the current coordinates; with with_pred the
// predicted wrapped positions px + (c4 v + dt^2/2 a) are refreshed too.
__device__ inline void wrap_bead(const DevPtrs &P, const DSys &s, const double *X, double *PXc, double *PXp,
                                 int g, int with_pred) {
#pragma clang fp contract(off)
  const int n = P.ntot;
  for (int d = 0; d < 3; ++d) {
    double pw = s.box[d], pmax = 0.5 * pw, pmin = -0.5 * pw;
    double tx = X[d * n + g];
    if (tx > pmax) tx = tx - pw * (double)((int)((tx - pmax) / pw) + 1);
    else if (tx < pmin) tx = tx + pw * (double)((int)((pmin - tx) / pw) + 1);
    PXc[d * n + g] = tx;
    if (with_pred) {
      double dx = P.lcls[s.lc0 + P.lcl[g]].r[3] * P.v[d * n + g] + P.tstep2 * P.a[d * n + g];
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

// Bounding box of bead tile or charge tile `it` (tile tl) from the
// coordinates the lists are built from (non-periodic systems only cull by
// boxes).
__device__ inline void bbox_tile(const DevPtrs &P, const DSys &s, const double *X, const Tile tl, int it,
                                 bool charge) {
  const int lane = threadIdx.x;
  const double *Y = X;
  const int n = P.ntot;
  const int li = tl.start + lane;
  const int lim = charge ? s.nq : s.nmp;
  const int first = charge ? P.qbead[s.qoff + tl.start] : s.off + tl.start;
  int g = first;
  if (li < lim) g = charge ? P.qbead[s.qoff + li] : s.off + li;
  tile_boxes(Y[g], Y[n + g], Y[2 * n + g], charge ? &P.qbbox[it] : &P.bbox[it], charge ? nullptr : &P.sbox[4 * it]);
}

// squared distance bounds between two boxes (FP64); conservative culling
// margins of 1e-9 relative cover the rounding of the per-pair test.
__device__ inline double box_dmin2(const BBox &a, const BBox &b) {
  double s2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    double g = fmax(0.0, fmax(b.lo[d] - a.hi[d], a.lo[d] - b.hi[d]));
    s2 += g * g;
  }
  return s2;
}
__device__ inline double box_dmax2(const BBox &a, const BBox &b) {
  double s2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    double g = fmax(fabs(b.hi[d] - a.lo[d]), fabs(a.hi[d] - b.lo[d]));
    s2 += g * g;
  }
  return s2;
}

// exv lists (neighbor_list.F90:56-87 + neighbor_assign.F90): every bead of
// the tile scans the system's bead tiles that its tile's box can reach,
// staged through LDS, with the FP64 test d^2 < (cutoff + margin)^2 rounded as
// the reference rounds it.
constexpr int kRl2Lds = 16;       // unit-pair radius tables up to 16 x 16 units are staged in LDS

struct ListLds {
  double sx[kTile], sy[kTile], sz[kTile];
  int su[kTile], stc[kTile];
  BBox sb[kTile / 16];                 // boxes of the staged tile's 16-bead groups
  double rl2[kRl2Lds * kRl2Lds];
};

// The pair vector x_hi - x_lo of the reference's lower-index-first loops
// from x_j - x_i (i: this lane's bead or charge): IEEE subtraction is
// antisymmetric, so a negation gives the bits of the reversed difference;
// non-periodic pairs need only |v|^2, which the sign does not change.
__device__ inline void pair_vec(const DSys &s, double xj0, double xj1, double xj2, const double xi[3], bool own_lo,
                                double v[3]) {
  v[0] = __dadd_rn(xj0, -xi[0]); v[1] = __dadd_rn(xj1, -xi[1]); v[2] = __dadd_rn(xj2, -xi[2]);
  if (s.periodic && !own_lo) { v[0] = -v[0]; v[1] = -v[1]; v[2] = -v[2]; }
}

// squared distance from a point to a box (FP64)
__device__ inline double point_box_d2(const double p[3], const BBox &b) {
  double s2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    double g = fmax(0.0, fmax(b.lo[d] - p[d], p[d] - b.hi[d]));
    s2 += g * g;
  }
  return s2;
}

// One wave per bead tile sweeps the system's bead tiles in order; tiles and
// 16-bead groups whose boxes no bead of the tile can reach within the
// largest list radius are skipped.  The tile boxes are tested 64 at a time
// (one lane per tile), the group boxes are staged with the tile, and the
// unit-pair radii sit in LDS, so neither the culling nor the per-pair tests
// wait on dependent global loads.
// One candidate pair of the exv list (neighbor_list.F90 distance test with
// the unit-pair radius, neighbor_assign.F90 exclusions): this lane's bead
// li against bead lj2 of the same system (staged coordinates, unit,
// type | class << 8).  On true, entry = mirror << 27 | global index.
template <bool RL2_LDS>
__device__ inline bool exv_candidate(const DevPtrs &P, const DSys &s, const double xi[3], int li, int ui, int ti,
                                     int ci, int lj2, double xj0, double xj1, double xj2, int uj, int tcj,
                                     const double *rl2l, uint32_t &entry) {
  if (lj2 == li) return false;
  const bool own_lo = li < lj2;
  double v[3];
  pair_vec(s, xj0, xj1, xj2, xi, own_lo, v);
  const int code = s.periodic ? pbneighbor(s, v) : 0;
  const double d2 = sq3_exact(v[0], v[1], v[2]);
  if (!(d2 < s.rl2max)) return false;   // rl2max bounds every unit-pair radius
  const int ulo = own_lo ? ui : uj, uhi = own_lo ? uj : ui;
  const double r2 = RL2_LDS ? rl2l[ulo * s.nunit + uhi] : P.rl2[s.uoff + ulo * s.nunit + uhi];
  if (!(d2 < r2)) return false;
  const int tj = tcj & 0xFF, cj = tcj >> 8;
  const bool keep = own_lo ? exv_keep(s, ui, uj, ci, ti, li, lj2) : exv_keep(s, uj, ui, cj, tj, lj2, li);
  if (!keep) return false;
  const int oc = own_lo ? code : mirror_flip(code);
  entry = ((uint32_t)oc << kMirrorShift) | (uint32_t)(s.off + lj2);
  return true;
}

template <bool RL2_LDS, bool ONE_UNIT>
__device__ inline void list_exv_tile_t(const DevPtrs &P, const DSys &s, const double *X, const double *PX,
                                       const Tile tl, int it, ListLds &L) {
  auto &sx = L.sx; auto &sy = L.sy; auto &sz = L.sz; auto &su = L.su; auto &stc = L.stc;
  const int lane = threadIdx.x;
  const int li = tl.start + lane;
  const bool valid = li < s.nmp;
  const int n = P.ntot;
  const double *Y = s.periodic ? PX : X;
  const int g = s.off + (valid ? li : 0);
  if (RL2_LDS)
    for (int k = lane; k < s.nunit * s.nunit; k += kTile) L.rl2[k] = P.rl2[s.uoff + k];
  double xi[3] = {0, 0, 0};
  int ui = 0, ti = 0, ci = 0;
  if (valid) { ld3(Y, n, g, xi); ui = P.unit[g]; ti = P.mptype[g]; ci = P.iclass[g]; }
  const BBox bi = P.bbox[it];
  const double cut2 = s.rl2max * (1.0 + 1e-9);
  int cnt = 0;
  for (int j0 = 0; j0 < s.nbt; j0 += kTile) {
    // tiles j0 .. j0 + 63 that a bead of this tile may reach, tested at once
    const int jl = j0 + lane;
    const bool reach = jl < s.nbt && (s.periodic || !(box_dmin2(bi, P.bbox[s.bt0 + jl]) > cut2));
    uint64_t todo = __ballot(reach);
    while (todo) {
      const int jtile = j0 + __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const int jt = jtile * kTile;
      int lj = jt + lane;
      if (lj < s.nmp) {
        int gj = s.off + lj;
        sx[lane] = Y[gj]; sy[lane] = Y[n + gj]; sz[lane] = Y[2 * n + gj];
        su[lane] = P.unit[gj];
        stc[lane] = (int)P.mptype[gj] | ((int)P.iclass[gj] << 8);
      }
      if (!s.periodic && lane < kTile / 16) L.sb[lane] = P.sbox[4 * (s.bt0 + jtile) + lane];
      __syncthreads();
      const int kend = min(kTile, s.nmp - jt);
      for (int q0 = 0; q0 < kend; q0 += 16) {
        if (!s.periodic) {
          const bool need = valid && point_box_d2(xi, L.sb[q0 / 16]) <= cut2;
          if (!__any(need)) continue;
        }
        const int q1 = min(q0 + 16, kend);
        if (valid) {
          // exv_candidate for a non-periodic system, every test evaluated
          // and combined (no branch per test), the group's staged values
          // read together
#pragma unroll
          for (int dk = 0; dk < 16; ++dk) {
            const int k = q0 + dk;
            if (k >= q1) continue;   // (the same for every lane)
            const int lj2 = jt + k;
            const double d2 = sq3_separate(sx[k] - xi[0], sy[k] - xi[1], sz[k] - xi[2]);
            const bool own_lo = li < lj2;
            // one unit: its radius is rl2max and both beads are in it
            const int uj = ONE_UNIT ? ui : su[k], tcj = stc[k];
            const int ulo = own_lo ? ui : uj, uhi = own_lo ? uj : ui;
            const double r2 = ONE_UNIT ? s.rl2max
                                       : (RL2_LDS ? L.rl2[ulo * s.nunit + uhi] : P.rl2[s.uoff + ulo * s.nunit + uhi]);
            const int tlo = own_lo ? ti : (tcj & 0xFF), clo = own_lo ? ci : (tcj >> 8);
            const int lo = own_lo ? li : lj2, hi = own_lo ? lj2 : li;
            int sep = -1;   // exv_keep's separation by the lower bead's type, as selects
            sep = tlo == 2 ? s.nsep[0] : sep;
            sep = tlo == 4 ? s.nsep[1] : sep;
            sep = tlo == 3 ? s.nsep[2] : sep;
            const bool keep = (ulo != uhi) | (clo != 2) | ((sep >= 0) & (hi >= lo + sep));
            if ((lj2 != li) & (d2 < s.rl2max) & (d2 < r2) & keep) {
              if (cnt < P.nb_cap) P.nb[(size_t)cnt * n + g] = (uint32_t)(s.off + lj2);
              ++cnt;
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (valid) {
    P.nb_cnt[g] = cnt;
    if (cnt > P.nb_cap) atomicMax(&P.counters[5], (unsigned long long)cnt);
  }
}

// the unit-pair radii from LDS when the table fits there (a compile-time
// choice: a run-time one makes every radius read a generic-pointer load).
// Non-periodic systems only (k_lpair builds the periodic ones).
__device__ inline void list_exv_tile(const DevPtrs &P, const DSys &s, const double *X, const double *PX,
                                     const Tile tl, int it, ListLds &L) {
  if (s.nunit == 1) list_exv_tile_t<true, true>(P, s, X, PX, tl, it, L);
  else if (s.nunit <= kRl2Lds) list_exv_tile_t<true, false>(P, s, X, PX, tl, it, L);
  else list_exv_tile_t<false, false>(P, s, X, PX, tl, it, L);
}

// ele bit matrix (neighbor_list_ele.F90:123-210): bit (qi, qj) set when the
// pair is in the list; radius (margin + cutoff)^2 per i_DH_cutoff_type.  Tile
// pairs whose boxes lie entirely inside (outside) the radius are filled with
// ones (zeros) without per-pair tests; the rest are tested pair by pair.
// A box of charge tile t (charges [64 t, 64 t + 64) of system s): the union
// of the boxes of the bead tiles that hold them (charges follow bead order).
__device__ inline BBox charge_tile_union_box(const DevPtrs &P, const DSys &s, int t) {
  const int q0 = t * kTile, q1 = min(q0 + kTile, s.nq) - 1;
  const int t0 = (P.qbead[s.qoff + q0] - s.off) / kTile, t1 = (P.qbead[s.qoff + q1] - s.off) / kTile;
  BBox b = P.bbox[s.bt0 + t0];
  for (int k = t0 + 1; k <= t1; ++k) {
    const BBox c = P.bbox[s.bt0 + k];
    for (int d = 0; d < 3; ++d) { b.lo[d] = fmin(b.lo[d], c.lo[d]); b.hi[d] = fmax(b.hi[d], c.hi[d]); }
  }
  return b;
}

// (union_boxes: the charge tiles' boxes from the bead tiles' boxes instead of
// qbbox, which only k_prep computes; conservative, so the decisions below
// stay exact)
__device__ inline void list_ele_tile(const DevPtrs &P, const DSys &s, const double *X, const double *PX,
                                     const Tile tl, int it, ListLds &L, bool union_boxes) {
  auto &sx = L.sx; auto &sy = L.sy; auto &sz = L.sz;
  const int lane = threadIdx.x;
  const int qi = tl.start + lane;
  const bool valid = qi < s.nq;
  const int n = P.ntot;
  const double *Y = s.periodic ? PX : X;
  double xi[3] = {0, 0, 0};
  if (valid) ld3(Y, n, P.qbead[s.qoff + qi], xi);
  uint32_t *row = P.qmask + s.moff + (int64_t)(valid ? qi : 0) * s.mwords;
  const BBox bi = union_boxes ? charge_tile_union_box(P, s, tl.start / kTile) : P.qbbox[it];
  for (int j0 = 0; j0 < s.nqt; j0 += kTile) {
    // mode of the tile pairs (it, j0 + lane), decided at once: 0 none, 1 all,
    // 2 test
    int lmode = 2;
    if (!s.periodic && j0 + lane < s.nqt) {
      const BBox bj = union_boxes ? charge_tile_union_box(P, s, j0 + lane) : P.qbbox[s.qt0 + j0 + lane];
      if (box_dmin2(bi, bj) > s.ele_rl2 * (1.0 + 1e-9)) lmode = 0;
      else if (box_dmax2(bi, bj) < s.ele_rl2 * (1.0 - 1e-9)) lmode = 1;
    }
    for (int jtile = j0; jtile < min(s.nqt, j0 + kTile); ++jtile) {
      const int jt = jtile * kTile;
      const int kend = min(kTile, s.nq - jt);
      const int mode = __shfl(lmode, jtile - j0);
      uint32_t w[2] = {0u, 0u};
      if (mode == 1) {
        w[0] = kend >= 32 ? 0xFFFFFFFFu : ((1u << kend) - 1u);
        w[1] = kend >= 64 ? 0xFFFFFFFFu : (kend > 32 ? ((1u << (kend - 32)) - 1u) : 0u);
        if (qi >= jt && qi < jt + kend) w[(qi - jt) >> 5] &= ~(1u << ((qi - jt) & 31));
      } else if (mode == 2) {
        int qj = jt + lane;
        if (qj < s.nq) {
          int gj = P.qbead[s.qoff + qj];
          sx[lane] = Y[gj]; sy[lane] = Y[n + gj]; sz[lane] = Y[2 * n + gj];
        }
        __syncthreads();
        if (valid)
          for (int k = 0; k < kend; ++k) {
            int qj2 = jt + k;
            if (qj2 == qi) continue;
            const bool own_lo = qi < qj2;
            double v[3];
            pair_vec(s, sx[k], sy[k], sz[k], xi, own_lo, v);
            int code = s.periodic ? pbneighbor(s, v) : 0;
            double d2 = sq3_exact(v[0], v[1], v[2]);
            if (!(d2 < s.ele_rl2)) continue;
            w[k >> 5] |= 1u << (k & 31);
            if (s.periodic) P.qmirror[s.mioff + (int64_t)qi * s.nq + qj2] = (uint8_t)(own_lo ? code : mirror_flip(code));
          }
        __syncthreads();
      }
      if (valid) {
        row[jt >> 5] = w[0];
        if ((jt >> 5) + 1 < s.mwords) row[(jt >> 5) + 1] = w[1];
      }
    }
  }
}

// hb candidate list (neighbor_list_hb.F90:59-78; radius mloop_dtrna.F90:149)
__device__ inline void list_hb_entry(const DevPtrs &P, const DSys &s, const double *X, const double *PX, int h) {
#pragma clang fp contract(off)
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


// ------------------------------------------------------------------ lists by tile pairs
// Periodic systems (no box culling, so one wave per tile would sweep every
// tile of the box in turn): every (bead tile, bead tile) and (charge tile,
// charge tile) pair is its own wave.  The excluded-volume lists take two passes: pass 0 counts each
// bead's entries per tile pair, pass 1 writes them after the counts of the
// earlier tile pairs of its row, so each list keeps the tile order of
// list_exv_tile.  mode bit 1 takes only the systems whose trigger fired,
// bit 2 only the periodic systems.

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// exv entries of bead tile ia (local index) against bead tile jb: counts
// them (write = false) or writes them from slot `off` on; returns the count
template <bool RL2_LDS>
__device__ inline int exv_pair_scan_t(const DevPtrs &P, const DSys &s, const double *X, const double *PX, int ia,
                                      int jb, bool write, int off, ListLds &L) {
  const int lane = threadIdx.x & (kTile - 1);
  const int li = ia * kTile + lane;
  const bool valid = li < s.nmp;
  const int n = P.ntot;
  const double *Y = s.periodic ? PX : X;
  const int g = s.off + (valid ? li : 0);
  const double cut2 = s.rl2max * (1.0 + 1e-9);
  if (!s.periodic && box_dmin2(P.bbox[s.bt0 + ia], P.bbox[s.bt0 + jb]) > cut2) return 0;
  if (RL2_LDS)
    for (int k = lane; k < s.nunit * s.nunit; k += kTile) L.rl2[k] = P.rl2[s.uoff + k];
  double xi[3] = {0, 0, 0};
  int ui = 0, ti = 0, ci = 0;
  if (valid) { ld3(Y, n, g, xi); ui = P.unit[g]; ti = P.mptype[g]; ci = P.iclass[g]; }
  const int jt = jb * kTile;
  const int lj = jt + lane;
  if (lj < s.nmp) {
    const int gj = s.off + lj;
    L.sx[lane] = Y[gj]; L.sy[lane] = Y[n + gj]; L.sz[lane] = Y[2 * n + gj];
    L.su[lane] = P.unit[gj];
    L.stc[lane] = (int)P.mptype[gj] | ((int)P.iclass[gj] << 8);
  }
  wave_lds_sync();
  int cnt = 0;
  const int kend = min(kTile, s.nmp - jt);
  for (int q0 = 0; q0 < kend; q0 += 16) {
    if (!s.periodic) {
      const bool need = valid && point_box_d2(xi, P.sbox[4 * (s.bt0 + jb) + q0 / 16]) <= cut2;
      if (!__any(need)) continue;
    }
    const int q1 = min(q0 + 16, kend);
    if (valid)
      for (int k = q0; k < q1; ++k) {
        const int lj2 = jt + k;
        if (lj2 == li) continue;
        const bool own_lo = li < lj2;
        double v[3];
        pair_vec(s, L.sx[k], L.sy[k], L.sz[k], xi, own_lo, v);
        const int code = s.periodic ? pbneighbor(s, v) : 0;
        const double d2 = sq3_exact(v[0], v[1], v[2]);
        if (!(d2 < s.rl2max)) continue;
        const int uj = L.su[k];
        const int ulo = own_lo ? ui : uj, uhi = own_lo ? uj : ui;
        const double r2 = RL2_LDS ? L.rl2[ulo * s.nunit + uhi] : P.rl2[s.uoff + ulo * s.nunit + uhi];
        if (!(d2 < r2)) continue;
        const int tj = L.stc[k] & 0xFF, cj = L.stc[k] >> 8;
        const bool keep = own_lo ? exv_keep(s, ui, uj, ci, ti, li, lj2) : exv_keep(s, uj, ui, cj, tj, lj2, li);
        if (!keep) continue;
        if (write) {
          const int slot = off + cnt;
          const int oc = own_lo ? code : mirror_flip(code);
          if (slot < P.nb_cap) P.nb[(size_t)slot * n + g] = ((uint32_t)oc << kMirrorShift) | (uint32_t)(s.off + lj2);
        }
        ++cnt;
      }
  }
  wave_lds_sync();
  return cnt;
}

__device__ inline int exv_pair_scan(const DevPtrs &P, const DSys &s, const double *X, const double *PX, int ia,
                                    int jb, bool write, int off, ListLds &L) {
  return s.nunit <= kRl2Lds ? exv_pair_scan_t<true>(P, s, X, PX, ia, jb, write, off, L)
                            : exv_pair_scan_t<false>(P, s, X, PX, ia, jb, write, off, L);
}

// ele bit words of charge tile I's rows for charge tile J (and, periodic,
// the mirror codes), as list_ele_tile does for one J
__device__ inline void ele_pair_rows(const DevPtrs &P, const DSys &s, const double *X, const double *PX, int I, int J,
                                     ListLds &L) {
  const int lane = threadIdx.x & (kTile - 1);
  const int qi = I * kTile + lane;
  const bool valid = qi < s.nq;
  const int n = P.ntot;
  const double *Y = s.periodic ? PX : X;
  double xi[3] = {0, 0, 0};
  if (valid) ld3(Y, n, P.qbead[s.qoff + qi], xi);
  uint32_t *row = P.qmask + s.moff + (int64_t)(valid ? qi : 0) * s.mwords;
  const int jt = J * kTile;
  const int kend = min(kTile, s.nq - jt);
  int mode = 2;  // 0 none, 1 all, 2 test
  if (!s.periodic) {
    const BBox bi = P.qbbox[s.qt0 + I], bj = P.qbbox[s.qt0 + J];
    if (box_dmin2(bi, bj) > s.ele_rl2 * (1.0 + 1e-9)) mode = 0;
    else if (box_dmax2(bi, bj) < s.ele_rl2 * (1.0 - 1e-9)) mode = 1;
  }
  uint32_t w[2] = {0u, 0u};
  if (mode == 1) {
    w[0] = kend >= 32 ? 0xFFFFFFFFu : ((1u << kend) - 1u);
    w[1] = kend >= 64 ? 0xFFFFFFFFu : (kend > 32 ? ((1u << (kend - 32)) - 1u) : 0u);
    if (qi >= jt && qi < jt + kend) w[(qi - jt) >> 5] &= ~(1u << ((qi - jt) & 31));
  } else if (mode == 2) {
    const int qj = jt + lane;
    if (qj < s.nq) {
      const int gj = P.qbead[s.qoff + qj];
      L.sx[lane] = Y[gj]; L.sy[lane] = Y[n + gj]; L.sz[lane] = Y[2 * n + gj];
    }
    wave_lds_sync();
    if (valid)
      for (int k = 0; k < kend; ++k) {
        const int qj2 = jt + k;
        if (qj2 == qi) continue;
        const bool own_lo = qi < qj2;
        double v[3];
        pair_vec(s, L.sx[k], L.sy[k], L.sz[k], xi, own_lo, v);
        const int code = s.periodic ? pbneighbor(s, v) : 0;
        const double d2 = sq3_exact(v[0], v[1], v[2]);
        if (!(d2 < s.ele_rl2)) continue;
        w[k >> 5] |= 1u << (k & 31);
        if (s.periodic) P.qmirror[s.mioff + (int64_t)qi * s.nq + qj2] = (uint8_t)(own_lo ? code : mirror_flip(code));
      }
    wave_lds_sync();
  }
  if (valid) {
    row[jt >> 5] = w[0];
    if ((jt >> 5) + 1 < s.mwords) row[(jt >> 5) + 1] = w[1];
  }
}

__device__ inline bool lpair_take(const DevPtrs &P, int sys, int mode) {
  return (!(mode & 1) || P.sys_flag[sys] != 0) && (!(mode & 2) || P.sys[sys].periodic != 0);
}

// pass 0: exv counts of bead tile pairs, ele rows of charge tile pairs and
// hbond candidates (nh > 0); pass 1: exv entries.  One pair per wave.
__global__ __launch_bounds__(kTile * kWaves) void k_lpair(DevPtrs P, const double *X, const double *PX,
                                                          const EwTask *bpairs, int nbp, const EwTask *qpairs,
                                                          int nqp, int h0, int nh, int mode, int pass) {
  __shared__ ListLds lds[kWaves];
  if (run_stopped(P)) return;
  const int w = (int)threadIdx.x / kTile, lane = (int)threadIdx.x & (kTile - 1);
  const int nbb = (nbp + kWaves - 1) / kWaves, nqb = (nqp + kWaves - 1) / kWaves;
  const int b = (int)blockIdx.x;
  if (b < nbb) {
    const int p = b * kWaves + w;
    if (p >= nbp) return;
    const EwTask bp = bpairs[p];
    if (!lpair_take(P, bp.sys, mode)) return;
    const DSys s = P.sys[bp.sys];
    int *cnt = P.lcnt + (size_t)(s.lp_off + bp.a * s.nbt) * kTile;   // counts of this row of tile pairs
    if (pass == 0) {
      cnt[bp.b * kTile + lane] = exv_pair_scan(P, s, X, PX, bp.a, bp.b, false, 0, lds[w]);
      return;
    }
    int off = 0;
    for (int j = 0; j < bp.b; ++j) off += cnt[j * kTile + lane];
    const int c = exv_pair_scan(P, s, X, PX, bp.a, bp.b, true, off, lds[w]);
    const int li = bp.a * kTile + lane;
    if (bp.b == s.nbt - 1 && li < s.nmp) {
      const int tot = off + c, g = s.off + li;
      P.nb_cnt[g] = tot;
      if (tot > P.nb_cap) atomicMax(&P.counters[5], (unsigned long long)tot);
    }
    return;
  }
  if (pass == 1) return;
  if (b < nbb + nqb) {
    const int p = (b - nbb) * kWaves + w;
    if (p >= nqp) return;
    const EwTask qp = qpairs[p];
    if (!lpair_take(P, qp.sys, mode)) return;
    const DSys s = P.sys[qp.sys];
    ele_pair_rows(P, s, X, PX, qp.a, qp.b, lds[w]);
    return;
  }
  const int k = (b - nbb - nqb) * (kTile * kWaves) + (int)threadIdx.x;
  if (k >= nh) return;
  const int h = h0 + k;
  if (!lpair_take(P, P.hb_sys[h], mode)) return;
  list_hb_entry(P, P.sys[P.hb_sys[h]], X, PX, h);
}

// The list passes of one rebuild in two launches.  gate = 0 builds for every
// system of the launch range; gate = 1 only for the systems whose
// displacement trigger fires (decided in k_prep from the tile pairs the
// previous k_main wrote, recorded in sys_flag for k_lists); all other blocks
// return at once.
//
// k_prep: blocks [0, nbt) bead tiles (periodic: wrap into the box and refresh
// the predicted wrapped positions; else the tile's bounding box; with gate,
// dxyz restarts from the last step's displacement), blocks [nbt, nbt + nqt)
// charge-tile boxes.
// time_integral.F90:57-75: the two largest squared displacements of the
// system (merged over its tiles' pairs; the merge is exact in any order) and
// sqrt(m1) + sqrt(m2) > margin.
__device__ inline bool system_trigger(const DevPtrs &P, const DSys &s) {
  double a = 0.0, b = 0.0;
  for (int t = (int)threadIdx.x; t < s.nbt; t += kTile) {
    const double t1 = P.ttop[2 * (s.bt0 + t)], t2 = P.ttop[2 * (s.bt0 + t) + 1];
    const double hi = fmax(a, t1), lo = fmax(fmin(a, t1), fmax(b, t2));
    a = hi; b = lo;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double a2 = __shfl_xor(a, off), b2 = __shfl_xor(b, off);
    const double hi = fmax(a, a2), lo = fmax(fmin(a, a2), fmax(b, b2));
    a = hi; b = lo;
  }
  return sqrt(a) + sqrt(b) > s.margin;
}

__global__ __launch_bounds__(kTile) void k_prep(DevPtrs P, const double *X, double *PXc, double *PXp,
                                                const Tile *btiles, int bt0, int nbt, const Tile *qtiles, int qt0,
                                                int nqt, int with_pred, int gate) {
  if (run_stopped(P)) return;
  const int b = (int)blockIdx.x;
  const bool charge = b >= nbt;
  const int it = charge ? qt0 + b - nbt : bt0 + b;
  const Tile tl = charge ? qtiles[it] : btiles[it];
  const DSys s = P.sys[tl.sys];
  if (gate) {
    // every block of the system reaches the same decision; the first bead
    // tile records it for k_lists and k_main (which counts the rebuild)
    const bool fire = system_trigger(P, s);
    if (!charge && tl.start == 0 && threadIdx.x == 0) P.sys_flag[tl.sys] = fire ? 1 : 0;
    if (!fire) return;
  }
  if (!charge) {
    const int li = tl.start + (int)threadIdx.x;
    if (li < s.nmp) {
      const int g = s.off + li, n = P.ntot;
      if (s.periodic) wrap_bead(P, s, X, PXc, PXp, g, with_pred);
      if (gate)
        for (int d = 0; d < 3; ++d) P.dxyz[d * n + g] = P.dlast[d * n + g];
    }
  }
  if (!s.periodic) bbox_tile(P, s, X, tl, it, charge);
}

// k_lists: one wave per block; [0, nbt) exv lists of bead tiles, [nbt,
// nbt + nqt) ele bit rows of charge tiles, then 64 hbond candidates per
// block.
__global__ __launch_bounds__(kTile) void k_lists(DevPtrs P, const double *X, const double *PX, const Tile *btiles,
                                                 int bt0, int nbt, const Tile *qtiles, int qt0, int nqt, int h0,
                                                 int nh, int gate) {
  __shared__ ListLds lds;
  if (run_stopped(P)) return;
  const int b = (int)blockIdx.x;
  if (b < nbt + nqt) {
    const bool charge = b >= nbt;
    const int it = charge ? qt0 + b - nbt : bt0 + b;
    const Tile tl = charge ? qtiles[it] : btiles[it];
    if (gate == 1 && !P.sys_flag[tl.sys]) return;
    const DSys s = P.sys[tl.sys];
    if (gate == 2) {
      // without k_prep (non-periodic batch): k_prep's decision, taken by
      // every block of the system, and its displacement reset; the bead
      // tiles' boxes are k_main's of this step's coordinates
      const bool fire = system_trigger(P, s);
      if (!charge && tl.start == 0 && threadIdx.x == 0) P.sys_flag[tl.sys] = fire ? 1 : 0;
      if (!fire) return;
      if (!charge) {
        const int li = tl.start + (int)threadIdx.x, n = P.ntot;
        if (li < s.nmp)
          for (int d = 0; d < 3; ++d) P.dxyz[d * n + s.off + li] = P.dlast[d * n + s.off + li];
        // the system's hbond candidates, shared out over its bead tiles
        for (int k = tl.start + (int)threadIdx.x; k < s.nhb; k += s.nbt * kTile) list_hb_entry(P, s, X, PX, s.hoff + k);
      }
    }
    if (s.periodic) return;   // built by tile pairs (k_lpair)
    if (charge) list_ele_tile(P, s, X, PX, tl, it, lds, gate == 2);
    else list_exv_tile(P, s, X, PX, tl, it, lds);
    return;
  }
  const int k = (b - nbt - nqt) * kTile + (int)threadIdx.x;
  if (k >= nh) return;
  const int h = h0 + k;
  if (gate && !P.sys_flag[P.hb_sys[h]]) return;
  list_hb_entry(P, P.sys[P.hb_sys[h]], X, PX, h);
}

// st_status of the consecutive stacks [s0, s0 + ns) at the given
// coordinates (force_dtrna_stack_nlocal.F90:83-92 via stack_on), for output.
__global__ __launch_bounds__(256) void k_stack_status(DevPtrs P, const double *X, const double *PX, int s0, int ns,
                                                      int32_t *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ns) out[k] = stack_on(P, X, PX, s0 + k) ? 1 : 0;
}

// The step counter: by steps at once after a captured window, one step per
// eager step.  A stopped run does not advance it (the host reads it to know
// where to resume).
__global__ void k_advance(uint64_t *stepbase, uint64_t by, unsigned long long *counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && counters[5] == 0ull) { *stepbase += by; counters[4] += by; }
}

// One step of a captured window: the step counter moves per step, so a stop
// inside a window leaves it at the step to resume from.
__global__ void k_advance_step(uint64_t *stepbase, unsigned long long *counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && counters[5] == 0ull) { *stepbase += 1ull; counters[4] += 1ull; }
}

__global__ void k_count_rebuild(unsigned long long *counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && counters[5] == 0ull) counters[3] += 1;
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

// unit_mode 1: block (x, y) sums only the terms whose unit pair, taken as
// energy_sumup.F90's terms take it, is (y mod nunit, y / nunit): raw
// energy_unit entries (the Ewald terms add none, as in
// energy_ele_coulomb_ewld.F90:168-170).
__global__ __launch_bounds__(kEBlock) void k_energy(DevPtrs P, const double *X, const double *PX,
                                                    const int32_t *sysidx, double *out, int unit_mode) {
  __shared__ double red[kEBlock];
  const int t = threadIdx.x;
  const int is = sysidx[blockIdx.x];
  const DSys s = P.sys[is];
  const int n = P.ntot;
  const int ua = unit_mode ? (int)blockIdx.y % s.nunit : 0, ub = unit_mode ? (int)blockIdx.y / s.nunit : 0;
  auto take = [&](int gi, int gj) { return !unit_mode || (P.unit[gi] == ua && P.unit[gj] == ub); };
  double e[kNE];
  for (int k = 0; k < kNE; ++k) e[k] = 0.0;
  // bonds (energy_bond.F90:39-57)
  for (int b = s.boff + t; b < s.boff + s.nbd; b += kEBlock) {
    if (!take(P.bd_mp[2 * b], P.bd_mp[2 * b + 1])) continue;
    double v[3];
    vsub(X, n, P.bd_mp[2 * b + 1], P.bd_mp[2 * b], v);
    double dd = sqrt(dot3(v, v)) - P.bd_nat[b];
    e[0] += P.bd_coef[b] * (dd * dd);
  }
  // angles (energy_bangle.F90:36-52)
  for (int a = s.aoff + t; a < s.aoff + s.nba; a += kEBlock) {
    if (!take(P.ba_mp[3 * a], P.ba_mp[3 * a + 1])) continue;
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
    if (!take(P.st_mp[7 * st], P.st_mp[7 * st])) continue;
    if (!stack_on(P, X, PX, st)) continue;
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
    if (!take(P.tst_mp[6 * k], P.tst_mp[6 * k + 1])) continue;
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
    if (P.hb_inlist[h] && P.hb_status[h] && take(P.hb_mp[6 * h], P.hb_mp[6 * h + 1]))
      e[P.hb_tert[h] ? 5 : 4] += P.hb_ene[h];
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
        if (j < g || !take(g, j)) continue;
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
      if (unit_mode && P.unit[P.qbead[s.qoff + qi]] != ua) continue;
      int gi = P.qbead[s.qoff + qi];
      double xi[3];
      ld3(Y, n, gi, xi);
      const uint32_t *row = P.qmask + s.moff + (int64_t)qi * s.mwords;
      for (int qj = qi + 1; qj < s.nq; ++qj) {
        if (!((row[qj >> 5] >> (qj & 31)) & 1u)) continue;
        int gj = P.qbead[s.qoff + qj];
        if (unit_mode && P.unit[gj] != ub) continue;
        double v[3] = {Y[gj] - xi[0], Y[n + gj] - xi[1], Y[2 * n + gj] - xi[2]};
        if (s.periodic) { int c = P.qmirror[s.mioff + (int64_t)qi * s.nq + qj]; for (int d = 0; d < 3; ++d) v[d] += mirror_shift(s, c, d); }
        double d2 = dot3(v, v);
        if (d2 > s.dh_ecut2) continue;
        double d1 = sqrt(d2);
        e[7] += (P.charge[gi] * P.charge[gj]) * s.ele_coef / d1 * exp(-d1 * s.rcdist);
      }
    }
  }
  // plain Coulomb (energy_ele_coulomb.F90:55-66) and the Ewald real-space
  // term (energy_ele_coulomb_ewld.F90:126-138), each pair once
  if (s.ele_form == 1 || (s.ele_form == 2 && !unit_mode)) {
    const double *Y = s.periodic ? PX : X;
    for (int qi = t; qi < s.nq; qi += kEBlock) {
      if (unit_mode && P.unit[P.qbead[s.qoff + qi]] != ua) continue;
      int gi = P.qbead[s.qoff + qi];
      double xi[3];
      ld3(Y, n, gi, xi);
      const uint32_t *row = P.qmask + s.moff + (int64_t)qi * s.mwords;
      for (int qj = qi + 1; qj < s.nq; ++qj) {
        if (!((row[qj >> 5] >> (qj & 31)) & 1u)) continue;
        int gj = P.qbead[s.qoff + qj];
        if (unit_mode && P.unit[gj] != ub) continue;
        double v[3] = {Y[gj] - xi[0], Y[n + gj] - xi[1], Y[2 * n + gj] - xi[2]};
        if (s.periodic) { int c = P.qmirror[s.mioff + (int64_t)qi * s.nq + qj]; for (int d = 0; d < 3; ++d) v[d] += mirror_shift(s, c, d); }
        double d2 = dot3(v, v);
        if (d2 > s.cl_cut2) continue;
        double d1 = sqrt(d2);
        const double cij = (P.charge[gi] * P.charge[gj]) * s.ele_coef;
        e[7] += s.ele_form == 1 ? cij * (1.0 / d1 - s.cl_rcinv) : cij * erfc(s.ew_alpha * d1) / d1;
      }
    }
  }
  // Ewald Fourier sum (energy_ele_coulomb_ewld.F90:70-90) from the partial
  // structure factors of the last k_ewald_sk pass, plus the self term
  if (s.ele_form == 2 && !unit_mode) {
    for (int k = t; k < s.nk; k += kEBlock) {
      double sc = 0.0, ss = 0.0;
      for (int c = 0; c < s.nqc; ++c) {
        const double *ps = P.psk + s.psk_off + 2 * ((int64_t)c * s.nk + k);
        sc += ps[0]; ss += ps[1];
      }
      e[7] += s.ele_coef * P.ewc[s.ewk0 + k] * (sc * sc + ss * ss);
    }
    if (t == 0) e[7] += s.ew_self;
  }
  // kinetic (energy_kinetic.F90:33-44)
  for (int li = t; li < s.nmp; li += kEBlock) {
    int g = s.off + li;
    if (!take(g, g)) continue;
    double v[3] = {P.v[g], P.v[n + g], P.v[2 * n + g]};
    e[8] += 0.5 * P.mass[g] * dot3(v, v);
  }
  double tot[kNE];
  for (int k = 0; k < kNE; ++k) tot[k] = block_sum(e[k], red);
  if (t == 0) {
    double *o = out + ((size_t)blockIdx.x * gridDim.y + blockIdx.y) * 25;
    for (int k = 0; k < 25; ++k) o[k] = 0.0;
    o[2] = tot[0]; o[3] = tot[1]; o[15] = tot[2]; o[18] = tot[3]; o[16] = tot[4]; o[19] = tot[5];
    o[17] = tot[6]; o[7] = tot[7]; o[1] = tot[8];
    double sum = 0.0;
    for (int k = 2; k < 25; ++k) sum += o[k];
    o[0] = sum;
  }
}

// ------------------------------------------------------------------ labels
// Replica exchange: system isys[k] takes the temperature-dependent
// coefficients of label lab[k] (charges, stack U0, descriptor fields and the
// Langevin noise amplitude c1 of time_integral_pre.F90:163-164, as the
// refresh of time_integral_post.F90:264-370 does) and, with vscale, has its
// velocities scaled (simu_replica_exchange.F90:272-300).  One block per
// entry; the descriptor is written after every lane has read it.
__global__ __launch_bounds__(256) void k_label_apply(DevPtrs P, LabelDev L, const int32_t *isys, const int32_t *lab,
                                                     const double *vscale) {
#pragma clang fp contract(off)
  const int k = blockIdx.x;
  const int is = isys[k], lb = lab[k];
  const DSys t = P.sys[is];
  const DSys src = L.sys[lb];
  const int n = P.ntot;
  double *charge = const_cast<double *>(P.charge);
  double *stc = const_cast<double *>(P.st_coef);
  const double f = vscale ? vscale[k] : 1.0;
  for (int q = threadIdx.x; q < t.nlc; q += blockDim.x) {
    LangClass &c = P.lcls[t.lc0 + q];
    if (P.tstep > 0.0) c.r[0] = sqrt(2.0 * c.fric * kBoltzKcalMol * src.tempk / (P.tstep * c.mass));
  }
  for (int i = threadIdx.x; i < t.nmp; i += blockDim.x) {
    const int g = t.off + i;
    charge[g] = L.q[L.qoff[lb] + i];
    if (vscale)
      for (int d = 0; d < 3; ++d) P.v[(size_t)d * n + g] *= f;
  }
  for (int q = threadIdx.x; q < t.nst; q += blockDim.x) stc[4 * (size_t)(t.soff + q)] = L.u0[L.u0off[lb] + q];
  __syncthreads();
  if (threadIdx.x == 0) {
    DSys d = t;
    copy_t_fields(d, src);
    const_cast<DSys *>(P.sys)[is] = d;
  }
}

}  // namespace tis
