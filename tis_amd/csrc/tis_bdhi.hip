// tis_bdhi.hip -- Brownian dynamics with hydrodynamic interaction on one
// MI355X: RPY diffusion tensor fill, blocked Cholesky (FP64 diagonal blocks
// and triangular solves, FP32 matrix-core trailing updates), and the
// Ermak-McCammon step (two matrix-vector products per step, HBM-bound).
//
// Layout: coordinates, forces and noise are (3,n) column-major vectors of
// length 3n (the reshape of time_integral.F90:433-434); D and L are (3n,3n)
// column-major FP64, L lower with a zero upper part.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tis_bdhi.h"
#include "tis_dev.h"
#include "tis_internal.h"

namespace tis {

constexpr int kCb = 64;         // Cholesky block size
constexpr int kGemvRows = 256;  // rows per block of the step products
typedef float floatx16 __attribute__((ext_vector_type(16)));

struct BdDev {
  int n, n3;
  double *x, *f, *xi;            // [3n]
  double *D, *L;                 // [n3 * n3]
  const double *radius;
  const int32_t *fb_off, *fb;    // per-bead FENE incidence (bond | side << 31)
  const int32_t *fene_mp;
  const double *fene_nat, *fene_coef, *fene_dist2;
  const int32_t *wb_off, *wb;    // per-bead WCA incidence (pair | side << 31)
  const int32_t *wca_mp;
  const double *wca_nat2, *wca_coef;
  double wca_rc2;                // 1 / cutoff_wca^2
  double c6, c8;
  int type;
  uint64_t seed;
  uint32_t key;
  int *info;                     // first failing pivot + 1, or 0
  double *stash;                 // two 64x64 diagonal factors awaiting their store
};

// ---------------------------------------------------------------- forces
// force_fene.F90:40-57 and force_wca.F90:64-96, owner-computes: each bead
// sums its own bonds and pairs in list order (no atomics).  with_noise also
// draws the bead's Gaussian triple for step `step`.
__global__ __launch_bounds__(256) void k_bd_force(BdDev B, uint64_t step, int with_noise) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.n) return;
  double F[3] = {0.0, 0.0, 0.0};
  for (int e = B.fb_off[i]; e < B.fb_off[i + 1]; ++e) {
    const uint32_t ent = (uint32_t)B.fb[e];
    const int b = (int)(ent & 0x7fffffffu);
    const int i0 = B.fene_mp[2 * b], j0 = B.fene_mp[2 * b + 1];
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = B.x[3 * j0 + d] - B.x[3 * i0 + d];
    const double dist = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const double dd = dist - B.fene_nat[b];
    const double fr = -B.fene_coef[b] * dd / (1.0 - dd * dd / B.fene_dist2[b]) / dist;
    const double sg = (ent >> 31) ? 1.0 : -1.0;
    for (int d = 0; d < 3; ++d) F[d] += sg * (fr * v[d]);
  }
  for (int e = B.wb_off[i]; e < B.wb_off[i + 1]; ++e) {
    const uint32_t ent = (uint32_t)B.wb[e];
    const int k = (int)(ent & 0x7fffffffu);
    const int i0 = B.wca_mp[2 * k], j0 = B.wca_mp[2 * k + 1];
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = B.x[3 * j0 + d] - B.x[3 * i0 + d];
    const double d2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const double ro2 = B.wca_nat2[k] / d2;
    if (ro2 < B.wca_rc2) continue;
    const double ro8 = ro2 * ro2 * ro2 * ro2;
    const double ro14 = ro2 * ro2 * ro2 * ro2 * ro2 * ro2 * ro2;
    const double c = ro2 > 1.0 ? B.wca_coef[2 * k] : B.wca_coef[2 * k + 1];
    double dv = 12.0 * c / B.wca_nat2[k] * (ro14 - ro8);
    if (dv > kDeMax) dv = kDeMax;
    const double sg = (ent >> 31) ? 1.0 : -1.0;
    for (int d = 0; d < 3; ++d) F[d] += sg * (dv * v[d]);
  }
  for (int d = 0; d < 3; ++d) B.f[3 * i + d] = F[d];
  if (with_noise) {
    double g[3];
    gauss3(B.seed, B.key, step, (uint32_t)i, g);
    for (int d = 0; d < 3; ++d) B.xi[3 * i + d] = g[d];
  }
}

// ---------------------------------------------------------------- tensor
// simu_hydro_tensors.F90:50-302.  Thread (i, j), i < j, writes the 3x3 block
// and its transpose into D and the lower copy into L (upper zero); i == j
// writes the diagonal block.
__global__ __launch_bounds__(256) void k_rpy_fill(BdDev B) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 16 + threadIdx.x, j = blockIdx.y * 16 + threadIdx.y;
  if (i >= B.n || j >= B.n || i > j) return;
  const int n3 = B.n3;
  if (i == j) {
    const double dg = B.c6 / (B.type == 4 ? B.radius[i] : B.radius[0]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const size_t o = (size_t)(3 * i + a) + (size_t)(3 * i + b) * n3;
        B.D[o] = a == b ? dg : 0.0;
        B.L[o] = (a == b) ? dg : 0.0;
      }
    return;
  }
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = B.x[3 * j + d] - B.x[3 * i + d];
  const double r_sq = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
  const double r_abs = sqrt(r_sq);
  double pre1, pre2;
  if (B.type == 4) {   // Zuk et al.
    const double ri = B.radius[i], rj = B.radius[j];
    if (ri + rj < r_abs) {
      pre1 = B.c8 / r_abs * (1.0 + (ri * ri + rj * rj) / (3.0 * r_sq));
      pre2 = B.c8 / r_abs * (1.0 - (ri * ri + rj * rj) / r_sq) / r_sq;
    } else {
      const double rs = ri < rj ? ri : rj, rl = ri < rj ? rj : ri;
      if (rl - rs < r_abs) {
        const double r_cb = r_sq * r_abs, dd = (ri - rj) * (ri - rj);
        pre1 = B.c6 / ri / rj * (16.0 * r_cb * (ri + rj) - (dd + 3.0 * r_sq) * (dd + 3.0 * r_sq)) / 32.0 / r_cb;
        pre2 = B.c6 / ri / rj * 3.0 * (dd - r_sq) * (dd - r_sq) / 32.0 / r_cb / r_sq;
      } else {
        pre1 = B.c6 / rl;
        pre2 = 0.0;
      }
    }
  } else {
    const double a = B.radius[0], a2 = a * a;
    if (B.type == 1 || 2.0 * a < r_abs) {
      pre1 = B.c8 / r_abs * (1.0 + 2.0 * a2 / (3.0 * r_sq));
      pre2 = B.c8 / r_abs * (1.0 - 2.0 * a2 / r_sq) / r_sq;
    } else if (B.type == 2) {
      pre1 = B.c6 / a * (1.0 - 9.0 / 32.0 * r_abs / a);
      pre2 = B.c6 / a * 3.0 / 32.0 / a / r_abs;
    } else {
      pre1 = B.c8 / 2.0 / a * 7.0 / 6.0;
      pre2 = B.c8 / 2.0 / 2.0 / r_sq;
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double v = pre2 * r[a] * r[b] + (a == b ? pre1 : 0.0);
      const size_t up = (size_t)(3 * i + a) + (size_t)(3 * j + b) * n3;   // row < col
      const size_t lo = (size_t)(3 * j + b) + (size_t)(3 * i + a) * n3;   // row > col
      B.D[up] = v;
      B.D[lo] = v;
      B.L[lo] = v;
      B.L[up] = 0.0;
    }
}

__global__ void k_bd_reset(BdDev B) {
  if (threadIdx.x == 0 && blockIdx.x == 0) atomicExch(B.info, 0);
}

// ---------------------------------------------------------------- Cholesky
// Blocked right-looking factorisation of L (holding the lower part of D) in
// place, in outer blocks of kNb = 256 columns made of four 64-column
// panels.  Panel j of the block starting at column k0 is one launch of
// k_chol_panel: its columns are brought up to date with the block's earlier
// panels (K = 64 j, FP32 matrix cores), then factored (the 64x64 diagonal
// block and the triangular solve of the rows below, FP64), its FP32 copy
// going to columns [64 j, 64 j + 64) of P (P[q * mpad + r - k0]).
// After the block's last panel the trailing lower triangle is updated with
// all 256 columns at once (k_chol_update, K = 256), so each trailing element
// is read and written once per 256 columns instead of once per 64.
constexpr int kNb = 256;

// LDS ordering inside one wave (writes by some lanes, reads by others)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The value of v at lane l of the wave, to every lane (two 32-bit reads).
__device__ inline double lane_value(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 1/x: the hardware estimate refined by two Newton steps y <- y + y (1 - x y).
__device__ inline double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// 1/sqrt(x) for x > 0: the hardware estimate refined by two Newton steps
// y <- y + y (1/2 - (x/2) y^2), each doubling the correct bits (to about
// 1 ulp in FP64); x <= 0 gives NaN or infinity, as 1/sqrt(x) would.
__device__ inline double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = fma(y, fma(-h * y, y, 0.5), y);
  y = fma(y, fma(-h * y, y, 0.5), y);
  return y;
}

// acc += (the value of r at lane n of each row of 16 lanes) * nx, as one
// v_fmac_f64 with a DPP row_newbcast operand (the same rounding as fma).
// The memory clobber keeps loads where the source puts them (ahead of use).
template <int N>
__device__ __forceinline__ void fmac_row_bcast(double &acc, double r, double nx) {
  asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "+v"(acc)
      : "v"(r), "v"(nx), "i"(N)
      : "memory");
}

// x[j] += v_j nx for j = J .. 63, where the column v reaches the lanes as
// four registers, R[k] holding v[16 k + (lane % 16)] (the same 16 values in
// each row of 16 lanes), from which row_newbcast takes entry j for every
// lane: four LDS reads per column instead of one broadcast read per entry.
template <int J>
__device__ __forceinline__ void bcast_update(double (&x)[kCb], const double (&R)[4], double nx) {
  if constexpr (J < kCb) {
    fmac_row_bcast<J & 15>(x[J], R[J >> 4], nx);
    bcast_update<J + 1>(x, R, nx);
  }
}
// x[j] += v_j nx for the local columns j = J .. 31 of a half of the block,
// v reaching the lanes as two registers (R[k] = v[16 k + lane % 16] of the
// half's rows) as in bcast_update.
template <int J>
__device__ __forceinline__ void bcast_update_half(double (&x)[kCb], const double (&R)[2], double nx) {
  if constexpr (J < 32) {
    fmac_row_bcast<J & 15>(x[J], R[J >> 4], nx);
    bcast_update_half<J + 1>(x, R, nx);
  }
}

// The 64x64 diagonal block as L D L^T, right-looking, by two waves: lane r
// holds row r, wave 0 its columns 0-31 and wave 1 its columns 32-63 (local
// index j = column - 32 w, in a[0..31]; rows past kb are unit rows).
// Column c's pivot d = A[c][c] and the entry A[c+1][c] come from lanes c and
// c + 1 by lane reads and each lane updates its entry of column c + 1 at
// once; the rest of the update, A[r][j] -= A[r][c] A[j][c] / d for j > c + 1,
// follows one column later from column c as published in LDS (U[c][r]),
// taken as registers and DPP row broadcasts.  For c < 32 the chain runs in
// wave 0, which also publishes the multipliers S[c][r] = A[r][c] / d and a
// count of published columns; wave 1 meanwhile applies each published
// column to its 32 columns (in column order, as one wave would), then runs
// the chain for c >= 32.  Column c is scaled to the Cholesky factor once
// final: L[r][c] = A[r][c] / sqrt(d), L[c][c] = sqrt(d), zero above the
// diagonal; Rd[c] = 1 / L[c][c].  Each wave's pivot check gives kk + c + 1
// for its first column c < kb with a pivot that is not positive, else 0.
struct DiagLds {
  double U[kCb][kCb];
  double S[32][kCb];
  int pub;    // columns of wave 0 published
  int bad0;   // wave 0's pivot check
};

template <int C>
__device__ __forceinline__ void diag_lo_step(double (&a)[kCb], DiagLds &D, double *Rd, int kk, int kb, int lane,
                                             int &bad, double sp) {
  if constexpr (C < 32) {
    double R[2];   // column C - 1 as published, rows 0-31 (sp = A[lane][C - 1] / d_{C - 1})
    if constexpr (C >= 1) {
#pragma unroll
      for (int k = 0; k < 2; ++k) R[k] = D.U[C - 1][16 * k + (lane & 15)];
      // column C - 1's update of column C + 1 first: column C + 1 is next on the chain
      if constexpr (C + 1 < 32) fmac_row_bcast<(C + 1) & 15>(a[C + 1], R[(C + 1) >> 4], -sp);
    }
    const double d = lane_value(a[C], C);
    if (C < kb && !(d > 0.0) && bad == 0) bad = kk + C + 1;
    const double u = a[C];
    D.U[C][lane] = u;
    const double s = u * rcp_nr(d);
    D.S[C][lane] = s;
    if constexpr (C == 31) {
      if (lane == 0) D.bad0 = bad;
    }
    // the column goes to wave 1: its LDS writes complete before the count
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(&D.pub, C + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if constexpr (C + 1 < 32) a[C + 1] = fma(-s, lane_value(u, C + 1), a[C + 1]);
    if constexpr (C >= 1) bcast_update_half<C + 2>(a, R, -sp);
    const double rs = rsqrt_nr(d);
    a[C] = lane == C ? d * rs : (lane > C ? u * rs : 0.0);
    if (lane == 0) Rd[C] = rs;
    wave_lds_sync();
    diag_lo_step<C + 1>(a, D, Rd, kk, kb, lane, bad, s);
  }
}

// wave 1: wave 0's column C applied to columns 32-63 once published
template <int C>
__device__ __forceinline__ void diag_hi_lag(double (&a)[kCb], DiagLds &D, int lane) {
  if constexpr (C < 32) {
    // wave 0 publishes all 32 columns unconditionally, so this wait ends
    while (__hip_atomic_load(&D.pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) <= C)
      __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const double s = D.S[C][lane];
    double R[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) R[k] = D.U[C][32 + 16 * k + (lane & 15)];
    bcast_update_half<0>(a, R, -s);
    diag_hi_lag<C + 1>(a, D, lane);
  }
}

template <int C>
__device__ __forceinline__ void diag_hi_step(double (&a)[kCb], DiagLds &D, double *Rd, int kk, int kb, int lane,
                                             int &bad, double sp) {
  if constexpr (C < kCb) {
    constexpr int c = C - 32;   // local column
    double R[2];   // column C - 1 as published, rows 32-63
    if constexpr (C >= 33) {
#pragma unroll
      for (int k = 0; k < 2; ++k) R[k] = D.U[C - 1][32 + 16 * k + (lane & 15)];
      if constexpr (C + 1 < kCb) fmac_row_bcast<(c + 1) & 15>(a[c + 1], R[(c + 1) >> 4], -sp);
    }
    const double d = lane_value(a[c], C);
    if (C < kb && !(d > 0.0) && bad == 0) bad = kk + C + 1;
    const double u = a[c];
    D.U[C][lane] = u;
    const double s = u * rcp_nr(d);
    if constexpr (C + 1 < kCb) a[c + 1] = fma(-s, lane_value(u, C + 1), a[c + 1]);
    if constexpr (C >= 33) bcast_update_half<c + 2>(a, R, -sp);
    const double rs = rsqrt_nr(d);
    a[c] = lane == C ? d * rs : (lane > C ? u * rs : 0.0);
    if (lane == 0) Rd[C] = rs;
    wave_lds_sync();
    diag_hi_step<C + 1>(a, D, Rd, kk, kb, lane, bad, s);
  }
}

// wave w (0 or 1) of the two-wave factorisation; wave 1 returns the pivot
// check of the whole block, wave 0 its own
__device__ inline int diag_factor2(double (&a)[kCb], DiagLds &D, double *Rd, int kk, int kb, int w) {
  const int lane = threadIdx.x & (kCb - 1);
  int bad = 0;
  if (w == 0) {
    diag_lo_step<0>(a, D, Rd, kk, kb, lane, bad, 0.0);
    return bad;
  }
  diag_hi_lag<0>(a, D, lane);
  diag_hi_step<32>(a, D, Rd, kk, kb, lane, bad, 0.0);
  return D.bad0 ? D.bad0 : bad;
}

template <int C>
__device__ __forceinline__ void solve_cols(double (&x)[kCb], const double *Ls, const double *Rd, int lane,
                                           const double (&R)[4], double rd) {
  if constexpr (C < kCb) {
    // the next column's entries are loaded a column ahead (the fmacs keep
    // the compiler from moving the loads down to their use)
    double Rn[4], rdn = 0.0;
    if constexpr (C + 1 < kCb) {
#pragma unroll
      for (int k = 0; k < 4; ++k) Rn[k] = Ls[(C + 1) * kCb + 16 * k + (lane & 15)];
      rdn = Rd[C + 1];
    }
    x[C] *= rd;
    bcast_update<C + 1>(x, R, -x[C]);
    solve_cols<C + 1>(x, Ls, Rd, lane, Rn, rdn);
  }
}

// Sum over k < K of P[k][ra + i] P[k][rb + j] for the 32x32 quadrant i, j <
// 32 (P column k at P + k * mpad), in the accumulator layout of
// v_mfma_f32_32x32x2f32 with the columns' entries as the first operand: the
// operands and the order of the sums are those of k_chol_update.  Operands
// come straight from memory (L2), eight k pairs ahead.
__device__ inline floatx16 quad_product(const float *P, int mpad, int ra, int rb, int K, int lane) {
  floatx16 acc;
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
  constexpr int kU = 8;   // k pairs per batch of loads
  const int l = lane & 31;
  const float *pa = P + (size_t)(lane >> 5) * mpad + ra + l;
  const float *pb = P + (size_t)(lane >> 5) * mpad + rb + l;
  float fa[kU], fb[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    fa[u] = pa[(size_t)(2 * u) * mpad];
    fb[u] = pb[(size_t)(2 * u) * mpad];
  }
  for (int s0 = 0; s0 < K / 2; s0 += kU) {
    float na[kU], nb[kU];
    const bool more = s0 + kU < K / 2;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const size_t off = (size_t)(2 * (more ? s0 + kU + u : u)) * mpad;
      na[u] = pa[off];
      nb[u] = pb[off];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fb[u], fa[u], acc, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      fa[u] = na[u];
      fb[u] = nb[u];
    }
  }
  return acc;
}

// One panel at column kk of the outer block at k0, in blocks of four waves;
// block b > 0 owns the 64 rows from kk + 64 b.  First the panel's columns
// take the update by the outer block's earlier panels (K = kk - k0 columns
// of P, FP32 matrix cores, quad_product; the same bits as with a separate
// k_chol_update launch): the four waves form the product for the 64x64
// diagonal block, a quadrant each, which passes through LDS into the
// row-per-lane layout and is subtracted from the FP64 values that wave 0
// loaded at the start; wave 0 then factors the diagonal block (diag_factor)
// while waves 2-3 form the product for the block's rows, which wave 2 (its
// rows loaded at the start) subtracts before its solve.
// Block 0 keeps the factor in a stash (two buffers, by panel parity) and the
// next panel's block 0 stores it over the diagonal block, so no block of
// this panel can read a partly overwritten block; a panel of one block (the
// last) stores its own factor, and the kernels between two panels do not
// touch the diagonal blocks of earlier panels.  In blocks 1.. the factor
// goes to LDS and wave 2 solves the block's rows by forward substitution
// (row per lane; see solve_cols), storing the rows and their FP32 copy in P.
// Full tiles take branch-free stores.  (Four waves leave the diagonal
// factorisation the registers of one wave per SIMD.)
constexpr int kPanelThreads = 256;
constexpr int kPanelRowsPerBlock = kCb;   // rows of the panel solved per block

__global__ __launch_bounds__(kPanelThreads) void k_chol_panel(BdDev B, int kk, int k0, float *P, int mpad) {
  // LDS: the diagonal block's product (64 x 65 floats), later the diagonal
  // factor Ls[c * 64 + r] = L11[r][c]; the product for the block's rows
  __shared__ __align__(16) double smem[kCb * kCb];
  __shared__ float Tr[kCb][kCb + 1];
  __shared__ double Rd[kCb];         // 1 / L11[c][c]
  __shared__ DiagLds D;              // the diagonal factorisation's published columns
  static_assert(sizeof(smem) >= kCb * (kCb + 1) * sizeof(float), "the product does not fit");
  double *Ls = smem;
  float (*Td)[kCb + 1] = reinterpret_cast<float (*)[kCb + 1]>(smem);
  const int lane = threadIdx.x & (kCb - 1), w = threadIdx.x / kCb, n3 = B.n3;
  const int kb = min(kCb, n3 - kk);
  const int K = kk - k0;
  const int r0w = kk + kCb * (int)blockIdx.x;   // the block's first row (blocks 1..)
  const bool rows = blockIdx.x > 0 && r0w < n3;
  const int r = r0w + lane;
  const int p = kk / kCb;
  // waves 0 and 1: the diagonal block's rows, columns [32 w, 32 w + 32);
  // wave 2: the block's rows (unconditional loads from clamped, in-bounds
  // addresses)
  double v[kCb];
  if (w < 2) {
    const size_t row = (size_t)min(kk + lane, n3 - 1);
#pragma unroll
    for (int c = 0; c < 32; ++c) v[c] = B.L[row + (size_t)min(kk + 32 * w + c, n3 - 1) * n3];
  } else if (w == 2) {
    const size_t row = (size_t)min(r, n3 - 1);
#pragma unroll
    for (int c = 0; c < kCb; ++c) v[c] = B.L[row + (size_t)min(kk + c, n3 - 1) * n3];
  }
  if (threadIdx.x == 0) D.pub = 0;
  if (w == 0 && blockIdx.x == 0 && p > 0) {
    // the previous panel's diagonal factor, kept aside until no block of
    // that panel reads the diagonal block (a full block: not the last panel)
    const double *sp = B.stash + (size_t)((p - 1) & 1) * kCb * kCb;
    const int kp = kk - kCb;
#pragma unroll 8
    for (int c = 0; c < kCb; ++c) B.L[(size_t)(kp + lane) + (size_t)(kp + c) * n3] = sp[c * kCb + lane];
  }
  const int l32 = lane & 31, c32 = 4 * (lane >> 5);   // accumulator e: row l32, column c32 + 8 (e / 4) + e % 4
  if (K > 0) {
    const int qi = w >> 1, qj = w & 1;
    const floatx16 acc = quad_product(P, mpad, kk - k0 + 32 * qi, kk - k0 + 32 * qj, K, lane);
#pragma unroll
    for (int e = 0; e < 16; ++e) Td[32 * qi + l32][32 * qj + c32 + 8 * (e >> 2) + (e & 3)] = acc[e];
  }
  __syncthreads();   // the diagonal block's product; D.pub cleared
  if (w < 2) {
    const int c0 = 32 * w;
    if (K > 0) {
#pragma unroll
      for (int c = 0; c < 32; ++c) v[c] -= (double)Td[lane][c0 + c];
    }
#pragma unroll
    for (int c = 0; c < 32; ++c) v[c] = (lane < kb && c0 + c < kb) ? v[c] : (c0 + c == lane ? 1.0 : 0.0);
    const int bad = diag_factor2(v, D, Rd, kk, kb, w);
    if (blockIdx.x > 0) {
#pragma unroll
      for (int c = 0; c < 32; ++c) Ls[(c0 + c) * kCb + lane] = v[c];
    } else if (gridDim.x > 1) {
      // other blocks may still be reading the diagonal block: the factor
      // waits in the stash for the next panel's launch
      if (w == 1 && lane == 0 && bad) atomicCAS(B.info, 0, bad);
      double *sp = B.stash + (size_t)(p & 1) * kCb * kCb;
#pragma unroll
      for (int c = 0; c < 32; ++c) sp[(c0 + c) * kCb + lane] = v[c];
    } else {
      if (w == 1 && lane == 0 && bad) atomicCAS(B.info, 0, bad);
      if (kb == kCb) {
#pragma unroll
        for (int c = 0; c < 32; ++c) B.L[(size_t)(kk + lane) + (size_t)(kk + c0 + c) * n3] = v[c];
      } else if (lane < kb) {
#pragma unroll
        for (int c = 0; c < 32; ++c)
          if (c0 + c < kb) B.L[(size_t)(kk + lane) + (size_t)(kk + c0 + c) * n3] = v[c];
      }
    }
  }
  if (blockIdx.x == 0) return;
  if (K > 0 && w >= 2 && rows) {
    // the product for the block's rows, beside the diagonal factorisation
    const int qi = w & 1;
#pragma unroll
    for (int qj = 0; qj < 2; ++qj) {
      const floatx16 acc = quad_product(P, mpad, r0w - k0 + 32 * qi, kk - k0 + 32 * qj, K, lane);
#pragma unroll
      for (int e = 0; e < 16; ++e) Tr[32 * qi + l32][32 * qj + c32 + 8 * (e >> 2) + (e & 3)] = acc[e];
    }
  }
  __syncthreads();
  if (w != 2 || !rows) return;
  if (K > 0) {
#pragma unroll
    for (int c = 0; c < kCb; ++c) v[c] -= (double)Tr[lane][c];
  }
  const bool live = r < n3;
  {
    double R0[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) R0[k] = Ls[16 * k + (lane & 15)];
    solve_cols<0>(v, Ls, Rd, lane, R0, Rd[0]);
  }
  const int rl = r - k0;   // row of P
  const int q0 = kk - k0;  // first column of P
  if (kb == kCb && r0w + kCb <= n3) {
#pragma unroll
    for (int c = 0; c < kCb; ++c) {
      B.L[(size_t)r + (size_t)(kk + c) * n3] = v[c];
      P[(size_t)(q0 + c) * mpad + rl] = (float)v[c];
    }
  } else if (rl < mpad) {
#pragma unroll
    for (int c = 0; c < kCb; ++c) {
      if (live && c < kb) B.L[(size_t)r + (size_t)(kk + c) * n3] = v[c];
      P[(size_t)(q0 + c) * mpad + rl] = (live && c < kb) ? (float)v[c] : 0.0f;
    }
  }
}

// C -= A B^T on 64x64 tiles of the lower part of L, FP32 matrix cores
// (v_mfma_f32_32x32x2f32, each of four waves a 32x32 quadrant), the sum
// over K columns of P in chunks of 64 staged through LDS, subtracted from
// the FP64 matrix (entries above the diagonal are not touched).
// The tiles are the lower triangle of tiles from row / column r0 (block b is
// the tile (ti, tj), tj <= ti, in row order).  Row r of A is P row r - k0.
// The products are formed transposed (the tile's columns as the first
// operand), so that each accumulator register holds 32 consecutive rows of
// one column: the tile's old values are loaded straight into that layout
// before the products (their latency hides behind the matrix work) and the
// results are stored from it, both as 256-byte column pieces.  The next K
// chunk's loads are issued before the current chunk's products.
__global__ __launch_bounds__(256) void k_chol_update(BdDev B, const float *P, int mpad, int k0, int r0, int K) {
  __shared__ float Pa[kCb][kCb + 1], Pb[kCb][kCb + 1];
  const int tid = threadIdx.x, n3 = B.n3;
  const int bidx = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)bidx + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= bidx) ++ti;
  while (ti * (ti + 1) / 2 > bidx) --ti;
  const int tj = bidx - ti * (ti + 1) / 2;
  const int ra = r0 - k0 + ti * kCb, rb = r0 - k0 + tj * kCb;   // P rows of the tile's rows / columns
  const int w = tid / 64, lane = tid & 63;
  const int qi = w >> 1, qj = w & 1;
  // accumulator v of this lane: row gr, column gc0 + 32 qj + 8 (v / 4) + 4 (lane / 32) + v % 4
  const int gr = r0 + ti * kCb + 32 * qi + (lane & 31);
  const int gcl = r0 + tj * kCb + 32 * qj + 4 * (lane >> 5);
  double old[16];
  {
    const size_t grc = (size_t)min(gr, n3 - 1);   // clamped: in bounds, unconditional
#pragma unroll
    for (int v = 0; v < 16; ++v)
      old[v] = B.L[grc + (size_t)min(gcl + 8 * (v >> 2) + (v & 3), n3 - 1) * n3];
  }
  floatx16 acc;
  for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
  const int ar = 32 * qi + (lane & 31), br = 32 * qj + (lane & 31), kq = lane >> 5;
  float ta[kCb * kCb / 256], tb[kCb * kCb / 256];
#pragma unroll
  for (int q = 0; q < kCb * kCb / 256; ++q) {
    const int idx = tid + 256 * q, r = idx % kCb, c = idx / kCb;
    ta[q] = P[(size_t)c * mpad + ra + r];
    tb[q] = P[(size_t)c * mpad + rb + r];
  }
  for (int q0 = 0; q0 < K; q0 += kCb) {
    if (q0) __syncthreads();   // the last chunk's products have read the LDS tiles
#pragma unroll
    for (int q = 0; q < kCb * kCb / 256; ++q) {
      const int idx = tid + 256 * q, r = idx % kCb, c = idx / kCb;
      Pa[r][c] = ta[q];
      Pb[r][c] = tb[q];
    }
    __syncthreads();
    if (q0 + kCb < K) {
#pragma unroll
      for (int q = 0; q < kCb * kCb / 256; ++q) {
        const int idx = tid + 256 * q, r = idx % kCb, c = idx / kCb;
        ta[q] = P[(size_t)(q0 + kCb + c) * mpad + ra + r];
        tb[q] = P[(size_t)(q0 + kCb + c) * mpad + rb + r];
      }
    }
#pragma unroll 8
    for (int s2 = 0; s2 < kCb / 2; ++s2)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Pb[br][2 * s2 + kq], Pa[ar][2 * s2 + kq], acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int gc = gcl + 8 * (v >> 2) + (v & 3);
    if (gr < n3 && gc < n3 && gr >= gc) B.L[(size_t)gr + (size_t)gc * n3] = old[v] - (double)acc[v];
  }
}

__global__ __launch_bounds__(256) void k_copy_lower(const double *D, double *L, int n3) {
  const int r = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
  if (r < n3) L[(size_t)r + (size_t)c * n3] = r >= c ? D[(size_t)r + (size_t)c * n3] : 0.0;
}

// ---------------------------------------------------------------- no-alloc scratch
// Scratch of tis_bdhi_factor_noalloc: the FP32 copy P of an outer block's
// panels (256 columns of at most kNoAllocMaxN3 rows) followed by the pivot
// flag and the stash of two diagonal factors.  Each device holds its own copy until the process ends; every call
// on a device shares it.
constexpr int kNoAllocMaxN3 = 9216;
constexpr size_t kScratchP = (size_t)kNb * kNoAllocMaxN3;
constexpr size_t kScratchStash = kScratchP + 64;   // two 64x64 FP64 blocks
static __device__ __align__(256) float scratch[kScratchStash + 4 * kCb * kCb];   // P, info, stash

// ---------------------------------------------------------------- step
// Partial products over a column chunk: yD = D[:, chunk] F[chunk],
// yL = L[:, chunk] xi[chunk] (lower part only); one thread per row.
__global__ __launch_bounds__(kGemvRows) void k_bd_gemv(BdDev B, double *ypart, int chunk) {
  const int k = blockIdx.x * kGemvRows + threadIdx.x;
  const int c = blockIdx.y, n3 = B.n3;
  const int l0 = c * chunk, l1 = min(n3, l0 + chunk);
  if (k >= n3) return;
  double aD = 0.0, aL = 0.0;
  for (int l = l0; l < l1; ++l) aD += B.D[(size_t)l * n3 + k] * B.f[l];
  const int lL = min(l1, k + 1);
  for (int l = l0; l < lL; ++l) aL += B.L[(size_t)l * n3 + k] * B.xi[l];
  ypart[(size_t)(2 * c) * n3 + k] = aD;
  ypart[(size_t)(2 * c + 1) * n3 + k] = aL;
}

// time_integral.F90:440-446: dx = (dt/kT) D F + sqrt(2 dt) L xi
__global__ __launch_bounds__(kGemvRows) void k_bd_update(BdDev B, const double *ypart, int nchunk, double c1,
                                                          double c2) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kGemvRows + threadIdx.x;
  if (k >= B.n3) return;
  double a = 0.0, b = 0.0;
  for (int c = 0; c < nchunk; ++c) { a += ypart[(size_t)(2 * c) * B.n3 + k]; b += ypart[(size_t)(2 * c + 1) * B.n3 + k]; }
  B.x[k] += c1 * a + c2 * b;
}

// ||L L^T - D||_F^2 and ||D||_F^2 over 64x64 lower tiles (off-diagonal
// tiles counted twice); FP64, K staged through LDS in chunks of 16.
__global__ __launch_bounds__(256) void k_bd_llt_err(BdDev B, double *part) {
  __shared__ double Li[kCb][17], Lj[kCb][17];
  __shared__ double red[2][256];
  const int tid = threadIdx.x, n3 = B.n3;
  const int bidx = blockIdx.x;
  int ti = (int)((sqrt(8.0 * (double)bidx + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= bidx) ++ti;
  while (ti * (ti + 1) / 2 > bidx) --ti;
  const int tj = bidx - ti * (ti + 1) / 2;
  const int r0 = ti * kCb, c0 = tj * kCb;
  // thread owns a 4x4 sub-block
  const int sr = (tid % 16) * 4, sc = (tid / 16) * 4;
  double acc[4][4] = {};
  const int kend = min(n3, c0 + kCb);  // L[j, l] = 0 for l > j
  for (int l0 = 0; l0 < kend; l0 += 16) {
    for (int idx = tid; idx < kCb * 16; idx += 256) {
      const int r = idx % kCb, l = idx / kCb;
      const int gl = l0 + l;
      Li[r][l] = (r0 + r < n3 && gl < n3) ? B.L[(size_t)(r0 + r) + (size_t)gl * n3] : 0.0;
      Lj[r][l] = (c0 + r < n3 && gl < n3) ? B.L[(size_t)(c0 + r) + (size_t)gl * n3] : 0.0;
    }
    __syncthreads();
    for (int l = 0; l < 16; ++l)
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) acc[a][b] += Li[sr + a][l] * Lj[sc + b][l];
    __syncthreads();
  }
  double e2 = 0.0, d2 = 0.0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const int gr = r0 + sr + a, gc = c0 + sc + b;
      if (gr >= n3 || gc >= n3 || gr < gc) continue;
      const double dv = B.D[(size_t)gr + (size_t)gc * n3];
      const double e = acc[a][b] - dv;
      const double wgt = gr == gc ? 1.0 : 2.0;
      e2 += wgt * e * e;
      d2 += wgt * dv * dv;
    }
  red[0][tid] = e2; red[1][tid] = d2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) { red[0][tid] += red[0][tid + s]; red[1][tid] += red[1][tid + s]; }
    __syncthreads();
  }
  if (tid == 0) { part[2 * bidx] = red[0][0]; part[2 * bidx + 1] = red[1][0]; }
}

// ---------------------------------------------------------------- host state
struct BdState {
  BdDev d{};
  std::vector<void *> allocs;
  float *P = nullptr;
  int mpad = 0;
  double *ypart = nullptr;
  int nchunk = 1, chunk = 1;
  double *llt_part = nullptr;
  double tempk = 0.0, visc = 0.0;
  hipGraphExec_t fill_graph = nullptr, factor_graph = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  double last_total_ms = 0.0, last_fill_ms = 0.0, last_factor_ms = 0.0;
};

void bd_free(tis_ctx *c) {
  BdState *&b = ctx_bd(c);
  if (!b) return;
  hipStream_t st = ctx_stream(c);
  if (st) (void)hipStreamSynchronize(st);
  if (b->fill_graph) (void)hipGraphExecDestroy(b->fill_graph);
  if (b->factor_graph) (void)hipGraphExecDestroy(b->factor_graph);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->e2) (void)hipEventDestroy(b->e2);
  delete b;
  b = nullptr;
}

namespace {

int bfail(tis_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return ctx_fail(c, code, buf);
}

#define BCHK(ctx, expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return bfail(ctx, TIS_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int balloc(tis_ctx *c, BdState *b, T **p, size_t n) {
  *p = nullptr;
  const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)p, bytes);
  if (e != hipSuccess) return bfail(c, TIS_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  b->allocs.push_back((void *)*p);
  e = hipMemsetAsync(*p, 0, bytes, ctx_stream(c));
  if (e != hipSuccess) return bfail(c, TIS_ERR_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  return 0;
}

template <typename T>
int bupload(tis_ctx *c, BdState *b, T **p, const std::vector<T> &h) {
  int rc = balloc(c, b, p, h.size());
  if (rc) return rc;
  if (!h.empty()) BCHK(c, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx_stream(c)));
  return 0;
}

BdState *need(tis_ctx *c) {
  BdState *b = ctx_bd(c);
  if (!b) ctx_fail(c, TIS_ERR_STATE, "tis_bdhi_setup not called");
  return b;
}

// fill, and the factorisation, each captured once as a graph (the launch
// sequence depends only on n)
void launch_fill(tis_ctx *c, BdState *b) {
  hipStream_t st = ctx_stream(c);
  const BdDev &d = b->d;
  hipLaunchKernelGGL(k_bd_reset, dim3(1), dim3(64), 0, st, d);
  const int gb = (d.n + 15) / 16;
  hipLaunchKernelGGL(k_rpy_fill, dim3(gb, gb), dim3(16, 16), 0, st, d);
}

// The factorisation's launches, outer block by outer block: its panels
// (each first updating its columns with the block's earlier panels), then
// the trailing update with the whole block.  P holds kNb columns of at most
// n3 rows (rows from the block's start).
void factor_launches(const BdDev &d, float *P, hipStream_t st) {
  const int n3 = d.n3;
  for (int k0 = 0; k0 < n3; k0 += kNb) {
    const int mpad = (n3 - k0 + kCb - 1) / kCb * kCb;
    for (int j = 0; j < kNb / kCb && k0 + j * kCb < n3; ++j) {
      const int kk = k0 + j * kCb;
      const int m = n3 - kk - kCb;
      const int nbp = m > 0 ? (m + kPanelRowsPerBlock - 1) / kPanelRowsPerBlock : 0;
      hipLaunchKernelGGL(k_chol_panel, dim3(1 + nbp), dim3(kPanelThreads), 0, st, d, kk, k0, P, mpad);
    }
    const int kt = k0 + kNb;
    if (kt >= n3) break;
    const int mt = (n3 - kt + kCb - 1) / kCb;
    hipLaunchKernelGGL(k_chol_update, dim3(mt * (mt + 1) / 2), dim3(256), 0, st, d, (const float *)P, mpad, k0, kt,
                       kNb);
  }
}

void launch_factor(tis_ctx *c, BdState *b) { factor_launches(b->d, b->P, ctx_stream(c)); }

int capture(tis_ctx *c, BdState *b, void (*fn)(tis_ctx *, BdState *), hipGraphExec_t *out) {
  hipStream_t st = ctx_stream(c);
  hipGraph_t g = nullptr;
  BCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  fn(c, b);
  hipError_t e = hipStreamEndCapture(st, &g);
  if (e != hipSuccess) return bfail(c, TIS_ERR_HIP, "capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return bfail(c, TIS_ERR_HIP, "instantiate failed: %s", hipGetErrorString(e));
  return 0;
}

// events e0 | fill | e1 | factorisation | e2 on the context's stream
int run_tensors(tis_ctx *c, BdState *b) {
  hipStream_t st = ctx_stream(c);
  int rc = 0;
  if (!b->fill_graph && (rc = capture(c, b, launch_fill, &b->fill_graph))) return rc;
  if (!b->factor_graph && (rc = capture(c, b, launch_factor, &b->factor_graph))) return rc;
  BCHK(c, hipEventRecord(b->e0, st));
  BCHK(c, hipGraphLaunch(b->fill_graph, st));
  BCHK(c, hipEventRecord(b->e1, st));
  BCHK(c, hipGraphLaunch(b->factor_graph, st));
  BCHK(c, hipEventRecord(b->e2, st));
  return 0;
}

int check_info(tis_ctx *c, BdState *b) {
  int info = 0;
  BCHK(c, hipMemcpyAsync(&info, b->d.info, sizeof(int), hipMemcpyDeviceToHost, ctx_stream(c)));
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  if (info) return bfail(c, TIS_ERR_STATE, "Cholesky failed: non-positive pivot at row %d", info - 1);
  return 0;
}

// device address of `scratch` per device, taken once per device with that
// device current (the array's name is a host-side stand-in)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_ptr(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TIS_ERR_HIP;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return TIS_OK;
}

}  // namespace
}  // namespace tis

using namespace tis;

extern "C" {

int32_t tis_bdhi_factor_noalloc(int32_t device, const double *D, double *L, int32_t n3, void *stream,
                                int32_t *info) {
  if (!D || !L || n3 <= 0 || n3 > kNoAllocMaxN3 || device < 0 || device >= kMaxDevices) return TIS_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return TIS_ERR_HIP;
  void *ptr = nullptr;
  if (int rc = scratch_ptr(device, &ptr)) return rc;
  hipStream_t st = (hipStream_t)stream;
  BdDev d{};
  d.n3 = n3;
  d.n = n3 / 3;
  d.L = L;
  d.D = const_cast<double *>(D);
  d.info = (int *)((float *)ptr + kScratchP);
  d.stash = (double *)((float *)ptr + kScratchStash);
  hipLaunchKernelGGL(k_bd_reset, dim3(1), dim3(64), 0, st, d);
  hipLaunchKernelGGL(k_copy_lower, dim3((n3 + 255) / 256, n3), dim3(256), 0, st, D, L, n3);
  factor_launches(d, (float *)ptr, st);
  if (hipGetLastError() != hipSuccess) return TIS_ERR_HIP;
  if (info) {
    int v = 0;
    if (hipMemcpyAsync(&v, d.info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return TIS_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return TIS_ERR_HIP;
    *info = v;
    if (v) return TIS_ERR_STATE;
  }
  return TIS_OK;
}

int32_t tis_bdhi_setup(tis_ctx *c, const tis_bd_system *s, uint32_t key, uint64_t seed) {
  if (!c || !s || s->nmp <= 1 || !s->radius) return TIS_ERR_ARG;
  if (s->hydro_tensor < 1 || s->hydro_tensor > 4) return bfail(c, TIS_ERR_ARG, "hydro_tensor must be 1-4");
  if (!(s->viscosity > 0.0) || !(s->tempk > 0.0)) return bfail(c, TIS_ERR_ARG, "bad temperature or viscosity");
  if (s->nwca > 0 && !(s->cutoff_wca > 0.0)) return bfail(c, TIS_ERR_ARG, "cutoff_wca must be > 0");
  const int64_t n3 = 3 * (int64_t)s->nmp;
  if (n3 > 46000) return bfail(c, TIS_ERR_ARG, "too many beads for a dense tensor (%d)", s->nmp);
  BCHK(c, hipSetDevice(ctx_device(c)));
  bd_free(c);
  BdState *b = new BdState();
  ctx_bd(c) = b;
  BdDev &d = b->d;
  const int n = s->nmp;
  d.n = n; d.n3 = (int)n3;
  for (int k = 0; k < s->nfene; ++k)
    for (int e = 0; e < 2; ++e)
      if (s->fene_mp[2 * k + e] < 0 || s->fene_mp[2 * k + e] >= n) return bfail(c, TIS_ERR_ARG, "bad FENE bead index");
  for (int k = 0; k < s->nwca; ++k)
    for (int e = 0; e < 2; ++e)
      if (s->wca_mp[2 * k + e] < 0 || s->wca_mp[2 * k + e] >= n) return bfail(c, TIS_ERR_ARG, "bad WCA bead index");
  // per-bead incidence lists (bond order, pair order)
  auto csr = [n](int m, const int32_t *mp, std::vector<int32_t> &off, std::vector<int32_t> &ent) {
    std::vector<std::vector<int32_t>> inc(n);
    for (int k = 0; k < m; ++k) {
      inc[mp[2 * k]].push_back(k);
      inc[mp[2 * k + 1]].push_back((int32_t)((uint32_t)k | 0x80000000u));
    }
    off.assign(1, 0);
    for (int i = 0; i < n; ++i) {
      ent.insert(ent.end(), inc[i].begin(), inc[i].end());
      off.push_back((int32_t)ent.size());
    }
  };
  std::vector<int32_t> fb_off, fb, wb_off, wb;
  csr(s->nfene, s->fene_mp, fb_off, fb);
  csr(s->nwca, s->wca_mp, wb_off, wb);
  std::vector<double> rad(s->radius, s->radius + n);
  std::vector<int32_t> fmp(s->fene_mp, s->fene_mp + 2 * s->nfene), wmp(s->wca_mp, s->wca_mp + 2 * s->nwca);
  std::vector<double> fnat(s->fene_nat, s->fene_nat + s->nfene), fcoef(s->fene_coef, s->fene_coef + s->nfene),
      fd2(s->fene_dist2, s->fene_dist2 + s->nfene), wn2(s->wca_nat2, s->wca_nat2 + s->nwca),
      wcf(s->wca_coef, s->wca_coef + 2 * s->nwca);
  int rc = 0;
#define BUP(ptr, vec) do { rc = bupload(c, b, &ptr, vec); if (rc) return rc; } while (0)
  { int32_t *p; BUP(p, fb_off); d.fb_off = p; BUP(p, fb); d.fb = p; BUP(p, fmp); d.fene_mp = p;
    BUP(p, wb_off); d.wb_off = p; BUP(p, wb); d.wb = p; BUP(p, wmp); d.wca_mp = p; }
  { double *p; BUP(p, rad); d.radius = p; BUP(p, fnat); d.fene_nat = p; BUP(p, fcoef); d.fene_coef = p;
    BUP(p, fd2); d.fene_dist2 = p; BUP(p, wn2); d.wca_nat2 = p; BUP(p, wcf); d.wca_coef = p; }
#undef BUP
  if ((rc = balloc(c, b, &d.x, (size_t)n3))) return rc;
  if ((rc = balloc(c, b, &d.f, (size_t)n3))) return rc;
  if ((rc = balloc(c, b, &d.xi, (size_t)n3))) return rc;
  if ((rc = balloc(c, b, &d.D, (size_t)n3 * n3))) return rc;
  if ((rc = balloc(c, b, &d.L, (size_t)n3 * n3))) return rc;
  const int mt = (int)((n3 + kCb - 1) / kCb);
  if ((rc = balloc(c, b, &d.info, 1))) return rc;
  if ((rc = balloc(c, b, &d.stash, (size_t)2 * kCb * kCb))) return rc;
  b->mpad = mt * kCb;
  if ((rc = balloc(c, b, &b->P, (size_t)kNb * b->mpad))) return rc;
  // step products: ~1024+ blocks
  const int rows = (int)((n3 + kGemvRows - 1) / kGemvRows);
  b->nchunk = std::max(1, std::min((int)((n3 + 127) / 128), (1024 + rows - 1) / rows));
  b->chunk = (int)((n3 + b->nchunk - 1) / b->nchunk);
  if ((rc = balloc(c, b, &b->ypart, (size_t)2 * b->nchunk * n3))) return rc;
  if ((rc = balloc(c, b, &b->llt_part, (size_t)2 * mt * (mt + 1) / 2))) return rc;
  d.wca_rc2 = s->nwca > 0 ? 1.0 / (s->cutoff_wca * s->cutoff_wca) : 0.0;
  d.c6 = kBoltzKcalMol * s->tempk / (6.0 * kPi * s->viscosity);
  d.c8 = kBoltzKcalMol * s->tempk / (8.0 * kPi * s->viscosity);
  d.type = s->hydro_tensor;
  d.seed = seed;
  d.key = key;
  b->tempk = s->tempk;
  b->visc = s->viscosity;
  BCHK(c, hipEventCreate(&b->e0));
  BCHK(c, hipEventCreate(&b->e1));
  BCHK(c, hipEventCreate(&b->e2));
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  return TIS_OK;
}

int32_t tis_bdhi_state_set(tis_ctx *c, int32_t which, const double *buf) {
  if (!c || !buf) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  if (which != 0) return bfail(c, TIS_ERR_ARG, "only coordinates can be set");
  BCHK(c, hipSetDevice(ctx_device(c)));
  BCHK(c, hipMemcpyAsync(b->d.x, buf, sizeof(double) * 3 * (size_t)b->d.n, hipMemcpyHostToDevice, ctx_stream(c)));
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  return TIS_OK;
}

int32_t tis_bdhi_state_get(tis_ctx *c, int32_t which, double *buf) {
  if (!c || !buf) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  if (which != 0 && which != 1) return bfail(c, TIS_ERR_ARG, "which must be 0 or 1");
  BCHK(c, hipSetDevice(ctx_device(c)));
  BCHK(c, hipMemcpyAsync(buf, which == 0 ? b->d.x : b->d.f, sizeof(double) * 3 * (size_t)b->d.n,
                         hipMemcpyDeviceToHost, ctx_stream(c)));
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  return TIS_OK;
}

int32_t tis_bdhi_forces(tis_ctx *c, double *force_out) {
  if (!c) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  BCHK(c, hipSetDevice(ctx_device(c)));
  hipLaunchKernelGGL(k_bd_force, dim3((b->d.n + 255) / 256), dim3(256), 0, ctx_stream(c), b->d, (uint64_t)0, 0);
  BCHK(c, hipGetLastError());
  if (force_out) return tis_bdhi_state_get(c, 1, force_out);
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  return TIS_OK;
}

int32_t tis_bdhi_tensors(tis_ctx *c) {
  if (!c) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  BCHK(c, hipSetDevice(ctx_device(c)));
  int rc = run_tensors(c, b);
  if (rc) return rc;
  BCHK(c, hipGetLastError());
  rc = check_info(c, b);
  float a = 0.0f, f = 0.0f;
  (void)hipEventElapsedTime(&a, b->e0, b->e1);
  (void)hipEventElapsedTime(&f, b->e1, b->e2);
  b->last_fill_ms = a;
  b->last_factor_ms = f;
  b->last_total_ms = (double)a + (double)f;
  return rc;
}

int32_t tis_bdhi_tensor_get(tis_ctx *c, int32_t which, double *buf) {
  if (!c || !buf) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  if (which != 0 && which != 1) return bfail(c, TIS_ERR_ARG, "which must be 0 (D) or 1 (L)");
  BCHK(c, hipSetDevice(ctx_device(c)));
  BCHK(c, hipMemcpyAsync(buf, which == 0 ? b->d.D : b->d.L, sizeof(double) * (size_t)b->d.n3 * b->d.n3,
                         hipMemcpyDeviceToHost, ctx_stream(c)));
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  return TIS_OK;
}

int32_t tis_bdhi_factor_error(tis_ctx *c, double *rel_err) {
  if (!c || !rel_err) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  BCHK(c, hipSetDevice(ctx_device(c)));
  const int mt = (b->d.n3 + kCb - 1) / kCb;
  const int nt = mt * (mt + 1) / 2;
  hipLaunchKernelGGL(k_bd_llt_err, dim3(nt), dim3(256), 0, ctx_stream(c), b->d, b->llt_part);
  BCHK(c, hipGetLastError());
  std::vector<double> part(2 * (size_t)nt);
  BCHK(c, hipMemcpyAsync(part.data(), b->llt_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx_stream(c)));
  BCHK(c, hipStreamSynchronize(ctx_stream(c)));
  double e2 = 0.0, d2 = 0.0;
  for (int k = 0; k < nt; ++k) { e2 += part[2 * k]; d2 += part[2 * k + 1]; }
  *rel_err = d2 > 0.0 ? std::sqrt(e2 / d2) : 0.0;
  return TIS_OK;
}

int32_t tis_bdhi_steps(tis_ctx *c, double tstep, int64_t nsteps, int32_t refactor_every, uint64_t step0) {
  if (!c || !(tstep > 0.0) || nsteps < 0 || refactor_every <= 0) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  BCHK(c, hipSetDevice(ctx_device(c)));
  hipStream_t st = ctx_stream(c);
  const BdDev &d = b->d;
  // time_integral_pre.F90:313-317
  const double c1 = tstep / (kBoltzKcalMol * b->tempk), c2 = std::sqrt(2.0 * tstep);
  hipEvent_t t0, t1;
  BCHK(c, hipEventCreate(&t0));
  BCHK(c, hipEventCreate(&t1));
  BCHK(c, hipEventRecord(t0, st));
  double fact_ms = 0.0, fill_ms = 0.0;
  const int rows = (d.n3 + kGemvRows - 1) / kGemvRows;
  for (int64_t t = 0; t < nsteps; ++t) {
    if (t % refactor_every == 0) {
      int rc = run_tensors(c, b);
      if (rc) return rc;
      if (nsteps <= 64 * refactor_every || t == 0) {   // pivots checked on the first refactor and short runs
        rc = check_info(c, b);
        if (rc) { (void)hipEventDestroy(t0); (void)hipEventDestroy(t1); return rc; }
        float a = 0.0f, f = 0.0f;
        (void)hipEventElapsedTime(&a, b->e0, b->e1);
        (void)hipEventElapsedTime(&f, b->e1, b->e2);
        fill_ms += a; fact_ms += f;
      }
    }
    hipLaunchKernelGGL(k_bd_force, dim3((d.n + 255) / 256), dim3(256), 0, st, d, step0 + (uint64_t)t + 1u, 1);
    hipLaunchKernelGGL(k_bd_gemv, dim3(rows, b->nchunk), dim3(kGemvRows), 0, st, d, b->ypart, b->chunk);
    hipLaunchKernelGGL(k_bd_update, dim3(rows), dim3(kGemvRows), 0, st, d, (const double *)b->ypart, b->nchunk,
                       c1, c2);
  }
  BCHK(c, hipGetLastError());
  BCHK(c, hipEventRecord(t1, st));
  BCHK(c, hipEventSynchronize(t1));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  b->last_total_ms = ms;
  b->last_fill_ms = fill_ms;
  b->last_factor_ms = fact_ms;
  return check_info(c, b);
}

int32_t tis_bdhi_timing(tis_ctx *c, double *total_ms, double *fill_ms, double *factor_ms) {
  if (!c) return TIS_ERR_ARG;
  BdState *b = need(c);
  if (!b) return TIS_ERR_STATE;
  if (total_ms) *total_ms = b->last_total_ms;
  if (fill_ms) *fill_ms = b->last_fill_ms;
  if (factor_ms) *factor_ms = b->last_factor_ms;
  return TIS_OK;
}

}  // extern "C"
