// tis_dev.h -- device-side data layout and helpers of libtis_hip (gfx950).
//
// Layout in HBM (one context = a batch of systems, beads numbered globally):
//   bead SoA:   x[2][3][N] FP64 (ping-pong: current / predicted positions),
//               px[2][3][N] (periodic systems only: wrapped positions),
//               v[3][N], a[3][N], f[3][N] FP64, a Langevin class byte,
//               per-bead type/unit/class bytes, exv radius/epsilon, charge.
//   incidence:  inc_off[N+1], inc[] -- the bonded terms each bead takes part
//               in, so that a bead's force is gathered by one lane in a fixed
//               order (no atomics: bitwise-reproducible, alone == in a batch).
//   exv list:   nb_cnt[N], nb[cap][N] slot-major (lane-coalesced); each entry
//               = (mirror code << 27) | j, both directions of every pair.
//   ele list:   one bit per (charge i, charge j) pair of a system, rows of
//               ceil(nq/32) words -- the reference's list at 1/32 B per pair.
//   hbonds:     per candidate: in-list flag, status, energy, force slots.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tis {

constexpr int kTile = 64;            // beads (or charges) per work tile = one wave
constexpr int kWaves = 4;            // waves per block of the force pass, k_lpair and the Ewald force pass
constexpr int kMirrorShift = 27;     // nb entry: (mirror << 27) | j
constexpr uint32_t kIdxMask = (1u << 27) - 1u;
constexpr double kDeMax = 50.0;      // const_physical.F90:35
constexpr double kPi = 3.14159265358979323846264338;
// k_B in kcal/mol/K: BOLTZ_J * JOUL2KCAL_MOL (const_physical.F90:21)
constexpr double kBoltzKcalMol = 1.38064852e-23 * ((1.0 / 4184.0) * 6.022140857e23);
constexpr double kHighEnergy = 1.0e10;

// incidence entry: type(3) | slot(3) | term(26)
enum IncType : uint32_t { INC_BOND = 0, INC_ANGLE = 1, INC_STACK = 2, INC_TSTACK = 3, INC_HBOND = 4 };
__host__ __device__ inline uint32_t inc_pack(uint32_t type, uint32_t slot, uint32_t term) {
  return (type << 29) | (slot << 26) | term;
}

struct DSys {
  int32_t off, nmp;          // beads [off, off+nmp)
  int32_t qoff, nq;          // charges [qoff, qoff+nq) in qbead[]
  int32_t hoff, nhb;         // hbond candidates
  int32_t toff, ntst;        // tertiary stacks
  int32_t soff, nst;         // consecutive stacks
  int32_t boff, nbd;         // bonds
  int32_t aoff, nba;         // angles
  int32_t sitoff, nsite;     // hbond sites
  int32_t hb_excess_possible; // some site has more candidates than its valence
  int32_t sptab, nspecies;   // offset and order of this system's species-pair table
  int32_t bt0, nbt, qt0, nqt; // its bead tiles and charge tiles
  double rl2max;             // largest exv list radius^2 over its unit pairs
  int32_t uoff, nunit;       // unit-pair list radius table rl2[uoff + a*nunit + b]
  int64_t moff;              // ele bit-matrix words
  int32_t mwords;            // words per row = ceil(nq/32)
  int64_t mioff;             // periodic: ele mirror bytes (nq*nq), else -1
  int32_t periodic;
  uint32_t key;              // identity in the random streams
  int32_t init_accel;        // mode-4 initialisation: 1 = new accelerations, 0 = keep
  int32_t lc0, nlc;          // its Langevin classes [lc0, lc0 + nlc)
  uint64_t step_origin;      // global step at which this system's step count is zero
  int32_t nsep[3];           // P, S, B
  int32_t ele_form, dh_type;
  double box[3];
  double exv_a, exv_dist, exv_dist_PS, exv_inf;
  double hb_cut, margin;
  double cutoff_ele, ele_coef, cdist, rcdist, ele_rl2, dh_cut2, dh_ecut2;
  double beta_hb, tempk;
  // plain Coulomb / Ewald (ele_form 1 / 2): cutoff_ele^2, 1/cutoff_ele
  double cl_cut2, cl_rcinv;
  // Ewald: alpha, 2 alpha / sqrt(pi), alpha^2, 1/L, 2 C (2 pi / L), self energy
  double ew_alpha, ew_beta, ew_a2, ew_invL, ew_fscale, ew_self;
  int32_t ewk0, nk;          // reciprocal vectors [ewk0, ewk0 + nk)
  int32_t nqc, nkc;          // charge chunks of the S pass, k chunks of the force pass
  int64_t psk_off, pf_off;   // partial structure factors [nqc][nk][2], partial forces [nkc][3][nq]
  // electrostatics by chunk-pair tasks: partial forces
  // [nqt(nqt+1)/2][2][3][kTile] at pp_off
  int64_t pp_off;
  // tile-pair list build: counts [nbt * nbt][kTile] at lp_off * kTile
  int64_t lp_off;
};

// Ewald work item: (system, tile a, chunk b)
struct EwTask { int32_t sys, a, b; };

// Beads of one system with the same friction and mass share their Langevin
// constants (time_integral_pre.F90:160-170): r[0] the noise amplitude c1,
// r[1..3] c2, c3, c4, and 1/m.
struct LangClass { double r[4], rcm, fric, mass, pad; };
constexpr int kEwKChunk = 256;   // reciprocal vectors per block of the force pass

struct Tile { int32_t sys, start; };

// bonds [b0, b0+nb), angles [a0, a0+na), consecutive stacks [s0, s0+ns);
// hbond lanes: entries [hd0, hd0+nhd) of DevPtrs::hb_dense
struct BondedRange { int32_t b0, nb, a0, na, s0, ns, hd0, nhd; };


// excluded-volume parameters of a species pair: D_ij (FP64 and its square)
// and eps_i*eps_j (force_exv_dt15.F90:60-80,92)
struct SpPair { double dij, dij2, coef; };

// axis-aligned bounding box of a tile of beads (or charges)
struct BBox { double lo[3], hi[3]; };

// All device pointers of a context, passed by value to kernels.
struct DevPtrs {
  const DSys *sys;
  int32_t nsys, ntot, nqtot;
  // beads
  double *x[2], *px[2];      // [3][ntot] each
  double *v, *a, *f;         // [3][ntot]
  LangClass *lcls;           // Langevin classes of every system (DSys::lc0 ..)
  const uint8_t *lcl;        // [ntot] class of each bead within its system
  const double *mass;
  const int8_t *mptype, *iclass;
  const int32_t *unit, *qidx;   // qidx: charge index (global) or -1
  const uint8_t *species;       // excluded-volume species of each bead (within its system)
  const SpPair *sptab;
  BBox *bbox, *qbbox;          // per bead tile / per charge tile
  BBox *sbox;                  // per 16-bead group of each bead tile [4 * nbtile]
  int *lcnt;                   // exv entry counts per (bead tile pair, lane) of the tile-pair build
  const double *exv_rad;
  const float *exv_eps;
  const double *charge;
  const double *rl2;
  // incidence: per bead the slots of the bonded terms it takes part in,
  // inc_slot[inc_off[i] .. inc_off[i+1]) in a fixed order.  Slots are stored
  // term-major by type, slot k of term t of a type at sb_<type> + K t + k
  // (K = 2, 3, 7, 6, 6 for bonds, angles, stacks, tertiary stacks, hbonds),
  // so that a term's lane writes its slots to consecutive addresses.
  const int32_t *inc_off;
  const int32_t *inc_slot;
  float *fslot;              // [slots][3] slot forces written by the term passes (summed in FP64)
  int64_t sb_bd, sb_ba, sb_st, sb_tst, sb_hb;
  // terms (global bead indices)
  const int32_t *bd_mp; const double *bd_nat, *bd_coef; int32_t nbd;
  const int32_t *ba_mp; const double *ba_nat, *ba_coef; int32_t nba;
  const int32_t *st_mp; const double *st_nat, *st_coef; int32_t nst;
  const int32_t *tst_mp; const double *tst_nat, *tst_coef; const int32_t *tst2st, *tst_excl; int32_t ntst;
  const int32_t *hb_mp; const double *hb_nat, *hb_coef; const int32_t *hb_site; const int32_t *hb_tert; int32_t nhb;
  const int32_t *hb_sys, *tst_sys;  // owning system of each candidate / tertiary stack
  const int32_t *site_val;
  // tertiary stacks that switch consecutive stack st off: st_ex[st_exoff[st] .. st_exoff[st+1])
  const int32_t *st_exoff, *st_ex;
  // per-step state of terms
  uint8_t *hb_inlist, *hb_status; // [nhb]
  double *hb_ene;            // [nhb]
  double *for_hb;            // [18][nhb]
  int32_t *site_cnt;         // [nsite]
  int32_t *hb_seq;           // [nhb] valence-pass scratch: a site's hbonds, per system at its hoff
  const int32_t *hb_dense;   // hbonds of the systems none of whose sites can exceed its valence, system order
  // lists
  int32_t *nb_cnt; uint32_t *nb; int32_t nb_cap;
  const int32_t *qbead;      // [nq] global bead of each charge
  double *pp;                // partial forces of the electrostatic chunk-pair tasks
  uint32_t *qmask;           // bit matrix
  uint8_t *qmirror;          // periodic ele mirror codes
  // Ewald reciprocal space
  const double *ewh;         // [3][nktot] integer h as FP64
  const double *ewc;         // [nktot] f(h)
  int32_t nktot;
  double *psk;               // partial structure factors
  double *pf;                // partial reciprocal forces
  // displacement-triggered rebuilds (time_integral.F90:57-93): displacement
  // since the last rebuild and of the last step per bead, the two largest
  // squared displacements of each bead tile, and the rebuild decision of
  // the current step per system
  int32_t dyn;
  double *dxyz, *dlast;      // [3][ntot]
  double *ttop;              // [nbtile][2]
  uint8_t *sys_flag;         // [nsys]
  // counters / timing
  unsigned long long *counters;  // TIS_CNT_*
  uint64_t *stepbase;
  uint64_t seed;
  double tstep2;
  double tstep;
};

// temperature-dependent fields of a system descriptor (simu_para2.F90:8-108)
__host__ __device__ inline void copy_t_fields(DSys &d, const DSys &s) {
  d.tempk = s.tempk; d.ele_coef = s.ele_coef; d.cdist = s.cdist; d.rcdist = s.rcdist; d.ele_rl2 = s.ele_rl2;
  d.dh_cut2 = s.dh_cut2; d.dh_ecut2 = s.dh_ecut2; d.beta_hb = s.beta_hb;
  d.ew_fscale = s.ew_fscale; d.ew_self = s.ew_self;
}

// parameter sets of temperature labels (replica exchange)
struct LabelDev {
  const DSys *sys;
  const double *q;   const int64_t *qoff;    // per-bead charges of each label
  const double *u0;  const int64_t *u0off;   // per-stack U0 of each label
};

// ---------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon et al., SC'11); identical to the oracle's definition.
__device__ inline void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}
__device__ inline double u01_53(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)a << 21) ^ ((uint64_t)b >> 11);
  return ((double)m + 0.5) * 0x1.0p-53;
}
__device__ inline void philox_at(uint64_t seed, uint32_t sys, uint64_t step, uint32_t lane,
                                 uint32_t tag, uint32_t out[4]) {
  out[0] = lane; out[1] = (uint32_t)step; out[2] = sys;
  out[3] = ((uint32_t)(step >> 32) & 0xFFFFu) | (tag << 16);
  philox4x32(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}
// Box-Muller N(0,1) triple for (system key, step, bead)
__device__ inline void gauss3(uint64_t seed, uint32_t sys, uint64_t step, uint32_t bead, double g[3]) {
  uint32_t r[4];
  philox_at(seed, sys, step, bead, 0u, r);
  double u1 = u01_53(r[0], r[1]), u2 = u01_53(r[2], r[3]);
  double rad = sqrt(-2.0 * log(u1));
  double th = (2.0 * kPi) * u2;
  double sn, cs;
  sincos(th, &sn, &cs);
  g[0] = rad * cs;
  g[1] = rad * sn;
  philox_at(seed, sys, step, bead, 1u, r);
  u1 = u01_53(r[0], r[1]); u2 = u01_53(r[2], r[3]);
  rad = sqrt(-2.0 * log(u1));
  th = (2.0 * kPi) * u2;
  g[2] = rad * cos(th);
}
__device__ inline double uniform01(uint64_t seed, uint32_t sys, uint64_t step, uint32_t stream,
                                   uint32_t idx) {
  uint32_t r[4];
  philox_at(seed, sys, step, idx, 2u + stream, r);
  return u01_53(r[0], r[1]);
}

// ---------------------------------------------------------------- exact FP64
// Cutoff tests reproduce the reference's left-to-right v1**2+v2**2+v3**2
// with two roundings per a*b+c (no fused multiply-add).
__device__ inline double sq3_exact(double a, double b, double c) {
  return __dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)), __dmul_rn(c, c));
}
// (a^2 + b^2) + c^2 with each product and sum rounded on its own (no fused
// multiply-add), the reference's evaluation of |v|^2
__device__ inline double sq3_separate(double a, double b, double c) {
#pragma clang fp contract(off)
  return (a * a + b * b) + c * c;
}

// util_pbneighbor.F90:22-35; returns mirror code 9*i1+3*i2+i3 (0-based)
__device__ inline int pbneighbor(const DSys &s, double v[3]) {
  int code = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double L = s.box[d], h = L / 2.0;
    int imi;
    if (v[d] > h) { v[d] = __dadd_rn(v[d], -L); imi = 0; }
    else if (v[d] < -h) { v[d] = __dadd_rn(v[d], L); imi = 1; }
    else imi = 2;
    code = code * 3 + imi;
  }
  return code;
}
// shift of mirror code (setp_periodic.F90:80-115)
__device__ inline double mirror_shift(const DSys &s, int code, int d) {
  int i = d == 0 ? code / 9 : (d == 1 ? (code / 3) % 3 : code % 3);
  // -L, +L or +0 for i = 0, 1, 2, without a branch
  return (double)((i == 1) - (i == 0)) * s.box[d];
}
// the same pair seen from the other bead: swap -L and +L per component
__device__ inline int mirror_flip(int code) {
  int c0 = code / 9, c1 = (code / 3) % 3, c2 = code % 3;
  c0 = c0 == 2 ? 2 : 1 - c0; c1 = c1 == 2 ? 2 : 1 - c1; c2 = c2 == 2 ? 2 : 1 - c2;
  return 9 * c0 + 3 * c1 + c2;
}

}  // namespace tis
