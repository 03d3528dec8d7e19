// tis_capi.hip -- host side of libtis_hip: context, batch upload into the
// HBM layout of tis_dev.h, and the C-ABI of include/tis_hip.h.  Kernel
// launches are stream-ordered on the context's own stream (default
// priority); the batched Langevin loop is captured once into a hipGraph and
// replayed.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/tis_hip.h"
#include "tis_dev.h"
#include "tis_internal.h"

namespace tis {
template <int PPW, int TPL>
__global__ void k_force(DevPtrs, const double *, const double *, const int32_t *, int, const EwTask *, int,
                        BondedRange, int32_t);
__global__ void k_ewald_sk(DevPtrs, const double *, const EwTask *, int);
__global__ void k_ewald_f(DevPtrs, const double *, const EwTask *);
__global__ void k_main(DevPtrs, const double *, const double *, double *, double *, const Tile *, int, int32_t, int, int,
                       int);
__global__ void k_prep(DevPtrs, const double *, double *, double *, const Tile *, int, int, const Tile *, int, int,
                       int, int);
__global__ void k_lists(DevPtrs, const double *, const double *, const Tile *, int, int, const Tile *, int, int, int,
                        int, int);
__global__ void k_lpair(DevPtrs, const double *, const double *, const EwTask *, int, const EwTask *, int, int, int,
                        int, int);
__global__ void k_advance(uint64_t *, uint64_t, unsigned long long *);
__global__ void k_count_rebuild(unsigned long long *);
__global__ void k_energy(DevPtrs, const double *, const double *, const int32_t *, double *, int);
__global__ void k_label_apply(DevPtrs, LabelDev, const int32_t *, const int32_t *, const double *);
__global__ void k_stack_status(DevPtrs, const double *, const double *, int, int, int32_t *);
}  // namespace tis

using namespace tis;

namespace {


struct SysCopy;

// Host staging of an arena: the used prefix of every array, global indices
// (slot positions of terms kept term-major here, [term][slot]).
struct Stage {
  int32_t nbead = 0, nq = 0, nbd = 0, nba = 0, nst = 0, ntst = 0, nhb = 0, nsite = 0;
  int64_t moff = 0, mioff = 0, pp_off = 0, lp_off = 0, psk_off = 0, pf_off = 0;
  std::vector<double> rl2, exv_rad, charge, mass;
  std::vector<int8_t> mptype, iclass;
  std::vector<int32_t> unit, qidx, qbead, inc_off{0};
  std::vector<uint32_t> inc;
  std::vector<uint8_t> lcl, species;
  std::vector<LangClass> lcls;
  std::vector<float> exv_eps;
  std::vector<SpPair> sptab;
  std::vector<int32_t> bd_mp, ba_mp, st_mp, tst_mp, tst2st, tst_excl, tst_sys, hb_mp, hb_site, hb_tert, hb_sys,
      site_val;
  std::vector<double> bd_nat, bd_coef, ba_nat, ba_coef, st_nat, st_coef, tst_nat, tst_coef, hb_nat, hb_coef;
  std::vector<int32_t> st_exoff{0}, st_ex;
  std::vector<double> ewhx, ewhy, ewhz, ewc;
};

// Sizes of a Stage's arrays at one moment (where an append starts).
struct StageMarks {
  size_t bead = 0, q = 0, bd = 0, ba = 0, st = 0, tst = 0, hb = 0, site = 0, rl2 = 0, inc = 0, lc = 0, spt = 0, k = 0,
         ex = 0;
};

// Capacities of the arena's arrays (elements; strides of the strided ones).
struct Caps {
  int64_t bead = 1, q = 1, bd = 1, ba = 1, st = 1, tst = 1, hb = 1, site = 1, rl2 = 1, inc = 1, lc = 1, spt = 1,
          k = 1, m = 1, mi = 1, pp = 1, lp = 1, psk = 1, pf = 1, ex = 1, nsys = 1;
  bool periodic = false;
};

// Writable device pointers of the arena's uploaded arrays.
struct ArenaBufs {
  DSys *sys = nullptr;
  uint8_t *lcl = nullptr, *species = nullptr;
  LangClass *lcls = nullptr;
  double *mass = nullptr, *exv_rad = nullptr, *charge = nullptr, *rl2 = nullptr;
  int8_t *mptype = nullptr, *iclass = nullptr;
  int32_t *unit = nullptr, *qidx = nullptr, *qbead = nullptr, *inc_off = nullptr;
  int32_t *inc_slot = nullptr;
  float *exv_eps = nullptr;
  SpPair *sptab = nullptr;
  int32_t *bd_mp = nullptr, *ba_mp = nullptr, *st_mp = nullptr, *tst_mp = nullptr, *hb_mp = nullptr;
  double *bd_nat = nullptr, *bd_coef = nullptr, *ba_nat = nullptr, *ba_coef = nullptr, *st_nat = nullptr,
         *st_coef = nullptr, *tst_nat = nullptr, *tst_coef = nullptr, *hb_nat = nullptr, *hb_coef = nullptr;
  int32_t *tst2st = nullptr, *tst_excl = nullptr, *tst_sys = nullptr, *hb_site = nullptr, *hb_tert = nullptr,
          *hb_sys = nullptr, *site_val = nullptr, *st_exoff = nullptr, *st_ex = nullptr, *ststat = nullptr;
  double *ewh = nullptr, *ewc = nullptr, *energy = nullptr;
  unsigned long long *counters = nullptr;
  uint64_t *stepbase = nullptr;
};

struct HostSys {
  DSys d;
  bool live = false;                    // in the launch tables
  bool stored = false;                  // its segments are in the arena
  int32_t live_pos = -1;                // index in the live list
  std::shared_ptr<SysCopy> src;         // host copy (for a new layout)
  int32_t bt_begin = 0, bt_count = 0;   // bead tiles
  int32_t qt_begin = 0, qt_count = 0;   // charge tiles
  bool has_task = false;                // tertiary stacks or valence-limited hbonds (terms_wave)
  int32_t hd_begin = 0, hd_count = 0;   // its entries of the hbond-lane list
  int32_t esk_begin = 0, esk_count = 0;  // Ewald structure-factor tasks
  int32_t ef_begin = 0, ef_count = 0;    // Ewald force tasks
  int32_t pt_begin = 0, pt_count = 0;    // periodic electrostatic chunk-pair tasks
  int32_t bp_begin = 0, bp_count = 0;    // bead tile pairs of the tile-pair list build
  int32_t qp_begin = 0, qp_count = 0;    // charge tile pairs
  std::vector<double> mass, fric;
  std::vector<double> lc_fric, lc_mass;   // its Langevin classes
  std::vector<int32_t> stack_ids;       // global stack indices
};

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
};

}  // namespace

constexpr int kTev = 6;  // timing events per eager step

struct tis_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority): the term passes run beside the
  // electrostatic pass; fork/join by events (captured into the graphs too)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::string err;
  // batch
  int32_t nsys = 0, nbeads = 0, nqtot = 0, nhbtot = 0, ntsttot = 0, nsttot = 0;  // handles; used counts
  int32_t nlive = 0;                     // systems in the launch tables
  std::vector<HostSys> hs;
  std::vector<void *> allocs;            // the arena (and list capacity growth)
  std::vector<void *> tab_allocs;        // the launch tables
  size_t dev_bytes = 0;                  // bytes held in allocs and tab_allocs
  size_t tab_bytes = 0;
  Stage stage;
  Caps caps;
  ArenaBufs ab{};
  bool have_arena = false;
  int64_t relayouts = 0;
  DevPtrs P{};
  DSys *d_sys = nullptr;
  Tile *d_btiles = nullptr, *d_qtiles = nullptr;
  Tile *d_btiles_xcd = nullptr;          // the bead tiles in k_main's whole-batch launch order
  int32_t nbtile = 0, nqtile = 0;
  int32_t *d_systask = nullptr, *d_sysidx = nullptr;
  int32_t nsystask = 0;
  int32_t *d_hbdense = nullptr;          // hbond lanes (DevPtrs::hb_dense)
  int32_t nhbdense = 0;
  std::vector<int32_t> systask_of;       // index into d_systask per system or -1
  int32_t *d_ststat = nullptr;           // st_status output of tis_term_status [nsttot]
  double *d_energy = nullptr;
  int32_t *d_eset = nullptr;             // tis_energy_sumup_set: system list and energies
  double *d_eset_out = nullptr;
  size_t eset_cap = 0, eset_out_cap = 0;
  double *d_eunit = nullptr;   // raw per-unit-pair energies [nunit^2][25]
  size_t eunit_cap = 0;
  unsigned long long *d_counters = nullptr;
  uint64_t *d_stepbase = nullptr;
  bool any_ele = false, any_periodic = false;
  EwTask *d_ewsk = nullptr, *d_ewf = nullptr, *d_ptask = nullptr, *d_ptask_lpt = nullptr, *d_bpair = nullptr,
         *d_qpair = nullptr;
  int32_t n_ewsk = 0, n_ewf = 0, n_ptask = 0, n_bpair = 0, n_qpair = 0;
  // state
  int cur = 0;
  bool lists_valid = false;
  // Langevin
  bool lv_init = false;
  double tstep = 0.0;
  int32_t rebuild_every = 10;
  int64_t t_since_init = 0;
  uint64_t step_host = 0;                // the device step counter (stepbase) as the host knows it
  // payload bytes since creation: uploads of topology and launch tables,
  // state uploads (tis_state_set), state downloads (tis_state_get), state
  // carried device to device by a new arena layout
  int64_t h2d_bytes = 0, h2d_state = 0, d2h_bytes = 0, d2d_state = 0;
  // launch shape of the force pass (tis_tune): chunk pairs per wave, bonded
  // terms per lane, off-diagonal chunk pairs first (1) or system order (0)
  struct Shape { int ppw = 1, tpl = 1, lpt = 1; } shape;
  std::vector<double> tune_ms;           // k_force ms per candidate of the last tis_tune
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_len = 0;
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_total_ms = 0.0, last_main_ms = 0.0, last_pre_ms = 0.0, last_force_ms = 0.0;
  int64_t last_nlaunch = 0;
  // per-kernel events of eager runs (kTev per step, see launch_force_pair)
  bool timing = false;
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  // Brownian dynamics with hydrodynamic interaction (tis_bdhi.hip)
  tis::BdState *bd = nullptr;
  // temperature labels (replica exchange)
  struct Label { bool set = false; DSys d{}; int32_t nmp = 0, nst = 0; std::vector<double> q, u0; };
  std::vector<Label> labels;
  bool labels_dirty = false;
  std::vector<void *> label_bufs;         // device copies of the label sets
  LabelDev label_dev{};
  int32_t *d_lab_args = nullptr;          // [2][cap] isys, label
  double *d_lab_scale = nullptr;          // [cap]
  int32_t lab_cap = 0;
};

namespace tis {
hipStream_t ctx_stream(tis_ctx *c) { return c->stream; }
int ctx_device(tis_ctx *c) { return c->device; }
BdState *&ctx_bd(tis_ctx *c) { return c->bd; }
int ctx_fail(tis_ctx *c, int code, const char *msg) {
  if (c) c->err = msg;
  return code;
}
}  // namespace tis

namespace {

int fail(tis_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                               \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, TIS_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

template <typename T>
int dalloc(tis_ctx *c, T **p, size_t n) {
  *p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)p, bytes);
  if (e != hipSuccess) return fail(c, TIS_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  c->allocs.push_back((void *)*p);
  c->dev_bytes += bytes;
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e != hipSuccess) return fail(c, TIS_ERR_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  return 0;
}

// Allocation into a given list; the bytes count in *acc (or the context's
// total).  The memory starts zeroed.
template <typename T>
int dalloc_in(tis_ctx *c, std::vector<void *> &list, T **p, size_t n, size_t *acc = nullptr) {
  *p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)p, bytes);
  if (e != hipSuccess) return fail(c, TIS_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  list.push_back((void *)*p);
  *(acc ? acc : &c->dev_bytes) += bytes;
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e != hipSuccess) return fail(c, TIS_ERR_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  return 0;
}

template <typename T>
int dupload_in(tis_ctx *c, std::vector<void *> &list, T **p, const std::vector<T> &h, size_t *acc = nullptr) {
  int rc = dalloc_in(c, list, p, h.size(), acc);
  if (rc) return rc;
  if (!h.empty()) HIPCHK(c, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  c->h2d_bytes += (int64_t)(h.size() * sizeof(T));
  return 0;
}

template <typename T>
int dupload(tis_ctx *c, T **p, const std::vector<T> &h) {
  int rc = dalloc(c, p, h.size());
  if (rc) return rc;
  if (!h.empty()) HIPCHK(c, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  c->h2d_bytes += (int64_t)(h.size() * sizeof(T));
  return 0;
}

void free_all(tis_ctx *c) {
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 2; ++k)
    if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  for (void *p : c->allocs) (void)hipFree(p);
  for (void *p : c->tab_allocs) (void)hipFree(p);
  c->allocs.clear();
  c->tab_allocs.clear();
  c->dev_bytes = c->tab_bytes = 0;
  c->hs.clear();
  c->P = DevPtrs{};
  c->ab = ArenaBufs{};
  c->stage = Stage{};
  c->caps = Caps{};
  c->have_arena = false;
  c->d_eunit = nullptr;
  c->eunit_cap = 0;
  c->nsys = c->nbeads = c->nlive = 0;
  c->lists_valid = false;
  c->lv_init = false;
}

void fill_derived(DSys &d, const tis_system &s) {
  const tis_params &p = s.p;
  d.periodic = p.periodic;
  for (int k = 0; k < 3; ++k) d.box[k] = p.box[k];
  d.nsep[0] = p.n_sep_P; d.nsep[1] = p.n_sep_S; d.nsep[2] = p.n_sep_B;
  d.ele_form = p.ele_form; d.dh_type = p.dh_cutoff_type;
  d.exv_a = p.exv_adjust; d.exv_dist = p.exv_dist; d.exv_dist_PS = p.exv_dist_PS; d.exv_inf = p.exv_inf;
  d.hb_cut = p.hb_cutoff_dist; d.margin = p.neigh_margin;
  d.cutoff_ele = p.cutoff_ele; d.ele_coef = p.ele_coef; d.cdist = p.ele_cdist;
  d.rcdist = 1.0 / p.ele_cdist;
  // list radius, neighbor_list_ele.F90:73-90
  double r;
  if (p.ele_form == TIS_ELE_DH && p.dh_cutoff_type == 0) r = p.neigh_margin + p.cutoff_ele * p.ele_cdist;
  else r = p.neigh_margin + p.cutoff_ele;
  d.ele_rl2 = r * r;
  double c = p.cutoff_ele * p.ele_cdist;
  d.dh_cut2 = c * c;  // force cutoff ignores i_DH_cutoff_type (force_ele_DH.F90:44)
  d.dh_ecut2 = p.dh_cutoff_type == 0 ? c * c : p.cutoff_ele * p.cutoff_ele;  // energy_ele_DH.F90:33-37
  d.beta_hb = p.beta_hb;
  d.tempk = s.tempk;
  d.cl_cut2 = p.cutoff_ele * p.cutoff_ele;
  d.cl_rcinv = p.cutoff_ele > 0.0 ? 1.0 / p.cutoff_ele : 0.0;
  if (p.ele_form == TIS_ELE_EWALD) {
    const double a = p.ewld_alpha, L = p.box[0];
    d.ew_alpha = a;
    d.ew_beta = 2.0 * a / std::sqrt(kPi);
    d.ew_a2 = a * a;
    d.ew_invL = 1.0 / L;
    d.ew_fscale = 2.0 * p.ele_coef * (2.0 * kPi / L);
    // self term, mloop_ewld.F90:146-154 (charges of this system)
    double q2 = 0.0;
    for (int i = 0; i < s.nmp; ++i)
      if (std::fabs(s.charge[i]) >= 1.0e-6) q2 += s.charge[i] * s.charge[i];
    d.ew_self = p.ele_coef * (-a / std::sqrt(kPi)) * q2;
  }
}

// Reciprocal vectors of mloop_ewld.F90:52-136: integer h != 0 with
// |h|^2 <= hmax^2 ordered by |h|^2 (ties in generation order), and
// f(h) = 2 pi / L^3 exp(-k^2 / 4 alpha^2) / k^2, k = 2 pi h / L.
void ewald_kvectors(const tis_params &p, std::vector<double> &hx, std::vector<double> &hy,
                    std::vector<double> &hz, std::vector<double> &coef) {
  const int hmax = p.ewld_hmax;
  const double L = p.box[0], a = p.ewld_alpha, factor = 2.0 * kPi / L;
  for (int r2 = 1; r2 <= hmax * hmax; ++r2)
    for (int ix = -hmax; ix <= hmax; ++ix)
      for (int iy = -hmax; iy <= hmax; ++iy)
        for (int iz = -hmax; iz <= hmax; ++iz) {
          if (ix * ix + iy * iy + iz * iz != r2) continue;
          hx.push_back(ix); hy.push_back(iy); hz.push_back(iz);
          coef.push_back(2.0 * kPi / (L * L * L) * std::exp(-factor * factor * (double)r2 / (4.0 * a * a)) /
                         (factor * factor * (double)r2));
        }
}

// Langevin constants, time_integral_pre.F90:160-170
void langevin_consts(double tstep, double tempk, double fric, double mass, double r[4]) {
  double h = 0.5 * tstep * fric;
  double u1 = 1.0 - h, u2 = 1.0 - h + h * h;
  r[0] = std::sqrt(2.0 * fric * kBoltzKcalMol * tempk / (tstep * mass));
  r[1] = u1 * u2;
  r[2] = (0.5 * tstep) * u1;
  r[3] = tstep * u1;
}

int check_sys(tis_ctx *c, int32_t isys) {
  if (c->nsys == 0) return fail(c, TIS_ERR_STATE, "no systems uploaded");
  if (isys < -1 || isys >= c->nsys) return fail(c, TIS_ERR_ARG, "system index %d out of range", isys);
  if (isys >= 0 && !c->hs[isys].live) return fail(c, TIS_ERR_ARG, "system %d is retired", isys);
  if (c->nlive == 0) return fail(c, TIS_ERR_STATE, "no live systems");
  return 0;
}

double *Xc(tis_ctx *c) { return c->P.x[c->cur]; }
double *PXc(tis_ctx *c) { return c->P.px[c->cur]; }

// --- launch helpers on a system subset: isys = -1 all, else one system
struct Range {
  bool all;                 // the whole batch
  int isys;                 // else this system
  int bt0, btn, qt0, qtn, task0, taskn;
  int hd0, hdn;             // hbond lanes
  const int32_t *sysidx; int nsel;
  int b0, nb, a0, na, s0, ns;
  int ek0, ekn, ef0, efn;   // Ewald tasks
  int pt0, ptn;             // periodic electrostatic chunk-pair tasks
};

// Whole-batch launch order of the bead tiles for k_main: workgroup b runs on
// XCD b mod 8, each with its own L2, and a tile reads the coordinates of its
// system's beads (the exv neighbours), so every system's tiles go to one XCD
// (systems dealt largest first to the XCD with the fewest tiles) and that
// L2 holds one eighth of the coordinates rather than all of them.  The
// order changes no result: a tile's work depends on the tile only.
static std::vector<Tile> xcd_order(const std::vector<Tile> &tiles, int nsys) {
  constexpr int kXcd = 8;
  std::vector<int> ntile(nsys, 0);
  for (const Tile &t : tiles) ++ntile[t.sys];
  std::vector<int> sys_order(nsys);
  for (int i = 0; i < nsys; ++i) sys_order[i] = i;
  std::stable_sort(sys_order.begin(), sys_order.end(), [&](int a, int b) { return ntile[a] > ntile[b]; });
  std::vector<int> xcd_of(nsys, 0), load(kXcd, 0);
  for (int sy : sys_order) {
    if (!ntile[sy]) continue;
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    xcd_of[sy] = x;
    load[x] += ntile[sy];
  }
  std::vector<std::vector<Tile>> bin(kXcd);
  for (const Tile &t : tiles) bin[xcd_of[t.sys]].push_back(t);
  const size_t nt = tiles.size();
  std::vector<Tile> out(nt);
  std::vector<char> used(nt, 0);
  std::vector<Tile> rest;
  std::vector<size_t> next(kXcd, 0);
  for (size_t p = 0; p < nt; ++p) {
    std::vector<Tile> &b = bin[p % kXcd];
    size_t &k = next[p % kXcd];
    if (k < b.size()) { out[p] = b[k++]; used[p] = 1; }
  }
  for (int x = 0; x < kXcd; ++x)
    for (size_t k = next[x]; k < bin[x].size(); ++k) rest.push_back(bin[x][k]);
  size_t r = 0;
  for (size_t p = 0; p < nt; ++p)
    if (!used[p]) out[p] = rest[r++];
  return out;
}

Range range_of(tis_ctx *c, int32_t isys) {
  Range r{};
  if (isys < 0) {
    r.all = true;
    r.bt0 = 0; r.btn = c->nbtile; r.qt0 = 0; r.qtn = c->nqtile; r.task0 = 0; r.taskn = c->nsystask;
    r.hd0 = 0; r.hdn = c->nhbdense;
    r.sysidx = c->d_sysidx; r.nsel = c->nlive;
    r.b0 = 0; r.nb = c->stage.nbd; r.a0 = 0; r.na = c->stage.nba; r.s0 = 0; r.ns = c->stage.nst;
    r.ek0 = 0; r.ekn = c->n_ewsk; r.ef0 = 0; r.efn = c->n_ewf;
    r.pt0 = 0; r.ptn = c->n_ptask;
  } else {
    const HostSys &h = c->hs[isys];
    r.isys = isys;
    r.bt0 = h.bt_begin; r.btn = h.bt_count; r.qt0 = h.qt_begin; r.qtn = h.qt_count;
    r.task0 = c->systask_of[isys] < 0 ? 0 : c->systask_of[isys];
    r.taskn = c->systask_of[isys] < 0 ? 0 : 1;
    r.hd0 = h.hd_begin; r.hdn = h.hd_count;
    r.sysidx = c->d_sysidx + h.live_pos; r.nsel = 1;
    r.b0 = h.d.boff; r.nb = h.d.nbd; r.a0 = h.d.aoff; r.na = h.d.nba; r.s0 = h.d.soff; r.ns = h.d.nst;
    r.ek0 = h.esk_begin; r.ekn = h.esk_count; r.ef0 = h.ef_begin; r.efn = h.ef_count;
    r.pt0 = h.pt_begin; r.ptn = h.pt_count;
  }
  return r;
}

static void launch_lpair(tis_ctx *c, const Range &r, const double *X, const double *PX, int mode, int h0, int nh) {
  const int bp0 = r.all ? 0 : c->hs[r.isys].bp_begin, nbp = r.all ? c->n_bpair : c->hs[r.isys].bp_count;
  const int qp0 = r.all ? 0 : c->hs[r.isys].qp_begin, nqp = r.all ? c->n_qpair : c->hs[r.isys].qp_count;
  const int nbb = (nbp + kWaves - 1) / kWaves, nqb = (nqp + kWaves - 1) / kWaves;
  const int nhb = (nh + kTile * kWaves - 1) / (kTile * kWaves);
  if (nbb + nqb + nhb > 0)
    hipLaunchKernelGGL(k_lpair, dim3(nbb + nqb + nhb), dim3(kTile * kWaves), 0, c->stream, c->P, X, PX,
                       c->d_bpair + bp0, nbp, c->d_qpair + qp0, nqp, h0, nh, mode, 0);
  if (nbb > 0)
    hipLaunchKernelGGL(k_lpair, dim3(nbb), dim3(kTile * kWaves), 0, c->stream, c->P, X, PX, c->d_bpair + bp0, nbp,
                       c->d_qpair + qp0, nqp, h0, 0, mode, 1);
}

// The lists of the range from X (wrapped copies into PXcur, predicted
// wrapped positions into PXpred when with_pred): k_prep, then k_lists by
// bead / charge tile for non-periodic systems and the hbond candidates, and
// k_lpair by tile pair for periodic systems; with gate, only the systems
// whose displacement trigger fired.
int launch_lists(tis_ctx *c, const Range &r, const double *X, double *PXcur, double *PXpred, int with_pred,
                 int gate) {
  const int h0 = r.all ? 0 : c->hs[r.isys].d.hoff, nh = r.all ? c->nhbtot : c->hs[r.isys].d.nhb;
  const bool periodic = r.all ? c->any_periodic : c->hs[r.isys].d.periodic != 0;
  if (gate && !periodic) {
    // displacement-triggered rebuilds of non-periodic systems: k_lists takes
    // the decision and the displacement reset itself and culls by the boxes
    // k_main leaves at the end of each step (gate 2), so k_prep is skipped;
    // each system's bead tiles also rebuild its hbond candidates
    if (r.btn + r.qtn > 0)
      hipLaunchKernelGGL(k_lists, dim3(r.btn + r.qtn), dim3(kTile), 0, c->stream, c->P, X, PXcur, c->d_btiles, r.bt0,
                         r.btn, c->d_qtiles, r.qt0, r.qtn, h0, nh, 2);
    return 0;
  }
  if (r.btn + r.qtn > 0)
    hipLaunchKernelGGL(k_prep, dim3(r.btn + r.qtn), dim3(kTile), 0, c->stream, c->P, X, PXcur, PXpred, c->d_btiles,
                       r.bt0, r.btn, c->d_qtiles, r.qt0, r.qtn, with_pred, gate);
  const int nblk = r.btn + r.qtn + (nh + kTile - 1) / kTile;
  if (nblk > 0)
    hipLaunchKernelGGL(k_lists, dim3(nblk), dim3(kTile), 0, c->stream, c->P, X, PXcur, c->d_btiles, r.bt0, r.btn,
                       c->d_qtiles, r.qt0, r.qtn, h0, nh, gate);
  if (periodic) launch_lpair(c, r, X, PXcur, gate ? 3 : 2, h0, 0);
  return 0;
}

}  // namespace

// =====================================================================
// Index arrays of one system against their bounds (0-based): bead indices
// of every term in [0, nmp), units in [0, nunit), hbond sites in
// [-1, nhbsite), the tertiary-to-consecutive stack map in [-1, nst).  The
// message goes to *err.
int vfail(std::string *err, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (err) *err = buf;
  return TIS_ERR_ARG;
}

int check_range(std::string *err, int is, const char *what, const int32_t *a, int64_t n, int32_t lo,
                int32_t hi) {
  if (n > 0 && !a) return vfail(err, "system %d: %s is NULL", is, what);
  for (int64_t k = 0; k < n; ++k)
    if (a[k] < lo || a[k] >= hi)
      return vfail(err, "system %d: %s[%lld] = %d outside [%d, %d)", is, what, (long long)k, a[k], lo, hi);
  return TIS_OK;
}

int validate_system(std::string *err, int is, const tis_system &s) {
  if (s.nmp <= 0 || s.nunit <= 0) return vfail(err, "system %d has no beads or units", is);
  if (s.nbd < 0 || s.nba < 0 || s.nst < 0 || s.ntst < 0 || s.nhb < 0 || s.nhbsite < 0)
    return vfail(err, "system %d: negative term count", is);
  const void *bead_arrays[] = {s.unit, s.mptype, s.iclass, s.exv_rad, s.exv_eps, s.charge, s.mass, s.fric,
                               s.rlist2_unit};
  for (const void *a : bead_arrays)
    if (!a) return vfail(err, "system %d: a per-bead array is NULL", is);
  int rc;
  if ((rc = check_range(err, is, "unit", s.unit, s.nmp, 0, s.nunit))) return rc;
  if ((rc = check_range(err, is, "bd_mp", s.bd_mp, 2 * (int64_t)s.nbd, 0, s.nmp))) return rc;
  if ((rc = check_range(err, is, "ba_mp", s.ba_mp, 3 * (int64_t)s.nba, 0, s.nmp))) return rc;
  if ((rc = check_range(err, is, "st_mp", s.st_mp, 7 * (int64_t)s.nst, 0, s.nmp))) return rc;
  if ((rc = check_range(err, is, "tst_mp", s.tst_mp, 6 * (int64_t)s.ntst, 0, s.nmp))) return rc;
  if ((rc = check_range(err, is, "tst2st", s.tst2st, 2 * (int64_t)s.ntst, -1, s.nst))) return rc;
  if ((rc = check_range(err, is, "tst_excl", s.tst_excl, 2 * (int64_t)s.ntst, 0, 2))) return rc;
  if ((rc = check_range(err, is, "hb_mp", s.hb_mp, 6 * (int64_t)s.nhb, 0, s.nmp))) return rc;
  if ((rc = check_range(err, is, "hb_site", s.hb_site, 6 * (int64_t)s.nhb, -1, s.nhbsite))) return rc;
  if (s.nhbsite > 0 && !s.hbsite_valence) return vfail(err, "system %d: hbsite_valence is NULL", is);
  const void *nb[] = {s.bd_nat, s.bd_coef}, *na[] = {s.ba_nat, s.ba_coef}, *ns[] = {s.st_nat, s.st_coef},
             *nt[] = {s.tst_nat, s.tst_coef}, *nh[] = {s.hb_nat, s.hb_coef};
  for (int k = 0; k < 2; ++k)
    if ((s.nbd && !nb[k]) || (s.nba && !na[k]) || (s.nst && !ns[k]) || (s.ntst && !nt[k]) || (s.nhb && !nh[k]))
      return vfail(err, "system %d: a term parameter array is NULL", is);
  return TIS_OK;
}

// =====================================================================
// The batch arena.  A context's systems live in one set of device arrays:
// every per-bead, per-charge and per-term array is a concatenation of
// per-system segments, and a system keeps its offsets for as long as it
// stays.  tis_system_append stages new systems after the last one and
// uploads only their own segments when they fit the capacities;
// tis_system_retire drops a system from the launch tables (its segments
// stay until the arena is laid out again).  A new layout (compaction or
// growth) stages the live systems from the host copies kept of them and
// carries their positions, velocities and accelerations over on the
// device.  Strided arrays (x[3][*], nb[cap][*], for_hb[18][*], the term
// slot positions [slot][*], reciprocal vectors [3][*]) use the capacity as
// their stride, so an append moves nothing already there.
namespace {

template <typename T>
const T *keep(std::vector<T> &v, const T *src, int64_t n) {
  if (!src) return nullptr;
  v.assign(src, src + std::max<int64_t>(n, 0));
  return v.empty() ? src : v.data();
}

struct SysCopy {
  tis_system s{};
  std::vector<int32_t> unit, mptype, iclass, bd_mp, ba_mp, st_mp, tst_mp, tst2st, tst_excl, hb_mp, hb_site, hb_tert,
      valence;
  std::vector<double> exv_rad, exv_eps, charge, rl2, bd_nat, bd_coef, ba_nat, ba_coef, st_nat, st_coef, tst_nat,
      tst_coef, hb_nat, hb_coef, mass, fric, radius;
};

std::shared_ptr<SysCopy> copy_system(const tis_system &in) {
  auto c = std::make_shared<SysCopy>();
  tis_system &s = c->s;
  s = in;
  const int64_t n = in.nmp, nbd = in.nbd, nba = in.nba, nst = in.nst, ntst = in.ntst, nhb = in.nhb;
  s.unit = keep(c->unit, in.unit, n);
  s.mptype = keep(c->mptype, in.mptype, n);
  s.iclass = keep(c->iclass, in.iclass, n);
  s.exv_rad = keep(c->exv_rad, in.exv_rad, n);
  s.exv_eps = keep(c->exv_eps, in.exv_eps, n);
  s.charge = keep(c->charge, in.charge, n);
  s.rlist2_unit = keep(c->rl2, in.rlist2_unit, (int64_t)in.nunit * in.nunit);
  s.bd_mp = keep(c->bd_mp, in.bd_mp, 2 * nbd);
  s.bd_nat = keep(c->bd_nat, in.bd_nat, nbd);
  s.bd_coef = keep(c->bd_coef, in.bd_coef, nbd);
  s.ba_mp = keep(c->ba_mp, in.ba_mp, 3 * nba);
  s.ba_nat = keep(c->ba_nat, in.ba_nat, nba);
  s.ba_coef = keep(c->ba_coef, in.ba_coef, nba);
  s.st_mp = keep(c->st_mp, in.st_mp, 7 * nst);
  s.st_nat = keep(c->st_nat, in.st_nat, 3 * nst);
  s.st_coef = keep(c->st_coef, in.st_coef, 4 * nst);
  s.tst_mp = keep(c->tst_mp, in.tst_mp, 6 * ntst);
  s.tst_nat = keep(c->tst_nat, in.tst_nat, 6 * ntst);
  s.tst_coef = keep(c->tst_coef, in.tst_coef, 7 * ntst);
  s.tst2st = keep(c->tst2st, in.tst2st, 2 * ntst);
  s.tst_excl = keep(c->tst_excl, in.tst_excl, 2 * ntst);
  s.hb_mp = keep(c->hb_mp, in.hb_mp, 6 * nhb);
  s.hb_nat = keep(c->hb_nat, in.hb_nat, 6 * nhb);
  s.hb_coef = keep(c->hb_coef, in.hb_coef, 7 * nhb);
  s.hb_site = keep(c->hb_site, in.hb_site, 6 * nhb);
  s.hb_tertiary = keep(c->hb_tert, in.hb_tertiary, nhb);
  s.hbsite_valence = keep(c->valence, in.hbsite_valence, in.nhbsite);
  s.mass = keep(c->mass, in.mass, n);
  s.fric = keep(c->fric, in.fric, n);
  s.radius = keep(c->radius, in.radius, n);
  return c;
}

Caps caps_for(const Stage &g, int nsys, bool periodic, double grow) {
  auto c = [grow](int64_t n) { return std::max<int64_t>(1, (int64_t)std::ceil((double)n * grow)); };
  Caps k;
  k.bead = c(g.nbead); k.q = c(g.nq); k.bd = c(g.nbd); k.ba = c(g.nba); k.st = c(g.nst); k.tst = c(g.ntst);
  k.hb = c(g.nhb); k.site = c(g.nsite); k.rl2 = c((int64_t)g.rl2.size()); k.inc = c((int64_t)g.inc.size());
  k.lc = c((int64_t)g.lcls.size()); k.spt = c((int64_t)g.sptab.size()); k.k = c((int64_t)g.ewc.size());
  k.m = c(g.moff); k.mi = c(g.mioff); k.pp = c(g.pp_off); k.lp = c(g.lp_off); k.psk = c(g.psk_off);
  k.pf = c(g.pf_off); k.ex = c((int64_t)g.st_ex.size()); k.nsys = c(nsys);
  k.periodic = periodic;
  return k;
}

bool caps_fit(const Caps &k, const Stage &g, int nsys, bool periodic) {
  return g.nbead <= k.bead && g.nq <= k.q && g.nbd <= k.bd && g.nba <= k.ba && g.nst <= k.st && g.ntst <= k.tst &&
         g.nhb <= k.hb && g.nsite <= k.site && (int64_t)g.rl2.size() <= k.rl2 && (int64_t)g.inc.size() <= k.inc &&
         (int64_t)g.lcls.size() <= k.lc && (int64_t)g.sptab.size() <= k.spt && (int64_t)g.ewc.size() <= k.k &&
         g.moff <= k.m && g.mioff <= k.mi && g.pp_off <= k.pp && g.lp_off <= k.lp && g.psk_off <= k.psk &&
         g.pf_off <= k.pf && (int64_t)g.st_ex.size() <= k.ex && nsys <= k.nsys && (!periodic || k.periodic);
}

// Stage system s (handle is) after the arena's last system: its DSys, its
// segments of every array, with global indices.
int stage_system(tis_ctx *c, Stage &g, HostSys &h, const tis_system &s, int is, uint32_t key) {
  if (s.nmp <= 0 || s.nunit <= 0) return fail(c, TIS_ERR_ARG, "system %d has no beads", is);
  if (s.p.ele_form == TIS_ELE_EWALD) {
    if (!s.p.periodic || s.p.box[0] != s.p.box[1] || s.p.box[0] != s.p.box[2])
      return fail(c, TIS_ERR_ARG, "system %d: Ewald needs a periodic cubic box", is);
    if (!(s.p.ewld_alpha > 0.0) || s.p.ewld_hmax <= 0)
      return fail(c, TIS_ERR_ARG, "system %d: Ewald needs ewld_alpha > 0 and ewld_hmax > 0", is);
  }
  if (s.p.ele_form != TIS_ELE_NONE && s.p.ele_form != TIS_ELE_DH && s.p.ele_form != TIS_ELE_COULOMB &&
      s.p.ele_form != TIS_ELE_EWALD)
    return fail(c, TIS_ERR_ARG, "system %d: unknown ele_form %d", is, s.p.ele_form);
  if ((int64_t)g.nbead + s.nmp >= (1 << 27)) return fail(c, TIS_ERR_ARG, "too many beads in the arena");
  DSys &d = h.d;
  std::memset(&d, 0, sizeof(d));
  fill_derived(d, s);
  h.mass.clear(); h.fric.clear(); h.lc_fric.clear(); h.lc_mass.clear();
  d.key = key;
  const int32_t off = g.nbead;
  d.off = off; d.nmp = s.nmp;
  d.uoff = (int32_t)g.rl2.size(); d.nunit = s.nunit;
  d.rl2max = 0.0;
  for (int k = 0; k < s.nunit * s.nunit; ++k) {
    g.rl2.push_back(s.rlist2_unit[k]);
    d.rl2max = std::max(d.rl2max, s.rlist2_unit[k]);
  }
  // excluded-volume species of this system: distinct (mptype, class,
  // radius, epsilon), numbered within the system
  struct Spec { int8_t t, c; double r, e; };
  std::vector<Spec> specs;
  auto spec_of = [&](int8_t t, int8_t cl, double r, double e) -> int {
    for (size_t k = 0; k < specs.size(); ++k)
      if (specs[k].t == t && specs[k].c == cl && specs[k].r == r && specs[k].e == e) return (int)k;
    specs.push_back(Spec{t, cl, r, e});
    return (int)specs.size() - 1;
  };
  // beads
  d.qoff = g.nq;
  int nq = 0;
  for (int i = 0; i < s.nmp; ++i) {
    g.mptype.push_back((int8_t)s.mptype[i]);
    g.iclass.push_back((int8_t)s.iclass[i]);
    g.unit.push_back(s.unit[i]);
    g.exv_rad.push_back(s.exv_rad[i]);
    g.exv_eps.push_back((float)s.exv_eps[i]);
    const int sp = spec_of((int8_t)s.mptype[i], (int8_t)s.iclass[i], s.exv_rad[i], s.exv_eps[i]);
    if (sp > 255) return fail(c, TIS_ERR_ARG, "system %d: too many bead species", is);
    g.species.push_back((uint8_t)sp);
    g.charge.push_back(s.charge[i]);
    g.mass.push_back(s.mass[i]);
    h.mass.push_back(s.mass[i]);
    h.fric.push_back(s.fric[i]);
    // Langevin class: the system's (friction, mass) pairs in order of
    // first appearance
    int cl = 0;
    while (cl < (int)h.lc_fric.size() && !(h.lc_fric[cl] == s.fric[i] && h.lc_mass[cl] == s.mass[i])) ++cl;
    if (cl == (int)h.lc_fric.size()) {
      if (cl == 256) return fail(c, TIS_ERR_ARG, "system %d: more than 256 (friction, mass) classes", is);
      h.lc_fric.push_back(s.fric[i]);
      h.lc_mass.push_back(s.mass[i]);
    }
    g.lcl.push_back((uint8_t)cl);
    // charges of the ele list: |q| >= ZERO_JUDGE (neighbor_list_ele.F90:134,147)
    if (s.p.ele_form != TIS_ELE_NONE && std::fabs(s.charge[i]) >= 1.0e-6) {
      g.qidx.push_back(g.nq + nq);
      g.qbead.push_back(off + i);
      ++nq;
    } else {
      g.qidx.push_back(-1);
    }
  }
  // species-pair table of the system (force_exv_dt15.F90:60-80): D_ij and
  // eps_i eps_j
  const int NS = (int)specs.size();
  d.sptab = (int32_t)g.sptab.size();
  d.nspecies = NS;
  for (int a = 0; a < NS; ++a)
    for (int b = 0; b < NS; ++b) {
      const Spec &A = specs[a], &B = specs[b];
      double dij;
      if ((A.t == TIS_MP_RNA_PHOS && B.t == TIS_MP_RNA_SUGAR) || (A.t == TIS_MP_RNA_SUGAR && B.t == TIS_MP_RNA_PHOS))
        dij = d.exv_dist_PS;
      else if (A.c == TIS_CLASS_RNA && B.c == TIS_CLASS_RNA)
        dij = d.exv_dist;
      else
        dij = A.r + B.r;
      g.sptab.push_back(SpPair{dij, dij * dij, A.e * B.e});
    }
  // its Langevin classes (constants r[] set by tis_langevin_init)
  d.lc0 = (int32_t)g.lcls.size();
  d.nlc = (int32_t)h.lc_fric.size();
  for (int k = 0; k < d.nlc; ++k)
    g.lcls.push_back(LangClass{{0.0, 0.0, 0.0, 0.0}, 1.0 / h.lc_mass[k], h.lc_fric[k], h.lc_mass[k], 0.0});
  d.nq = nq;
  d.mwords = (nq + 31) / 32;
  d.moff = g.moff;
  g.moff += (int64_t)nq * d.mwords;
  d.mioff = s.p.periodic ? g.mioff : -1;
  if (s.p.periodic) g.mioff += (int64_t)nq * nq;
  g.nq += nq;
  // terms with global bead indices
  const int32_t nbd = g.nbd, nba = g.nba, nst = g.nst, ntst = g.ntst, nhb = g.nhb, nsite = g.nsite;
  d.boff = nbd; d.nbd = s.nbd;
  for (int b = 0; b < s.nbd; ++b) {
    g.bd_mp.push_back(off + s.bd_mp[2 * b]); g.bd_mp.push_back(off + s.bd_mp[2 * b + 1]);
    g.bd_nat.push_back(s.bd_nat[b]); g.bd_coef.push_back(s.bd_coef[b]);
  }
  d.aoff = nba; d.nba = s.nba;
  for (int a = 0; a < s.nba; ++a) {
    for (int k = 0; k < 3; ++k) g.ba_mp.push_back(off + s.ba_mp[3 * a + k]);
    g.ba_nat.push_back(s.ba_nat[a]); g.ba_coef.push_back(s.ba_coef[a]);
  }
  d.soff = nst; d.nst = s.nst;
  for (int t = 0; t < s.nst; ++t) {
    for (int k = 0; k < 7; ++k) g.st_mp.push_back(off + s.st_mp[7 * t + k]);
    for (int k = 0; k < 3; ++k) g.st_nat.push_back(s.st_nat[3 * t + k]);
    for (int k = 0; k < 4; ++k) g.st_coef.push_back(s.st_coef[4 * t + k]);
  }
  d.toff = ntst; d.ntst = s.ntst;
  for (int t = 0; t < s.ntst; ++t) {
    for (int k = 0; k < 6; ++k) g.tst_mp.push_back(off + s.tst_mp[6 * t + k]);
    for (int k = 0; k < 6; ++k) g.tst_nat.push_back(s.tst_nat[6 * t + k]);
    for (int k = 0; k < 7; ++k) g.tst_coef.push_back(s.tst_coef[7 * t + k]);
    for (int k = 0; k < 2; ++k) {
      g.tst2st.push_back(s.tst2st[2 * t + k] >= 0 ? nst + s.tst2st[2 * t + k] : -1);
      g.tst_excl.push_back(s.tst_excl[2 * t + k]);
    }
    g.tst_sys.push_back(is);
  }
  // per consecutive stack, the exclusive tertiary stacks that name it
  // (CSR st_exoff / st_ex; st_exoff keeps its end sentinel)
  {
    std::vector<std::vector<int32_t>> ex(s.nst);
    for (int t = 0; t < s.ntst; ++t)
      for (int k = 0; k < 2; ++k)
        if (s.tst_excl[2 * t + k] && s.tst2st[2 * t + k] >= 0) ex[s.tst2st[2 * t + k]].push_back(ntst + t);
    g.st_exoff.pop_back();
    for (int t = 0; t < s.nst; ++t) {
      g.st_exoff.push_back((int32_t)g.st_ex.size());
      g.st_ex.insert(g.st_ex.end(), ex[t].begin(), ex[t].end());
    }
    g.st_exoff.push_back((int32_t)g.st_ex.size());
  }
  d.sitoff = nsite; d.nsite = s.nhbsite;
  for (int k = 0; k < s.nhbsite; ++k) g.site_val.push_back(s.hbsite_valence[k]);
  d.hoff = nhb; d.nhb = s.nhb;
  std::vector<int> site_use(std::max(s.nhbsite, 1), 0);
  for (int t = 0; t < s.nhb; ++t) {
    for (int k = 0; k < 6; ++k) g.hb_mp.push_back(off + s.hb_mp[6 * t + k]);
    for (int k = 0; k < 6; ++k) g.hb_nat.push_back(s.hb_nat[6 * t + k]);
    for (int k = 0; k < 7; ++k) g.hb_coef.push_back(s.hb_coef[7 * t + k]);
    for (int k = 0; k < 6; ++k) {
      int st = s.hb_site[6 * t + k];
      g.hb_site.push_back(st >= 0 ? nsite + st : -1);
      if (st >= 0) site_use[st] += 1;
    }
    g.hb_tert.push_back(s.hb_tertiary ? s.hb_tertiary[t] : 0);
    g.hb_sys.push_back(is);
  }
  d.hb_excess_possible = 0;
  for (int k = 0; k < s.nhbsite; ++k)
    if (site_use[k] > s.hbsite_valence[k]) d.hb_excess_possible = 1;
  // incidence lists (bonds, angles, stacks, tertiary stacks, hbonds) and
  // the incidence position of every (term, slot), kept term-major here
  std::vector<std::vector<uint32_t>> incb(s.nmp);
  for (int b = 0; b < s.nbd; ++b)
    for (int k = 0; k < 2; ++k) incb[s.bd_mp[2 * b + k]].push_back(inc_pack(INC_BOND, k, nbd + b));
  for (int a = 0; a < s.nba; ++a)
    for (int k = 0; k < 3; ++k) incb[s.ba_mp[3 * a + k]].push_back(inc_pack(INC_ANGLE, k, nba + a));
  for (int t = 0; t < s.nst; ++t)
    for (int k = 0; k < 7; ++k) incb[s.st_mp[7 * t + k]].push_back(inc_pack(INC_STACK, k, nst + t));
  for (int t = 0; t < s.ntst; ++t)
    for (int k = 0; k < 6; ++k) incb[s.tst_mp[6 * t + k]].push_back(inc_pack(INC_TSTACK, k, ntst + t));
  for (int t = 0; t < s.nhb; ++t)
    for (int k = 0; k < 6; ++k) incb[s.hb_mp[6 * t + k]].push_back(inc_pack(INC_HBOND, k, nhb + t));
  g.inc_off.pop_back();
  for (int i = 0; i < s.nmp; ++i) {
    g.inc_off.push_back((int32_t)g.inc.size());
    g.inc.insert(g.inc.end(), incb[i].begin(), incb[i].end());
  }
  g.inc_off.push_back((int32_t)g.inc.size());
  g.nbd += s.nbd; g.nba += s.nba; g.nst += s.nst; g.ntst += s.ntst; g.nhb += s.nhb; g.nsite += s.nhbsite;
  // tile counts (the tiles themselves are in the launch tables)
  d.nbt = (s.nmp + kTile - 1) / kTile;
  d.nqt = (nq + kTile - 1) / kTile;
  // electrostatic chunk-pair partial forces
  d.pp_off = g.pp_off;
  if (s.p.ele_form != TIS_ELE_NONE && nq > 0) g.pp_off += (int64_t)d.nqt * (d.nqt + 1) / 2 * 6 * kTile;
  // counts of the periodic tile-pair list build
  d.lp_off = g.lp_off;
  if (s.p.periodic) g.lp_off += (int64_t)d.nbt * d.nbt;
  h.has_task = (s.ntst > 0 || (s.nhb > 0 && d.hb_excess_possible));
  // Ewald reciprocal vectors
  if (s.p.ele_form == TIS_ELE_EWALD && nq > 0) {
    d.ewk0 = (int32_t)g.ewc.size();
    ewald_kvectors(s.p, g.ewhx, g.ewhy, g.ewhz, g.ewc);
    d.nk = (int32_t)g.ewc.size() - d.ewk0;
    d.nqc = (nq + kTile - 1) / kTile;
    d.nkc = (d.nk + kEwKChunk - 1) / kEwKChunk;
    d.psk_off = g.psk_off; g.psk_off += 2 * (int64_t)d.nqc * d.nk;
    d.pf_off = g.pf_off; g.pf_off += 3 * (int64_t)d.nkc * nq;
  }
  g.nbead += s.nmp;
  h.stored = true;
  return 0;
}

// Device arrays of an arena at capacities k, in `list` (everything that
// tis_system_append may later fill in place).  The DSys array, counters
// and step base are allocated here too.
int alloc_arena(tis_ctx *c, const Caps &k, std::vector<void *> &list, DevPtrs &P, ArenaBufs &A, int nb_cap) {
  P = DevPtrs{};
  A = ArenaBufs{};
  const size_t B = (size_t)k.bead;
#define AL(ptr, n) do { int rc_ = dalloc_in(c, list, &(ptr), (size_t)(n)); if (rc_) return rc_; } while (0)
  AL(A.sys, k.nsys);
  P.sys = A.sys;
  for (int q = 0; q < 2; ++q) {
    AL(P.x[q], 3 * B);
    AL(P.px[q], k.periodic ? 3 * B : 1);
  }
  AL(P.v, 3 * B); AL(P.a, 3 * B); AL(P.f, 3 * B);
  AL(A.lcl, B); AL(A.lcls, k.lc); AL(A.mass, B); AL(A.exv_rad, B); AL(A.charge, B); AL(A.rl2, k.rl2);
  AL(A.mptype, B); AL(A.iclass, B); AL(A.unit, B); AL(A.qidx, B); AL(A.qbead, k.q); AL(A.inc_off, B + 1);
  AL(A.inc_slot, k.inc); AL(A.exv_eps, B); AL(A.species, B); AL(A.sptab, k.spt);
  // slot regions, term-major per type
  P.sb_bd = 0;
  P.sb_ba = P.sb_bd + 2 * k.bd;
  P.sb_st = P.sb_ba + 3 * k.ba;
  P.sb_tst = P.sb_st + 7 * k.st;
  P.sb_hb = P.sb_tst + 6 * k.tst;
  AL(P.fslot, 3 * (P.sb_hb + 6 * k.hb));
  AL(A.bd_mp, 2 * k.bd); AL(A.bd_nat, k.bd); AL(A.bd_coef, k.bd);
  AL(A.ba_mp, 3 * k.ba); AL(A.ba_nat, k.ba); AL(A.ba_coef, k.ba);
  AL(A.st_mp, 7 * k.st); AL(A.st_nat, 3 * k.st); AL(A.st_coef, 4 * k.st);
  AL(A.tst_mp, 6 * k.tst); AL(A.tst_nat, 6 * k.tst); AL(A.tst_coef, 7 * k.tst); AL(A.tst2st, 2 * k.tst);
  AL(A.tst_excl, 2 * k.tst); AL(A.tst_sys, k.tst);
  AL(A.hb_mp, 6 * k.hb); AL(A.hb_nat, 6 * k.hb); AL(A.hb_coef, 7 * k.hb); AL(A.hb_site, 6 * k.hb);
  AL(A.hb_tert, k.hb); AL(A.hb_sys, k.hb); AL(A.site_val, k.site);
  AL(A.st_exoff, k.st + 1); AL(A.st_ex, k.ex);
  AL(P.hb_inlist, k.hb); AL(P.hb_status, k.hb); AL(P.hb_ene, k.hb); AL(P.for_hb, 18 * k.hb);
  AL(P.site_cnt, k.site); AL(P.hb_seq, k.hb); AL(A.ststat, k.st);
  AL(P.nb_cnt, B);
  P.nb_cap = std::max(nb_cap, 64);
  AL(P.nb, (size_t)P.nb_cap * B);
  AL(P.dxyz, 3 * B); AL(P.dlast, 3 * B);
  AL(P.sys_flag, k.nsys);
  AL(P.pp, k.pp); AL(P.qmask, k.m); AL(P.qmirror, k.mi); AL(P.lcnt, k.lp * kTile);
  AL(A.ewh, 3 * k.k); AL(A.ewc, k.k); AL(P.psk, k.psk); AL(P.pf, k.pf);
  AL(A.counters, TIS_CNT_MAX); AL(A.stepbase, 1); AL(A.energy, 25 * k.nsys);
#undef AL
  P.lcl = A.lcl; P.lcls = A.lcls; P.mass = A.mass; P.exv_rad = A.exv_rad; P.charge = A.charge; P.rl2 = A.rl2;
  P.mptype = A.mptype; P.iclass = A.iclass; P.unit = A.unit; P.qidx = A.qidx; P.qbead = A.qbead;
  P.inc_off = A.inc_off; P.inc_slot = A.inc_slot; P.exv_eps = A.exv_eps; P.species = A.species; P.sptab = A.sptab;
  P.bd_mp = A.bd_mp; P.bd_nat = A.bd_nat; P.bd_coef = A.bd_coef;
  P.ba_mp = A.ba_mp; P.ba_nat = A.ba_nat; P.ba_coef = A.ba_coef;
  P.st_mp = A.st_mp; P.st_nat = A.st_nat; P.st_coef = A.st_coef;
  P.tst_mp = A.tst_mp; P.tst_nat = A.tst_nat; P.tst_coef = A.tst_coef; P.tst2st = A.tst2st;
  P.tst_excl = A.tst_excl; P.tst_sys = A.tst_sys;
  P.hb_mp = A.hb_mp; P.hb_nat = A.hb_nat; P.hb_coef = A.hb_coef; P.hb_site = A.hb_site; P.hb_tert = A.hb_tert;
  P.hb_sys = A.hb_sys; P.site_val = A.site_val; P.st_exoff = A.st_exoff; P.st_ex = A.st_ex;
  P.ewh = A.ewh; P.ewc = A.ewc;
  P.counters = A.counters; P.stepbase = A.stepbase;
  // strides
  P.ntot = (int32_t)k.bead; P.nqtot = (int32_t)k.q;
  P.nbd = (int32_t)k.bd; P.nba = (int32_t)k.ba; P.nst = (int32_t)k.st; P.ntst = (int32_t)k.tst;
  P.nhb = (int32_t)k.hb; P.nktot = (int32_t)k.k;
  return 0;
}

template <typename T>
int put_tail(tis_ctx *c, T *dev, const std::vector<T> &h, size_t from) {
  if (h.size() <= from) return 0;
  HIPCHK(c, hipMemcpyAsync(dev + from, h.data() + from, (h.size() - from) * sizeof(T), hipMemcpyHostToDevice,
                           c->stream));
  c->h2d_bytes += (int64_t)((h.size() - from) * sizeof(T));
  return 0;
}

// Upload the staged arrays from the marks m on (m all zero: everything).
int upload_stage(tis_ctx *c, const Stage &g, const StageMarks &m) {
  const ArenaBufs &A = c->ab;
  int rc = 0;
#define PT(dev, vec, from) do { rc = put_tail(c, dev, g.vec, from); if (rc) return rc; } while (0)
  PT(A.rl2, rl2, m.rl2);
  PT(A.mptype, mptype, m.bead); PT(A.iclass, iclass, m.bead); PT(A.unit, unit, m.bead); PT(A.qidx, qidx, m.bead);
  PT(A.exv_rad, exv_rad, m.bead); PT(A.exv_eps, exv_eps, m.bead); PT(A.species, species, m.bead);
  PT(A.charge, charge, m.bead); PT(A.mass, mass, m.bead); PT(A.lcl, lcl, m.bead);
  PT(A.inc_off, inc_off, m.bead);   // from the old end sentinel on
  PT(A.qbead, qbead, m.q); PT(A.lcls, lcls, m.lc); PT(A.sptab, sptab, m.spt);
  PT(A.bd_mp, bd_mp, 2 * m.bd); PT(A.bd_nat, bd_nat, m.bd); PT(A.bd_coef, bd_coef, m.bd);
  PT(A.ba_mp, ba_mp, 3 * m.ba); PT(A.ba_nat, ba_nat, m.ba); PT(A.ba_coef, ba_coef, m.ba);
  PT(A.st_mp, st_mp, 7 * m.st); PT(A.st_nat, st_nat, 3 * m.st); PT(A.st_coef, st_coef, 4 * m.st);
  PT(A.tst_mp, tst_mp, 6 * m.tst); PT(A.tst_nat, tst_nat, 6 * m.tst); PT(A.tst_coef, tst_coef, 7 * m.tst);
  PT(A.tst2st, tst2st, 2 * m.tst); PT(A.tst_excl, tst_excl, 2 * m.tst); PT(A.tst_sys, tst_sys, m.tst);
  PT(A.hb_mp, hb_mp, 6 * m.hb); PT(A.hb_nat, hb_nat, 6 * m.hb); PT(A.hb_coef, hb_coef, 7 * m.hb);
  PT(A.hb_site, hb_site, 6 * m.hb); PT(A.hb_tert, hb_tert, m.hb); PT(A.hb_sys, hb_sys, m.hb);
  PT(A.site_val, site_val, m.site);
  PT(A.st_exoff, st_exoff, m.st); PT(A.st_ex, st_ex, m.ex);
  PT(A.ewc, ewc, m.k);
  PT(A.ewh, ewhx, m.k); PT(A.ewh + c->P.nktot, ewhy, m.k); PT(A.ewh + 2 * (size_t)c->P.nktot, ewhz, m.k);
#undef PT
  // incidence entries (type | slot | term) -> slot positions of the layout
  const DevPtrs &P = c->P;
  if (g.inc.size() > m.inc) {
    std::vector<int32_t> pos(g.inc.size() - m.inc);
    const int64_t base[5] = {P.sb_bd, P.sb_ba, P.sb_st, P.sb_tst, P.sb_hb};
    const int K[5] = {2, 3, 7, 6, 6};
    for (size_t k = m.inc; k < g.inc.size(); ++k) {
      const uint32_t e = g.inc[k], type = e >> 29, slot = (e >> 26) & 7u, term = e & ((1u << 26) - 1u);
      pos[k - m.inc] = (int32_t)(base[type] + (int64_t)K[type] * term + slot);
    }
    HIPCHK(c, hipMemcpyAsync(A.inc_slot + m.inc, pos.data(), pos.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));   // pos is local
    c->h2d_bytes += (int64_t)(pos.size() * sizeof(int32_t));
  }
  return 0;
}

StageMarks marks_of(const Stage &g) {
  StageMarks m;
  m.bead = g.nbead; m.q = g.qbead.size(); m.bd = g.nbd; m.ba = g.nba; m.st = g.nst; m.tst = g.ntst; m.hb = g.nhb;
  m.site = g.nsite; m.rl2 = g.rl2.size(); m.inc = g.inc.size(); m.lc = g.lcls.size(); m.spt = g.sptab.size();
  m.k = g.ewc.size(); m.ex = g.st_ex.size();
  return m;
}

// The launch tables of the live systems (tiles, chunk pairs, tile pairs,
// Ewald tasks, hbond / tertiary-stack systems, the live list) and every
// system's DSys; tables are rebuilt whole on each change of the live set.
int build_tables(tis_ctx *c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (void *p : c->tab_allocs) { (void)hipFree(p); }
  c->dev_bytes -= c->tab_bytes;
  c->tab_allocs.clear();
  c->tab_bytes = 0;
  std::vector<Tile> btiles, qtiles;
  std::vector<EwTask> ptask, bpair, qpair, ewsk, ewf;
  std::vector<int32_t> systask, sysidx, hbdense;
  c->systask_of.assign(c->nsys, -1);
  c->any_ele = false;
  c->any_periodic = false;
  std::vector<DSys> dsys(c->nsys);
  for (int is = 0; is < c->nsys; ++is) {
    HostSys &h = c->hs[is];
    DSys &d = h.d;
    h.bt_begin = h.bt_count = h.qt_begin = h.qt_count = h.esk_begin = h.esk_count = 0;
    h.ef_begin = h.ef_count = h.pt_begin = h.pt_count = h.bp_begin = h.bp_count = h.qp_begin = h.qp_count = 0;
    h.hd_begin = h.hd_count = 0;
    if (!h.live) { std::memset(&dsys[is], 0, sizeof(DSys)); dsys[is].nmp = 0; continue; }
    if (d.periodic) c->any_periodic = true;
    if (d.ele_form >= 0) c->any_ele = true;
    h.bt_begin = (int32_t)btiles.size();
    for (int st0 = 0; st0 < d.nmp; st0 += kTile) btiles.push_back(Tile{is, st0});
    h.bt_count = (int32_t)btiles.size() - h.bt_begin;
    h.qt_begin = (int32_t)qtiles.size();
    for (int st0 = 0; st0 < d.nq; st0 += kTile) qtiles.push_back(Tile{is, st0});
    h.qt_count = (int32_t)qtiles.size() - h.qt_begin;
    d.bt0 = h.bt_begin; d.nbt = h.bt_count; d.qt0 = h.qt_begin; d.nqt = h.qt_count;
    // electrostatic pairs: one wave per chunk pair (I, J), I <= J
    h.pt_begin = (int32_t)ptask.size();
    if (d.ele_form != TIS_ELE_NONE && d.nq > 0)
      for (int I = 0; I < d.nqt; ++I)
        for (int J = I; J < d.nqt; ++J) ptask.push_back(EwTask{is, I, J});
    h.pt_count = (int32_t)ptask.size() - h.pt_begin;
    // tile pairs of the periodic list builds (k_lpair; periodic systems
    // only: their number and the count buffer grow as nbt^2)
    h.bp_begin = (int32_t)bpair.size();
    h.qp_begin = (int32_t)qpair.size();
    if (d.periodic) {
      for (int a = 0; a < d.nbt; ++a)
        for (int b = 0; b < d.nbt; ++b) bpair.push_back(EwTask{is, a, b});
      for (int a = 0; a < d.nqt; ++a)
        for (int b = 0; b < d.nqt; ++b) qpair.push_back(EwTask{is, a, b});
    }
    h.bp_count = (int32_t)bpair.size() - h.bp_begin;
    h.qp_count = (int32_t)qpair.size() - h.qp_begin;
    if (h.has_task) { c->systask_of[is] = (int32_t)systask.size(); systask.push_back(is); }
    h.hd_begin = (int32_t)hbdense.size();
    if (!d.hb_excess_possible)
      for (int k = 0; k < d.nhb; ++k) hbdense.push_back(d.hoff + k);
    h.hd_count = (int32_t)hbdense.size() - h.hd_begin;
    if (d.ele_form == TIS_ELE_EWALD && d.nq > 0) {
      h.esk_begin = (int32_t)ewsk.size();
      for (int kt = 0; kt < (d.nk + kTile - 1) / kTile; ++kt)
        for (int qc = 0; qc < d.nqc; ++qc) ewsk.push_back(EwTask{is, kt, qc});
      h.esk_count = (int32_t)ewsk.size() - h.esk_begin;
      h.ef_begin = (int32_t)ewf.size();
      for (int qt = 0; qt < d.nqc; ++qt)
        for (int kc = 0; kc < d.nkc; ++kc) ewf.push_back(EwTask{is, qt, kc});
      h.ef_count = (int32_t)ewf.size() - h.ef_begin;
    }
    h.live_pos = (int32_t)sysidx.size();
    sysidx.push_back(is);
    dsys[is] = d;
  }
  c->nlive = (int32_t)sysidx.size();
  c->P.nsys = c->nsys;
  c->nbtile = (int32_t)btiles.size(); c->nqtile = (int32_t)qtiles.size(); c->nsystask = (int32_t)systask.size();
  c->nhbdense = (int32_t)hbdense.size();
  c->n_ptask = (int32_t)ptask.size(); c->n_bpair = (int32_t)bpair.size(); c->n_qpair = (int32_t)qpair.size();
  c->n_ewsk = (int32_t)ewsk.size(); c->n_ewf = (int32_t)ewf.size();
  if (!dsys.empty()) {
    HIPCHK(c, hipMemcpyAsync(c->ab.sys, dsys.data(), dsys.size() * sizeof(DSys), hipMemcpyHostToDevice, c->stream));
    c->h2d_bytes += (int64_t)(dsys.size() * sizeof(DSys));
  }
  // whole-batch k_force launches take the off-diagonal chunk pairs (64
  // steps) before the diagonal ones (32): the order changes no result
  std::vector<EwTask> lpt(ptask);
  std::stable_partition(lpt.begin(), lpt.end(), [](const EwTask &t) { return t.a != t.b; });
  int rc = 0;
#define UPT(ptr, vec) do { rc = dupload_in(c, c->tab_allocs, &(ptr), vec, &c->tab_bytes); if (rc) return rc; } while (0)
  UPT(c->d_btiles, btiles); UPT(c->d_qtiles, qtiles);
  {
    const std::vector<Tile> xt = xcd_order(btiles, c->nsys);
    UPT(c->d_btiles_xcd, xt);
  }
  UPT(c->d_ptask, ptask); UPT(c->d_ptask_lpt, lpt); UPT(c->d_bpair, bpair); UPT(c->d_qpair, qpair);
  UPT(c->d_ewsk, ewsk); UPT(c->d_ewf, ewf); UPT(c->d_systask, systask); UPT(c->d_sysidx, sysidx);
  UPT(c->d_hbdense, hbdense);
#undef UPT
  c->P.hb_dense = c->d_hbdense;
  const size_t nbt = std::max<size_t>(btiles.size(), 1), nqt = std::max<size_t>(qtiles.size(), 1);
  if ((rc = dalloc_in(c, c->tab_allocs, &c->P.bbox, nbt, &c->tab_bytes))) return rc;
  if ((rc = dalloc_in(c, c->tab_allocs, &c->P.qbbox, nqt, &c->tab_bytes))) return rc;
  if ((rc = dalloc_in(c, c->tab_allocs, &c->P.sbox, 4 * nbt, &c->tab_bytes))) return rc;
  if ((rc = dalloc_in(c, c->tab_allocs, &c->P.ttop, 2 * nbt, &c->tab_bytes))) return rc;
  c->dev_bytes += c->tab_bytes;
  for (int k = 0; k < 2; ++k)
    if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  c->lists_valid = false;
  c->lv_init = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Lay the arena out again at capacities for the live systems (and `extra`
// systems about to be staged) times `grow`: the live systems are staged
// from their host copies in handle order and their positions (current
// slot), velocities and accelerations are copied over on the device; the
// counters and the step base carry over.  Retired handles keep their index
// with no storage.
int relayout(tis_ctx *c, const std::vector<const tis_system *> &extra, double grow) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  Stage g;
  std::vector<DSys> old_d(c->nsys);
  std::vector<HostSys> hs2(c->hs);   // committed once the new arena exists
  bool periodic = false;
  for (int is = 0; is < c->nsys; ++is) {
    HostSys &h = hs2[is];
    old_d[is] = h.d;
    if (!h.live) { h.stored = false; continue; }
    const uint64_t step_origin = h.d.step_origin;
    const int32_t init_accel = h.d.init_accel;
    int rc = stage_system(c, g, h, h.src->s, is, h.d.key);
    if (rc) return rc;
    h.d.step_origin = step_origin;
    h.d.init_accel = init_accel;
    periodic = periodic || h.d.periodic;
  }
  // capacity for what is about to be appended
  Stage probe = g;
  std::vector<HostSys> tmp(extra.size());
  for (size_t k = 0; k < extra.size(); ++k) {
    int rc = stage_system(c, probe, tmp[k], *extra[k], c->nsys + (int)k, 0);
    if (rc) return rc;
    periodic = periodic || extra[k]->p.periodic;
  }
  const Caps caps = caps_for(probe, c->nsys + (int)extra.size(), periodic, grow);
  std::vector<void *> list;
  DevPtrs P;
  ArenaBufs A;
  const size_t bytes0 = c->dev_bytes;
  c->dev_bytes = c->tab_bytes;
  int rc = alloc_arena(c, caps, list, P, A, c->P.nb_cap);
  if (rc) {
    for (void *p : list) (void)hipFree(p);
    c->dev_bytes = bytes0;
    return rc;
  }
  c->hs = std::move(hs2);
  // the new layout's topology from the host copies
  const DevPtrs O = c->P;
  const ArenaBufs OA = c->ab;
  const uint64_t seed = c->P.seed;
  std::vector<void *> old = std::move(c->allocs);
  c->allocs = list;
  c->P = P;
  c->P.seed = seed;
  c->ab = A;
  c->caps = caps;
  c->stage = std::move(g);
  c->d_sys = A.sys; c->d_ststat = A.ststat; c->d_energy = A.energy;
  c->d_counters = A.counters; c->d_stepbase = A.stepbase;
  c->d_eunit = nullptr;
  c->eunit_cap = 0;
  c->relayouts += 1;
  rc = upload_stage(c, c->stage, StageMarks{});
  // then, on the device: the dynamic state, and what temperature changes
  // (tis_set_system_T, replica labels) have made of charges, stack U0 and
  // the temperature fields; the counters and the step base
  const int cu = c->cur;
  for (int is = 0; rc == 0 && is < c->nsys; ++is) {
    HostSys &h = c->hs[is];
    if (!h.live) continue;
    const size_t n = (size_t)h.d.nmp, so = (size_t)old_d[is].off, sn = (size_t)h.d.off;
    for (int d = 0; d < 3 && rc == 0; ++d) {
      const size_t oo = d * (size_t)O.ntot + so, no = d * (size_t)P.ntot + sn;
      if (hipMemcpyAsync(P.x[cu] + no, O.x[cu] + oo, n * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
          hipMemcpyAsync(P.v + no, O.v + oo, n * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
          hipMemcpyAsync(P.a + no, O.a + oo, n * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        rc = fail(c, TIS_ERR_HIP, "state copy of a new layout failed");
    }
    c->d2d_state += (int64_t)(9 * n * 8);
    if (rc == 0 && (hipMemcpyAsync(A.charge + sn, OA.charge + so, n * 8, hipMemcpyDeviceToDevice, c->stream) !=
                        hipSuccess ||
                    (h.d.nst > 0 && hipMemcpyAsync(A.st_coef + 4 * (size_t)h.d.soff,
                                                   OA.st_coef + 4 * (size_t)old_d[is].soff,
                                                   32 * (size_t)h.d.nst, hipMemcpyDeviceToDevice,
                                                   c->stream) != hipSuccess)))
      rc = fail(c, TIS_ERR_HIP, "parameter copy of a new layout failed");
    copy_t_fields(h.d, old_d[is]);
  }
  if (rc == 0 && OA.counters &&
      (hipMemcpyAsync(A.counters, OA.counters, TIS_CNT_MAX * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                      c->stream) != hipSuccess ||
       hipMemcpyAsync(A.stepbase, OA.stepbase, sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream) != hipSuccess))
    rc = fail(c, TIS_ERR_HIP, "counter copy of a new layout failed");
  (void)hipStreamSynchronize(c->stream);
  for (void *p : old) (void)hipFree(p);
  return rc;
}

// Stage systems at the arena's end and upload their segments, laying the
// arena out again (with room to grow by `grow`) when they do not fit.
int append_systems(tis_ctx *c, int32_t nadd, const tis_system *sys, const uint32_t *keys, double grow) {
  for (int k = 0; k < nadd; ++k)
    if (int rc = validate_system(&c->err, c->nsys + k, sys[k])) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  // fits as it is?
  bool fits = false;
  {
    Stage t = c->stage;
    std::vector<HostSys> tmp(nadd);
    bool periodic = c->caps.periodic;
    bool ok = true;
    for (int k = 0; k < nadd && ok; ++k) {
      if (int rc = stage_system(c, t, tmp[k], sys[k], c->nsys + k, 0)) return rc;
      if (sys[k].p.periodic && !c->caps.periodic) ok = false;
      periodic = periodic || sys[k].p.periodic;
    }
    fits = ok && c->have_arena && caps_fit(c->caps, t, c->nsys + nadd, periodic);
  }
  if (!fits) {
    std::vector<const tis_system *> ex(nadd);
    for (int k = 0; k < nadd; ++k) ex[k] = &sys[k];
    int rc = relayout(c, ex, grow);
    if (rc) return rc;
    c->have_arena = true;
  }
  const StageMarks m = marks_of(c->stage);
  for (int k = 0; k < nadd; ++k) {
    const int is = c->nsys;
    c->hs.emplace_back();
    HostSys &h = c->hs.back();
    h.src = copy_system(sys[k]);
    h.live = true;
    c->nsys += 1;
    int rc = stage_system(c, c->stage, h, h.src->s, is, keys ? keys[k] : (uint32_t)is);
    if (rc) { c->hs.pop_back(); c->nsys -= 1; return rc; }
  }
  int rc = upload_stage(c, c->stage, m);
  if (rc) return rc;
  c->nbeads = c->stage.nbead; c->nqtot = c->stage.nq; c->nhbtot = c->stage.nhb; c->ntsttot = c->stage.ntst;
  c->nsttot = c->stage.nst;
  return build_tables(c);
}

}  // namespace


extern "C" {

int32_t tis_ctx_create(int32_t device, tis_ctx **out) {
  if (!out) return TIS_ERR_ARG;
  *out = nullptr;
  tis_ctx *c = new tis_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return TIS_ERR_HIP; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return TIS_ERR_HIP; }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) { delete c; return TIS_ERR_HIP; }
  (void)hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming);
  (void)hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming);
  (void)hipEventCreate(&c->ev0);
  (void)hipEventCreate(&c->ev1);
  *out = c;
  return TIS_OK;
}

int32_t tis_system_validate(int32_t nsys, const tis_system *systems, char *msg, int64_t msg_cap) {
  std::string err;
  int rc = TIS_OK;
  if (nsys <= 0 || !systems) {
    err = "bad arguments to tis_system_validate";
    rc = TIS_ERR_ARG;
  }
  for (int is = 0; rc == TIS_OK && is < nsys; ++is) rc = validate_system(&err, is, systems[is]);
  if (msg && msg_cap > 0) {
    std::strncpy(msg, err.c_str(), (size_t)msg_cap - 1);
    msg[msg_cap - 1] = '\0';
  }
  return rc;
}

int32_t tis_ctx_destroy(tis_ctx *c) {
  if (!c) return TIS_ERR_ARG;
  (void)hipSetDevice(c->device);
  free_all(c);
  bd_free(c);
  for (void *p : c->label_bufs) (void)hipFree(p);
  if (c->d_lab_args) (void)hipFree(c->d_lab_args);
  if (c->d_lab_scale) (void)hipFree(c->d_lab_scale);
  if (c->d_eset) (void)hipFree(c->d_eset);
  if (c->d_eset_out) (void)hipFree(c->d_eset_out);
  for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TIS_OK;
}

const char *tis_last_error(const tis_ctx *c) { return c ? c->err.c_str() : "null context"; }

int32_t tis_system_upload(tis_ctx *c, int32_t nsys, const tis_system *sys, const uint32_t *keys,
                          uint64_t seed) {
  if (!c || nsys <= 0 || !sys) return fail(c, TIS_ERR_ARG, "bad arguments to tis_system_upload");
  // every index a kernel or the incidence build dereferences, checked before
  // anything is freed or allocated (a failed upload keeps the previous batch)
  for (int is = 0; is < nsys; ++is)
    if (int rc = validate_system(&c->err, is, sys[is])) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  free_all(c);
  // an arena that fits the batch exactly
  int rc = append_systems(c, nsys, sys, keys, 1.0);
  if (rc) return rc;
  c->P.seed = seed;
  c->cur = 0;
  return TIS_OK;
}

int32_t tis_system_append(tis_ctx *c, int32_t nadd, const tis_system *sys, const uint32_t *keys,
                          int32_t *isys_out) {
  if (!c || nadd <= 0 || !sys) return fail(c, TIS_ERR_ARG, "bad arguments to tis_system_append");
  if (!c->have_arena) return fail(c, TIS_ERR_STATE, "no batch: tis_system_upload first");
  const int first = c->nsys;
  // room to grow: a new layout reserves twice what it holds
  int rc = append_systems(c, nadd, sys, keys, 2.0);
  if (rc) return rc;
  if (isys_out)
    for (int k = 0; k < nadd; ++k) isys_out[k] = first + k;
  return TIS_OK;
}

int32_t tis_system_retire(tis_ctx *c, int32_t n, const int32_t *isys) {
  if (!c || n < 0 || (n > 0 && !isys)) return TIS_ERR_ARG;
  for (int k = 0; k < n; ++k) {
    if (isys[k] < 0 || isys[k] >= c->nsys || !c->hs[isys[k]].live)
      return fail(c, TIS_ERR_ARG, "system %d is not live", isys[k]);
    for (int j = 0; j < k; ++j)
      if (isys[j] == isys[k]) return fail(c, TIS_ERR_ARG, "system %d listed twice", isys[k]);
  }
  if (n == 0) return TIS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  for (int k = 0; k < n; ++k) {
    HostSys &h = c->hs[isys[k]];
    h.live = false;
    h.src.reset();
  }
  return build_tables(c);
}

int32_t tis_arena_stats(tis_ctx *c, int64_t *stats) {
  if (!c || !stats) return TIS_ERR_ARG;
  int64_t live = 0;
  for (const HostSys &h : c->hs)
    if (h.live) live += h.d.nmp;
  stats[0] = c->nlive;
  stats[1] = live;
  stats[2] = c->stage.nbead;
  stats[3] = c->have_arena ? c->caps.bead : 0;
  stats[4] = c->relayouts;
  stats[5] = c->nsys;
  return TIS_OK;
}

int32_t tis_set_system_T(tis_ctx *c, int32_t isys, const tis_system *s) {
  if (!c || !s) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_set_system_T needs one system");
  HostSys &h = c->hs[isys];
  if (s->nmp != h.d.nmp || s->nst != h.d.nst) return fail(c, TIS_ERR_ARG, "topology mismatch");
  HIPCHK(c, hipSetDevice(c->device));
  fill_derived(h.d, *s);
  HIPCHK(c, hipMemcpyAsync(c->d_sys + isys, &h.d, sizeof(DSys), hipMemcpyHostToDevice, c->stream));
  std::vector<double> q(s->charge, s->charge + s->nmp);
  HIPCHK(c, hipMemcpyAsync((void *)(c->P.charge + h.d.off), q.data(), q.size() * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  std::vector<double> stc(s->st_coef, s->st_coef + 4 * s->nst);
  if (!stc.empty())
    HIPCHK(c, hipMemcpyAsync((void *)(c->P.st_coef + 4 * (size_t)h.d.soff), stc.data(), stc.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  if (c->lv_init) {
    std::vector<LangClass> lc(h.d.nlc);
    for (int k = 0; k < h.d.nlc; ++k) {
      lc[k] = LangClass{{0.0, 0.0, 0.0, 0.0}, 1.0 / h.lc_mass[k], h.lc_fric[k], h.lc_mass[k], 0.0};
      langevin_consts(c->tstep, s->tempk, h.lc_fric[k], h.lc_mass[k], lc[k].r);
    }
    if (!lc.empty())
      HIPCHK(c, hipMemcpyAsync((void *)(c->P.lcls + h.d.lc0), lc.data(), lc.size() * sizeof(LangClass),
                               hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TIS_OK;
}

int32_t tis_state_set(tis_ctx *c, int32_t isys, int32_t which, const double *buf) {
  if (!c || !buf) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_state_set needs one system");
  HIPCHK(c, hipSetDevice(c->device));
  const HostSys &h = c->hs[isys];
  double *base = which == 0 ? Xc(c) : (which == 1 ? c->P.v : c->P.a);
  if (which < 0 || which > 2) return fail(c, TIS_ERR_ARG, "bad state index %d", which);
  std::vector<double> comp(h.d.nmp);
  for (int d = 0; d < 3; ++d) {
    for (int i = 0; i < h.d.nmp; ++i) comp[i] = buf[3 * i + d];
    HIPCHK(c, hipMemcpyAsync(base + (size_t)d * c->P.ntot + h.d.off, comp.data(), comp.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    c->h2d_state += (int64_t)(comp.size() * sizeof(double));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  if (which == 0) { c->lists_valid = false; c->lv_init = false; }
  return TIS_OK;
}

int32_t tis_state_get(tis_ctx *c, int32_t isys, int32_t which, double *buf) {
  if (!c || !buf) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_state_get needs one system");
  if (which < 0 || which > 3) return fail(c, TIS_ERR_ARG, "bad state index %d", which);
  HIPCHK(c, hipSetDevice(c->device));
  const HostSys &h = c->hs[isys];
  const double *base = which == 0 ? Xc(c) : (which == 1 ? c->P.v : (which == 2 ? c->P.a : c->P.f));
  std::vector<double> comp(h.d.nmp);
  for (int d = 0; d < 3; ++d) {
    HIPCHK(c, hipMemcpyAsync(comp.data(), base + (size_t)d * c->P.ntot + h.d.off, comp.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    c->d2h_bytes += (int64_t)(comp.size() * sizeof(double));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < h.d.nmp; ++i) buf[3 * i + d] = comp[i];
  }
  return TIS_OK;
}

static int rebuild_lists(tis_ctx *c, int32_t isys) {
  Range r = range_of(c, isys);
  for (int attempt = 0; attempt < 4; ++attempt) {
    HIPCHK(c, hipMemsetAsync(c->d_counters + TIS_CNT_NB_OVERFLOW, 0, sizeof(unsigned long long), c->stream));
    launch_lists(c, r, Xc(c), PXc(c), nullptr, 0, 0);
    hipLaunchKernelGGL(k_count_rebuild, dim3(1), dim3(1), 0, c->stream, c->d_counters);
    HIPCHK(c, hipGetLastError());
    unsigned long long ovf = 0;
    HIPCHK(c, hipMemcpyAsync(&ovf, c->d_counters + TIS_CNT_NB_OVERFLOW, sizeof(ovf), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ovf == 0) { c->lists_valid = true; return 0; }
    // grow the per-bead capacity and rebuild
    int newcap = (int)(ovf + ovf / 2 + 16);
    uint32_t *nb = nullptr;
    int rc = dalloc(c, &nb, (size_t)newcap * c->P.ntot);
    if (rc) return rc;
    c->P.nb = nb;  // the old buffer stays in allocs until the batch is freed
    c->P.nb_cap = newcap;
    for (int k = 0; k < 2; ++k)
      if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  }
  (void)hipMemsetAsync(c->d_counters + TIS_CNT_NB_OVERFLOW, 0, sizeof(unsigned long long), c->stream);
  return fail(c, TIS_ERR_OVERFLOW, "neighbour list capacity could not be grown");
}

int32_t tis_neighbor(tis_ctx *c, int32_t isys) {
  if (!c) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  return rebuild_lists(c, isys);
}

int32_t tis_pairs_get(tis_ctx *c, int32_t isys, int32_t which, int32_t *pairs, double *coef, int64_t cap,
                      int64_t *nout) {
  if (!c || !nout) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_pairs_get needs one system");
  if (!c->lists_valid) return fail(c, TIS_ERR_STATE, "lists not built");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const DSys &d = c->hs[isys].d;
  int64_t n = 0;
  if (which == TIS_LIST_EXV) {
    std::vector<int32_t> cnt(d.nmp);
    std::vector<uint32_t> nb((size_t)c->P.nb_cap * d.nmp);
    HIPCHK(c, hipMemcpy(cnt.data(), c->P.nb_cnt + d.off, d.nmp * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int m = 0; m < c->P.nb_cap; ++m)
      HIPCHK(c, hipMemcpy(nb.data() + (size_t)m * d.nmp, c->P.nb + (size_t)m * c->P.ntot + d.off,
                          d.nmp * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < d.nmp; ++i)
      for (int m = 0; m < std::min(cnt[i], c->P.nb_cap); ++m) {
        uint32_t e = nb[(size_t)m * d.nmp + i];
        int j = (int)(e & kIdxMask) - d.off;
        if (j <= i) continue;
        if (n < cap && pairs) { pairs[3 * n] = i; pairs[3 * n + 1] = j; pairs[3 * n + 2] = (int32_t)(e >> kMirrorShift); }
        ++n;
      }
  } else if (which == TIS_LIST_ELE) {
    std::vector<uint32_t> mask((size_t)d.nq * d.mwords);
    std::vector<int32_t> qb(d.nq);
    std::vector<double> q(d.nmp);
    std::vector<uint8_t> mir(d.periodic ? (size_t)d.nq * d.nq : 0);
    if (!mask.empty()) HIPCHK(c, hipMemcpy(mask.data(), c->P.qmask + d.moff, mask.size() * 4, hipMemcpyDeviceToHost));
    if (d.nq) HIPCHK(c, hipMemcpy(qb.data(), c->P.qbead + d.qoff, d.nq * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(q.data(), c->P.charge + d.off, d.nmp * 8, hipMemcpyDeviceToHost));
    if (!mir.empty()) HIPCHK(c, hipMemcpy(mir.data(), c->P.qmirror + d.mioff, mir.size(), hipMemcpyDeviceToHost));
    for (int qi = 0; qi < d.nq; ++qi)
      for (int qj = qi + 1; qj < d.nq; ++qj) {
        if (!((mask[(size_t)qi * d.mwords + (qj >> 5)] >> (qj & 31)) & 1u)) continue;
        if (n < cap && pairs) {
          int bi = qb[qi] - d.off, bj = qb[qj] - d.off;
          pairs[3 * n] = bi; pairs[3 * n + 1] = bj;
          pairs[3 * n + 2] = d.periodic ? mir[(size_t)qi * d.nq + qj] : 0;
          if (coef) coef[n] = (q[bi] * q[bj]) * d.ele_coef;
        }
        ++n;
      }
  } else if (which == TIS_LIST_HB) {
    std::vector<uint8_t> in(std::max(d.nhb, 1));
    if (d.nhb) HIPCHK(c, hipMemcpy(in.data(), c->P.hb_inlist + d.hoff, d.nhb, hipMemcpyDeviceToHost));
    for (int k = 0; k < d.nhb; ++k)
      if (in[k]) { if (n < cap && pairs) pairs[n] = k; ++n; }
  } else {
    return fail(c, TIS_ERR_ARG, "unknown list %d", which);
  }
  *nout = n;
  return n > cap ? fail(c, TIS_ERR_OVERFLOW, "list of %lld entries exceeds cap %lld", (long long)n, (long long)cap)
                 : TIS_OK;
}

// k_force (terms, electrostatic chunk pairs, bonded terms), the Ewald
// passes, then k_main.  Steps draw noise/valence numbers for step
// stepbase+t+1; a stand-alone force evaluation or the initial acceleration
// uses stepbase.
static void launch_force_pair(tis_ctx *c, const Range &r, const double *X, const double *PX, double *XN,
                              double *PXN, int t_local, int mode, int count_step = 0) {
  const int32_t pre_t = mode == 1 ? (int32_t)t_local : -1;  // valence draws at the force step
  // timing events on this stream: e[0] start, e[3] before k_main; e[1]/e[2]
  // and e[4]/e[5] are bound to the k_force and k_main dispatches (start and
  // end of the kernel itself, no marker packets)
  hipEvent_t e3[kTev] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (c->timing) {
    while (c->tev.size() < c->tev_used + kTev) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      c->tev.push_back(e);
    }
    for (int k = 0; k < kTev; ++k) e3[k] = c->tev[c->tev_used + k];
    c->tev_used += kTev;
    (void)hipEventRecord(e3[0], c->stream);
  }
  // k_force: hbond lanes, per-system tertiary stacks and valence-limited
  // hbonds, the electrostatic chunk pairs, then the bonded terms in blocks of
  // 256 lanes
  // (launch shape: chunk pairs per wave, bonded terms per lane and the
  // order of the chunk pairs; tis_tune chooses it, no result depends on it)
  const int nterm = r.nb + r.na + r.ns;
  const int npt = c->any_ele ? r.ptn : 0;
  const int ppw = c->shape.ppw, tpl = c->shape.tpl;
  const int nblk = (r.hdn + kTile * kWaves - 1) / (kTile * kWaves) + (r.taskn + kWaves - 1) / kWaves +
                   (npt + kWaves * ppw - 1) / (kWaves * ppw) + (nterm + kTile * kWaves * tpl - 1) / (kTile * kWaves * tpl);
  if (nblk > 0) {
    const BondedRange br{r.b0, r.nb, r.a0, r.na, r.s0, r.ns, r.hd0, r.hdn};
    const int32_t *task = c->d_systask + r.task0;
    const EwTask *pt = r.all ? (c->shape.lpt ? c->d_ptask_lpt : c->d_ptask) : c->d_ptask + r.pt0;
    // one kernel per launch shape (each with its own register allocation)
    auto kf = ppw == 1 ? (tpl == 1 ? k_force<1, 1> : k_force<1, 2>) : (tpl == 1 ? k_force<2, 1> : k_force<2, 2>);
    if (c->timing)
      hipExtLaunchKernelGGL(kf, dim3(nblk), dim3(kTile * kWaves), 0, c->stream, e3[1], e3[2], 0, c->P, X, PX,
                            task, r.taskn, pt, npt, br, pre_t);
    else
      hipLaunchKernelGGL(kf, dim3(nblk), dim3(kTile * kWaves), 0, c->stream, c->P, X, PX, task, r.taskn, pt, npt,
                         br, pre_t);
  } else if (c->timing) {
    (void)hipEventRecord(e3[1], c->stream);
    (void)hipEventRecord(e3[2], c->stream);
  }
  if (r.ekn > 0) {
    hipLaunchKernelGGL(k_ewald_sk, dim3(r.ekn), dim3(kTile), 0, c->stream, c->P, PX, c->d_ewsk + r.ek0, 0);
    hipLaunchKernelGGL(k_ewald_f, dim3(r.efn), dim3(kTile * kWaves), 0, c->stream, c->P, PX, c->d_ewf + r.ef0);
  }
  if (c->timing) (void)hipEventRecord(e3[3], c->stream);
  const int32_t ele_on = c->any_ele ? 1 : 0;
  if (r.btn > 0 && c->timing)
    hipExtLaunchKernelGGL(k_main, dim3(r.btn), dim3(kTile * 2), 0, c->stream, e3[4], e3[5], 0, c->P, X, PX, XN,
                          PXN, r.all ? c->d_btiles_xcd : c->d_btiles + r.bt0, r.btn, (int32_t)t_local, mode, ele_on,
                          count_step);
  else if (r.btn > 0)
    hipLaunchKernelGGL(k_main, dim3(r.btn), dim3(kTile * 2), 0, c->stream, c->P, X, PX, XN, PXN,
                       r.all ? c->d_btiles_xcd : c->d_btiles + r.bt0, r.btn, (int32_t)t_local, mode, ele_on,
                       count_step);
  else if (c->timing) {
    (void)hipEventRecord(e3[4], c->stream);
    (void)hipEventRecord(e3[5], c->stream);
  }
}

int32_t tis_force_sumup(tis_ctx *c, int32_t isys, double *force_out) {
  if (!c) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->lists_valid) { rc = rebuild_lists(c, -1); if (rc) return rc; }
  Range r = range_of(c, isys);
  // force evaluation at the current coordinates; valence draws use step = stepbase
  launch_force_pair(c, r, Xc(c), PXc(c), nullptr, nullptr, 0, 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (isys >= 0 && force_out) return tis_state_get(c, isys, 3, force_out);
  return TIS_OK;
}

int32_t tis_force_get(tis_ctx *c, int32_t isys, double *force_out) {
  if (!c || isys < 0 || !force_out) return TIS_ERR_ARG;
  return tis_state_get(c, isys, 3, force_out);
}

int32_t tis_energy_sumup(tis_ctx *c, int32_t isys, double *energy_out) {
  if (!c || !energy_out) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->lists_valid) { rc = rebuild_lists(c, -1); if (rc) return rc; }
  // structure factors at the current coordinates (FP64 sin/cos)
  const int ek0 = isys < 0 ? 0 : c->hs[isys].esk_begin, ekn = isys < 0 ? c->n_ewsk : c->hs[isys].esk_count;
  if (ekn > 0)
    hipLaunchKernelGGL(k_ewald_sk, dim3(ekn), dim3(kTile), 0, c->stream, c->P, PXc(c), c->d_ewsk + ek0, 1);
  // one block per live system (in live-list order); the output is by
  // system index, zeros for retired systems
  const int nsel = isys < 0 ? c->nlive : 1;
  hipLaunchKernelGGL(k_energy, dim3(nsel), dim3(256), 0, c->stream, c->P, Xc(c), PXc(c),
                     c->d_sysidx + (isys < 0 ? 0 : c->hs[isys].live_pos), c->d_energy, 0);
  HIPCHK(c, hipGetLastError());
  std::vector<double> e(25 * (size_t)nsel);
  HIPCHK(c, hipMemcpyAsync(e.data(), c->d_energy, e.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (isys >= 0) {
    std::memcpy(energy_out, e.data(), 25 * sizeof(double));
  } else {
    for (int is = 0; is < c->nsys; ++is) {
      const HostSys &h = c->hs[is];
      for (int k = 0; k < 25; ++k) energy_out[25 * (size_t)is + k] = h.live ? e[25 * (size_t)h.live_pos + k] : 0.0;
    }
  }
  return TIS_OK;
}

int32_t tis_energy_sumup_set(tis_ctx *c, int32_t n, const int32_t *isys, double *energy_out) {
  if (!c || !isys || !energy_out || n < 1) return TIS_ERR_ARG;
  for (int k = 0; k < n; ++k) {
    if (isys[k] < 0) return fail(c, TIS_ERR_ARG, "tis_energy_sumup_set: system index %d", isys[k]);
    const int rc = check_sys(c, isys[k]);
    if (rc) return rc;
  }
  HIPCHK(c, hipSetDevice(c->device));
  int rc = 0;
  if (!c->lists_valid) { rc = rebuild_lists(c, -1); if (rc) return rc; }
  // structure factors of the selected Ewald systems
  for (int k = 0; k < n; ++k) {
    const HostSys &h = c->hs[isys[k]];
    if (h.esk_count > 0)
      hipLaunchKernelGGL(k_ewald_sk, dim3(h.esk_count), dim3(kTile), 0, c->stream, c->P, PXc(c),
                         c->d_ewsk + h.esk_begin, 1);
  }
  // k_energy takes a list of system indices, one block each (the same
  // kernel as the whole-batch launch)
  if ((size_t)n > c->eset_cap) {
    if (c->d_eset) { (void)hipFree(c->d_eset); c->d_eset = nullptr; c->dev_bytes -= c->eset_cap * sizeof(int32_t); }
    c->eset_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_eset, (size_t)n * sizeof(int32_t)));
    c->eset_cap = (size_t)n;
    c->dev_bytes += c->eset_cap * sizeof(int32_t);
  }
  if ((size_t)n * 25 > c->eset_out_cap) {
    if (c->d_eset_out) { (void)hipFree(c->d_eset_out); c->d_eset_out = nullptr; c->dev_bytes -= c->eset_out_cap * sizeof(double); }
    c->eset_out_cap = 0;
    HIPCHK(c, hipMalloc(&c->d_eset_out, (size_t)n * 25 * sizeof(double)));
    c->eset_out_cap = (size_t)n * 25;
    c->dev_bytes += c->eset_out_cap * sizeof(double);
  }
  HIPCHK(c, hipMemcpyAsync(c->d_eset, isys, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_energy, dim3(n), dim3(256), 0, c->stream, c->P, Xc(c), PXc(c), c->d_eset, c->d_eset_out, 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(energy_out, c->d_eset_out, (size_t)n * 25 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TIS_OK;
}

int32_t tis_energy_unit(tis_ctx *c, int32_t isys, int32_t style, double *energy_unit) {
  if (!c || !energy_unit || (style != 0 && style != 1)) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_energy_unit needs one system");
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->lists_valid) { rc = rebuild_lists(c, -1); if (rc) return rc; }
  const int nu = c->hs[isys].d.nunit;
  const size_t nn = (size_t)nu * nu;
  if (nu < 1 || nn > 65535) return fail(c, TIS_ERR_ARG, "system %d has %d units", isys, nu);
  if (c->eunit_cap < nn * 25) {
    if (c->d_eunit) {
      (void)hipFree(c->d_eunit);
      c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void *)c->d_eunit), c->allocs.end());
      c->dev_bytes -= c->eunit_cap * sizeof(double);
      c->d_eunit = nullptr;
    }
    rc = dalloc(c, &c->d_eunit, nn * 25);
    if (rc) return rc;
    c->eunit_cap = nn * 25;
  }
  hipLaunchKernelGGL(k_energy, dim3(1, (unsigned)nn), dim3(256), 0, c->stream, c->P, Xc(c), PXc(c),
                     c->d_sysidx + c->hs[isys].live_pos, c->d_eunit, 1);
  HIPCHK(c, hipGetLastError());
  std::vector<double> raw(nn * 25);
  HIPCHK(c, hipMemcpyAsync(raw.data(), c->d_eunit, raw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // energy_unit(iunit, junit, type), column-major, from the raw block sums
  // raw[(a + nu*b)*25 + k]
  std::vector<double> m(nn * 25);
  for (int b = 0; b < nu; ++b)
    for (int a = 0; a < nu; ++a)
      for (int k = 0; k < 25; ++k) m[(size_t)a + (size_t)nu * b + nn * k] = raw[((size_t)a + (size_t)nu * b) * 25 + k];
  auto E = [&](int a, int b, int k) -> double & { return m[(size_t)a + (size_t)nu * b + nn * k]; };
  for (int a = 0; a < nu; ++a)
    for (int b = 0; b < nu; ++b) E(a, b, TIS_E_TOTAL) = 0.0;
  // energy_sumup.F90:480-508
  if (style == 0) {
    for (int iu = 0; iu < nu; ++iu) {
      for (int ju = 0; ju < iu; ++ju)
        for (int k = 0; k < 25; ++k) E(iu, iu, k) = E(iu, iu, k) + 0.5 * E(ju, iu, k);
      for (int ju = iu + 1; ju < nu; ++ju)
        for (int k = 0; k < 25; ++k) E(iu, iu, k) = E(iu, iu, k) + 0.5 * E(iu, ju, k);
    }
  } else {
    for (int iu = 0; iu < nu; ++iu)
      for (int ju = iu + 1; ju < nu; ++ju)
        for (int k = 0; k < 25; ++k) {
          const double sume = E(iu, ju, k) + E(ju, iu, k);
          E(iu, ju, k) = sume;
          E(ju, iu, k) = sume;
        }
  }
  for (int k = 0; k < 25; ++k) {
    if (k == TIS_E_TOTAL || k == TIS_E_VELO) continue;
    for (size_t q = 0; q < nn; ++q) m[q + nn * TIS_E_TOTAL] += m[q + nn * k];
  }
  std::memcpy(energy_unit, m.data(), m.size() * sizeof(double));
  return TIS_OK;
}

int32_t tis_system_origin(tis_ctx *c, int32_t isys, uint64_t step_origin, int32_t init_accel) {
  if (!c) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_system_origin needs one system");
  HIPCHK(c, hipSetDevice(c->device));
  HostSys &h = c->hs[isys];
  h.d.step_origin = step_origin;
  h.d.init_accel = init_accel ? 1 : 0;
  // kernels (and captured graphs) read DSys from device memory
  HIPCHK(c, hipMemcpyAsync(c->d_sys + isys, &h.d, sizeof(DSys), hipMemcpyHostToDevice, c->stream));
  c->h2d_bytes += (int64_t)sizeof(DSys);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TIS_OK;
}

int32_t tis_label_define(tis_ctx *c, int32_t label, int32_t isys, const tis_system *s) {
  if (!c || !s || label < 0 || label > 1000000) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_label_define needs a template system");
  const HostSys &h = c->hs[isys];
  if (s->nmp != h.d.nmp || s->nst != h.d.nst) return fail(c, TIS_ERR_ARG, "label topology differs from the template");
  if ((int)c->labels.size() <= label) c->labels.resize(label + 1);
  tis_ctx::Label &L = c->labels[label];
  L.set = true;
  L.d = h.d;
  fill_derived(L.d, *s);
  L.nmp = s->nmp;
  L.nst = s->nst;
  L.q.assign(s->charge, s->charge + s->nmp);
  L.u0.resize(s->nst);
  for (int t = 0; t < s->nst; ++t) L.u0[t] = s->st_coef[4 * t];
  c->labels_dirty = true;
  return TIS_OK;
}

int32_t tis_label_apply(tis_ctx *c, int32_t n, const int32_t *isys, const int32_t *label, const double *vscale) {
  if (!c || n < 0 || (n > 0 && (!isys || !label))) return TIS_ERR_ARG;
  if (n == 0) return TIS_OK;
  int rc = check_sys(c, -1);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) {
    if (isys[k] < 0 || isys[k] >= c->nsys || !c->hs[isys[k]].live)
      return fail(c, TIS_ERR_ARG, "system %d out of range or retired", isys[k]);
    if (label[k] < 0 || label[k] >= (int)c->labels.size() || !c->labels[label[k]].set)
      return fail(c, TIS_ERR_ARG, "label %d not defined", label[k]);
    const tis_ctx::Label &L = c->labels[label[k]];
    if (L.nmp != c->hs[isys[k]].d.nmp || L.nst != c->hs[isys[k]].d.nst)
      return fail(c, TIS_ERR_ARG, "label %d does not fit system %d", label[k], isys[k]);
    for (int j = 0; j < k; ++j)
      if (isys[j] == isys[k]) return fail(c, TIS_ERR_ARG, "system %d listed twice", isys[k]);
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (c->labels_dirty) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (void *p : c->label_bufs) (void)hipFree(p);
    c->label_bufs.clear();
    std::vector<DSys> ds;
    std::vector<double> q, u0;
    std::vector<int64_t> qo, uo;
    for (const auto &L : c->labels) {
      ds.push_back(L.d);
      qo.push_back((int64_t)q.size());
      uo.push_back((int64_t)u0.size());
      q.insert(q.end(), L.q.begin(), L.q.end());
      u0.insert(u0.end(), L.u0.begin(), L.u0.end());
    }
    auto up = [&](const void *src, size_t bytes, void **dst) -> int {
      HIPCHK(c, hipMalloc(dst, std::max<size_t>(bytes, 8)));
      c->label_bufs.push_back(*dst);
      if (bytes) HIPCHK(c, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
      return 0;
    };
    void *p;
    if ((rc = up(ds.data(), ds.size() * sizeof(DSys), &p))) return rc; c->label_dev.sys = (const DSys *)p;
    if ((rc = up(q.data(), q.size() * sizeof(double), &p))) return rc; c->label_dev.q = (const double *)p;
    if ((rc = up(qo.data(), qo.size() * sizeof(int64_t), &p))) return rc; c->label_dev.qoff = (const int64_t *)p;
    if ((rc = up(u0.data(), u0.size() * sizeof(double), &p))) return rc; c->label_dev.u0 = (const double *)p;
    if ((rc = up(uo.data(), uo.size() * sizeof(int64_t), &p))) return rc; c->label_dev.u0off = (const int64_t *)p;
    c->labels_dirty = false;
  }
  if (c->lab_cap < n) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->d_lab_args) (void)hipFree(c->d_lab_args);
    if (c->d_lab_scale) (void)hipFree(c->d_lab_scale);
    c->d_lab_args = nullptr; c->d_lab_scale = nullptr;
    HIPCHK(c, hipMalloc((void **)&c->d_lab_args, 2 * sizeof(int32_t) * (size_t)n));
    HIPCHK(c, hipMalloc((void **)&c->d_lab_scale, sizeof(double) * (size_t)n));
    c->lab_cap = n;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_lab_args, isys, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_lab_args + c->lab_cap, label, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  if (vscale) HIPCHK(c, hipMemcpyAsync(c->d_lab_scale, vscale, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_label_apply, dim3(n), dim3(256), 0, c->stream, c->P, c->label_dev, c->d_lab_args,
                     c->d_lab_args + c->lab_cap, vscale ? (const double *)c->d_lab_scale : nullptr);
  HIPCHK(c, hipGetLastError());
  for (int k = 0; k < n; ++k) copy_t_fields(c->hs[isys[k]].d, c->labels[label[k]].d);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TIS_OK;
}

int32_t tis_transfer_bytes(tis_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return TIS_ERR_ARG;
  bytes[0] = c->h2d_bytes;
  bytes[1] = c->h2d_state;
  bytes[2] = c->d2h_bytes;
  bytes[3] = c->d2d_state;
  return TIS_OK;
}

int32_t tis_device_bytes(tis_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return TIS_ERR_ARG;
  *bytes = (int64_t)c->dev_bytes;
  return TIS_OK;
}

int32_t tis_langevin_init(tis_ctx *c, double tstep, int32_t rebuild_every, uint64_t step0, int32_t init_accel) {
  if (!c || tstep <= 0.0 || rebuild_every < 0) return fail(c, TIS_ERR_ARG, "bad Langevin parameters");
  int rc = check_sys(c, -1);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  c->tstep = tstep;
  c->rebuild_every = rebuild_every;
  c->P.dyn = rebuild_every == 0 ? 1 : 0;
  c->P.tstep2 = 0.5 * tstep * tstep;
  c->P.tstep = tstep;
  for (int k = 0; k < 2; ++k)
    if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  // Langevin constants of every system's (friction, mass) classes
  std::vector<LangClass> lc(c->stage.lcls);
  for (int is = 0; is < c->nsys; ++is) {
    const HostSys &h = c->hs[is];
    if (!h.live) continue;
    for (int k = 0; k < h.d.nlc; ++k)
      langevin_consts(tstep, h.d.tempk, h.lc_fric[k], h.lc_mass[k], lc[h.d.lc0 + k].r);
  }
  c->h2d_bytes += (int64_t)(lc.size() * sizeof(LangClass));
  if (!lc.empty())
    HIPCHK(c, hipMemcpyAsync((void *)c->P.lcls, lc.data(), lc.size() * sizeof(LangClass), hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_stepbase, &step0, sizeof(step0), hipMemcpyHostToDevice, c->stream));
  c->step_host = step0;
  rc = rebuild_lists(c, -1);
  if (rc) return rc;
  // displacements since this rebuild start at zero, no rebuild pending
  HIPCHK(c, hipMemsetAsync(c->P.dxyz, 0, 3 * sizeof(double) * (size_t)std::max(c->P.ntot, 1), c->stream));
  HIPCHK(c, hipMemsetAsync(c->P.sys_flag, 0, (size_t)c->caps.nsys, c->stream));
  HIPCHK(c, hipMemsetAsync(c->P.ttop, 0, 2 * sizeof(double) * (size_t)std::max(c->nbtile, 1), c->stream));
  Range r = range_of(c, -1);
  const int cu = c->cur;
  if (init_accel < 0) {
    // per system: new accelerations (mode 2) or kept ones (mode 3)
    launch_force_pair(c, r, c->P.x[cu], c->P.px[cu], c->P.x[cu ^ 1], c->P.px[cu ^ 1], 0, 4);
  } else if (init_accel) {
    // a = F(x)/m + c1 xi(step0), then x_pred (mode 2)
    launch_force_pair(c, r, c->P.x[cu], c->P.px[cu], c->P.x[cu ^ 1], c->P.px[cu ^ 1], 0, 2);
  } else {
    // keep the stored a; forces at x give the hbond state, then predict
    launch_force_pair(c, r, c->P.x[cu], c->P.px[cu], c->P.x[cu ^ 1], c->P.px[cu ^ 1], 0, 3);
  }
  HIPCHK(c, hipGetLastError());
  if (c->any_periodic)
    hipLaunchKernelGGL(k_prep, dim3(c->nbtile), dim3(kTile), 0, c->stream, c->P, c->P.x[cu], c->P.px[cu],
                       c->P.px[cu ^ 1], c->d_btiles, 0, c->nbtile, c->d_qtiles, 0, 0, 1, 0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->lv_init = true;
  c->t_since_init = 0;
  return TIS_OK;
}

// The list part of one Langevin step of the batch (or of system isys) in
// eager launches: the displacement trigger's rebuilds, or the whole range
// every rebuild_every steps.
static void eager_lists(tis_ctx *c, int32_t isys) {
  const int cu = c->cur;
  Range r = range_of(c, isys);
  if (c->P.dyn) {
    launch_lists(c, r, c->P.x[cu], c->P.px[cu], c->P.px[cu ^ 1], 1, 1);
  } else if (c->t_since_init % c->rebuild_every == 0) {
    launch_lists(c, r, c->P.x[cu], c->P.px[cu], c->P.px[cu ^ 1], 1, 0);
    if (isys <= 0) hipLaunchKernelGGL(k_count_rebuild, dim3(1), dim3(1), 0, c->stream, c->d_counters);
  }
}

// The force and integration part (count_step: this launch completes the step).
static void eager_forces(tis_ctx *c, int32_t isys, int count_step) {
  const int cu = c->cur;
  Range r = range_of(c, isys);
  launch_force_pair(c, r, c->P.x[cu ^ 1], c->P.px[cu ^ 1], c->P.x[cu], c->P.px[cu], 0, 1, count_step);
}

static int eager_step(tis_ctx *c, int32_t isys) {
  eager_lists(c, isys);
  eager_forces(c, isys, 1);
  return 0;
}

// steps per captured graph: even (the ping-pong parity returns) and a
// multiple of the fixed rebuild interval
static int steps_per_graph(const tis_ctx *c) {
  if (c->P.dyn) return 10;
  return (c->rebuild_every % 2 == 0) ? c->rebuild_every : 2 * c->rebuild_every;
}

static int capture_graph(tis_ctx *c, int parity) {
  const int L = steps_per_graph(c);
  hipGraph_t g = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int cu = parity;
  Range r = range_of(c, -1);
  for (int t = 0; t < L; ++t) {
    if (c->P.dyn) {
      launch_lists(c, r, c->P.x[cu], c->P.px[cu], c->P.px[cu ^ 1], 1, 1);
    } else if (t % c->rebuild_every == 0) {
      launch_lists(c, r, c->P.x[cu], c->P.px[cu], c->P.px[cu ^ 1], 1, 0);
      hipLaunchKernelGGL(k_count_rebuild, dim3(1), dim3(1), 0, c->stream, c->d_counters);
    }
    launch_force_pair(c, r, c->P.x[cu ^ 1], c->P.px[cu ^ 1], c->P.x[cu], c->P.px[cu], t, 1, 1);
    cu ^= 1;
  }
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_stepbase, (uint64_t)L, c->d_counters);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return fail(c, TIS_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&c->graph[parity], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, TIS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  c->graph_len = L;
  return 0;
}

// The steps of one run in the given mode; a stop (list overflow) makes the
// remaining launches return at once.
static int run_steps(tis_ctx *c, int64_t nsteps, int32_t mode, int64_t *nlaunch) {
  int64_t done = 0;
  const int L = steps_per_graph(c);
  if (mode == TIS_MODE_GRAPH) {
    while (done < nsteps) {
      if (c->t_since_init % L == 0 && nsteps - done >= L) {
        if (!c->graph[c->cur]) { int rc = capture_graph(c, c->cur); if (rc) return rc; }
        HIPCHK(c, hipGraphLaunch(c->graph[c->cur], c->stream));
        done += L; c->t_since_init += L;  // L even: parity unchanged
        *nlaunch += 1;
      } else {
        eager_step(c, -1);
        hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_stepbase, (uint64_t)1, c->d_counters);
        c->cur ^= 1; ++done; ++c->t_since_init; *nlaunch += 2;
      }
    }
  } else if (mode == TIS_MODE_EAGER) {
    c->timing = true;
    c->tev_used = 0;
    for (; done < nsteps; ++done) {
      eager_step(c, -1);
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_stepbase, (uint64_t)1, c->d_counters);
      c->cur ^= 1; ++c->t_since_init; *nlaunch += 2;
    }
  } else {
    // one system per launch: every system's lists, then every system's
    // force and integration launches (the last one completes the step)
    int last = -1;
    for (int is = 0; is < c->nsys; ++is)
      if (c->hs[is].live) last = is;
    for (; done < nsteps; ++done) {
      for (int is = 0; is < c->nsys; ++is)
        if (c->hs[is].live) eager_lists(c, is);
      for (int is = 0; is < c->nsys; ++is)
        if (c->hs[is].live) eager_forces(c, is, is == last);
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_stepbase, (uint64_t)1, c->d_counters);
      c->cur ^= 1; ++c->t_since_init; *nlaunch += 2 * c->nlive;
    }
  }
  HIPCHK(c, hipGetLastError());
  return TIS_OK;
}

// Larger per-bead list capacity; the entries built so far are kept (the
// layout nb[slot][ntot] keeps its stride), so systems whose lists are not
// rebuilt keep theirs.  Captured graphs hold the old capacity: dropped.
static int grow_nb(tis_ctx *c, unsigned long long need) {
  const int newcap = (int)(need + need / 2 + 16);
  if (newcap <= c->P.nb_cap) return TIS_OK;
  uint32_t *nb = nullptr;
  int rc = dalloc(c, &nb, (size_t)newcap * c->P.ntot);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(nb, c->P.nb, sizeof(uint32_t) * (size_t)c->P.nb_cap * c->P.ntot, hipMemcpyDeviceToDevice,
                           c->stream));
  c->P.nb = nb;  // the old buffer stays in allocs until the batch is freed
  c->P.nb_cap = newcap;
  for (int k = 0; k < 2; ++k)
    if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  return TIS_OK;
}

// ---------------------------------------------------------------- launch-shape tuner
// The force pass of the whole batch at the current coordinates (mode 0:
// forces only, no state changes) under each candidate launch shape: after
// warm-up passes, reps rounds that time every candidate once (events bound
// to the k_force dispatch), so that clock ramps and drifts fall on all
// candidates alike; the candidate with the smallest median is kept.  The
// counters the passes bump are put back.
int32_t tis_tune(tis_ctx *c, int32_t reps, int32_t *shape_out, double *ms_out) {
  if (!c || reps < 1 || reps > 1000) return TIS_ERR_ARG;
  int rc = check_sys(c, -1);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->lists_valid) { rc = rebuild_lists(c, -1); if (rc) return rc; }
  unsigned long long cnt[TIS_CNT_MAX];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
  constexpr int kCand = 8;
  const tis_ctx::Shape cand[kCand] = {{1, 1, 1}, {2, 1, 1}, {1, 2, 1}, {2, 2, 1},
                                      {1, 1, 0}, {2, 1, 0}, {1, 2, 0}, {2, 2, 0}};
  const tis_ctx::Shape keep = c->shape;
  const Range r = range_of(c, -1);
  for (int w = 0; w < 4; ++w)
    for (int k = 0; k < kCand; ++k) {
      c->shape = cand[k];
      launch_force_pair(c, r, Xc(c), PXc(c), nullptr, nullptr, 0, 0);
    }
  std::vector<std::vector<double>> t(kCand);
  for (int rep = 0; rep < reps; ++rep) {
    // an untimed pass after the host's wait, then every candidate once,
    // starting from a different one each round
    launch_force_pair(c, r, Xc(c), PXc(c), nullptr, nullptr, 0, 0);
    c->timing = true;
    c->tev_used = 0;
    for (int i = 0; i < kCand; ++i) {
      c->shape = cand[(rep + i) % kCand];
      launch_force_pair(c, r, Xc(c), PXc(c), nullptr, nullptr, 0, 0);
    }
    c->timing = false;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
      c->shape = keep;
      return fail(c, TIS_ERR_HIP, "tuning pass failed");
    }
    for (int i = 0; i < kCand; ++i) {
      float e = 0.0f;
      if (hipEventElapsedTime(&e, c->tev[kTev * i + 1], c->tev[kTev * i + 2]) != hipSuccess) e = 0.0f;
      t[(rep + i) % kCand].push_back(e);
    }
    c->tev_used = 0;
  }
  c->tune_ms.assign(kCand, 0.0);
  int best = 0;
  for (int k = 0; k < kCand; ++k) {
    std::sort(t[k].begin(), t[k].end());
    const size_t m = t[k].size();
    c->tune_ms[k] = m % 2 ? t[k][m / 2] : 0.5 * (t[k][m / 2 - 1] + t[k][m / 2]);
    if (c->tune_ms[k] < c->tune_ms[best]) best = k;
  }
  c->shape = cand[best];
  HIPCHK(c, hipMemcpy(c->d_counters, cnt, sizeof(cnt), hipMemcpyHostToDevice));
  for (int k = 0; k < 2; ++k)
    if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  if (shape_out) { shape_out[0] = c->shape.ppw; shape_out[1] = c->shape.tpl; shape_out[2] = c->shape.lpt; }
  if (ms_out)
    for (int k = 0; k < kCand; ++k) ms_out[k] = c->tune_ms[k];
  return TIS_OK;
}

int32_t tis_launch_shape(tis_ctx *c, const int32_t *set, int32_t *get) {
  if (!c) return TIS_ERR_ARG;
  if (set) {
    if ((set[0] != 1 && set[0] != 2) || (set[1] != 1 && set[1] != 2) || (set[2] != 0 && set[2] != 1))
      return fail(c, TIS_ERR_ARG, "launch shape (%d, %d, %d) not offered", set[0], set[1], set[2]);
    c->shape = tis_ctx::Shape{set[0], set[1], set[2]};
    for (int k = 0; k < 2; ++k)
      if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  }
  if (get) { get[0] = c->shape.ppw; get[1] = c->shape.tpl; get[2] = c->shape.lpt; }
  return TIS_OK;
}

int32_t tis_term_status(tis_ctx *c, int32_t isys, int32_t *hb_status, int32_t *st_status) {
  if (!c) return TIS_ERR_ARG;
  int rc = check_sys(c, isys);
  if (rc) return rc;
  if (isys < 0) return fail(c, TIS_ERR_ARG, "tis_term_status needs one system");
  HIPCHK(c, hipSetDevice(c->device));
  const DSys &d = c->hs[isys].d;
  if (hb_status && d.nhb > 0) {
    std::vector<uint8_t> h((size_t)d.nhb);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->P.hb_status + d.hoff, h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < d.nhb; ++k) hb_status[k] = h[k];
  }
  if (st_status && d.nst > 0) {
    hipLaunchKernelGGL(k_stack_status, dim3((d.nst + 255) / 256), dim3(256), 0, c->stream, c->P, Xc(c), PXc(c),
                       d.soff, d.nst, c->d_ststat);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(st_status, c->d_ststat, sizeof(int32_t) * (size_t)d.nst, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return TIS_OK;
}

int32_t tis_list_capacity(tis_ctx *c, int32_t cap, int32_t *cap_out) {
  if (!c) return TIS_ERR_ARG;
  int rc = check_sys(c, -1);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if (cap > 0 && cap != c->P.nb_cap) {
    int need = 0;
    if (c->lists_valid) {
      std::vector<int32_t> cnt((size_t)c->P.ntot);
      HIPCHK(c, hipMemcpy(cnt.data(), c->P.nb_cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost));
      for (int32_t v : cnt) need = std::max(need, v);
    }
    if (cap < need)
      return fail(c, TIS_ERR_ARG, "capacity %d below the %d entries of the current lists", cap, need);
    uint32_t *nb = nullptr;
    rc = dalloc(c, &nb, (size_t)cap * c->P.ntot);
    if (rc) return rc;
    const int keep = std::min(cap, c->P.nb_cap);
    HIPCHK(c, hipMemcpyAsync(nb, c->P.nb, sizeof(uint32_t) * (size_t)keep * c->P.ntot, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->P.nb = nb;
    c->P.nb_cap = cap;
    for (int k = 0; k < 2; ++k)
      if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
  }
  if (cap_out) *cap_out = c->P.nb_cap;
  return TIS_OK;
}

int32_t tis_langevin_steps(tis_ctx *c, int64_t nsteps, int32_t mode) {
  if (!c || nsteps < 0) return TIS_ERR_ARG;
  if (!c->lv_init) return fail(c, TIS_ERR_STATE, "tis_langevin_init not called");
  if (mode != TIS_MODE_GRAPH && mode != TIS_MODE_EAGER && mode != TIS_MODE_PLAIN)
    return fail(c, TIS_ERR_ARG, "unknown mode %d", mode);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemsetAsync(c->d_counters + TIS_CNT_NB_OVERFLOW, 0, 2 * sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  int64_t nlaunch = 0, remaining = nsteps;
  for (int attempt = 0;; ++attempt) {
    const int cur0 = c->cur;
    const int64_t t0 = c->t_since_init;
    const uint64_t step0 = c->step_host;
    int rc = run_steps(c, remaining, mode, &nlaunch);
    if (rc) return rc;
    unsigned long long st[2] = {0ull, 0ull};  // overflow need, steps completed
    HIPCHK(c, hipMemcpyAsync(st, c->d_counters + TIS_CNT_NB_OVERFLOW, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st[0] == 0ull) {
      c->step_host = step0 + (uint64_t)remaining;
      HIPCHK(c, hipMemsetAsync(c->d_counters + TIS_CNT_NB_OVERFLOW, 0, 2 * sizeof(unsigned long long), c->stream));
      break;
    }
    // a rebuild overflowed the list capacity: the run stopped at the start
    // of the step after the st[1] completed ones; put the step counters
    // there, grow the capacity and resume
    if (attempt >= 8) {
      (void)hipMemsetAsync(c->d_counters + TIS_CNT_NB_OVERFLOW, 0, 2 * sizeof(unsigned long long), c->stream);
      return fail(c, TIS_ERR_OVERFLOW, "neighbour list capacity could not be grown");
    }
    const int64_t completed = (int64_t)st[1];
    c->cur = cur0 ^ (int)(completed & 1);
    c->t_since_init = t0 + completed;
    c->step_host = step0 + (uint64_t)completed;
    unsigned long long cnt[TIS_CNT_MAX];
    uint64_t sb = 0;
    HIPCHK(c, hipMemcpy(cnt, c->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&sb, c->d_stepbase, sizeof(sb), hipMemcpyDeviceToHost));
    cnt[TIS_CNT_STEPS] += c->step_host - sb;
    cnt[TIS_CNT_NB_OVERFLOW] = 0ull;
    cnt[TIS_CNT_NB_OVERFLOW + 1] = 0ull;
    cnt[TIS_CNT_RESUMES] += 1ull;
    sb = c->step_host;
    rc = grow_nb(c, st[0]);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(c->d_counters, cnt, sizeof(cnt), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_stepbase, &sb, sizeof(sb), hipMemcpyHostToDevice));
    remaining -= completed;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->last_total_ms = ms;
  c->last_nlaunch = nlaunch;
  c->last_main_ms = c->last_pre_ms = 0.0;
  if (c->timing) {
    double sp = 0.0, sm = 0.0, se = 0.0;
    size_t nstep = c->tev_used / kTev;
    for (size_t k = 0; k < nstep; ++k) {
      float a = 0.0f, b = 0.0f, e = 0.0f;
      if (hipEventElapsedTime(&a, c->tev[kTev * k], c->tev[kTev * k + 3]) != hipSuccess) a = 0.0f;
      if (hipEventElapsedTime(&e, c->tev[kTev * k + 1], c->tev[kTev * k + 2]) != hipSuccess) e = 0.0f;
      if (hipEventElapsedTime(&b, c->tev[kTev * k + 4], c->tev[kTev * k + 5]) != hipSuccess) b = 0.0f;
      sp += a; sm += b; se += e;
    }
    if (nstep) { c->last_pre_ms = sp / nstep; c->last_main_ms = sm / nstep; c->last_force_ms = se / nstep; }
    c->timing = false;
    c->tev_used = 0;
  }
  return TIS_OK;
}

int32_t tis_status_counters(tis_ctx *c, int64_t *counters) {
  if (!c || !counters) return TIS_ERR_ARG;
  if (c->nsys == 0) { for (int k = 0; k < TIS_CNT_MAX; ++k) counters[k] = 0; return TIS_OK; }
  HIPCHK(c, hipSetDevice(c->device));
  unsigned long long h[TIS_CNT_MAX];
  HIPCHK(c, hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < TIS_CNT_MAX; ++k) counters[k] = (int64_t)h[k];
  return TIS_OK;
}

int32_t tis_synchronize(tis_ctx *c) {
  if (!c) return TIS_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TIS_OK;
}

int32_t tis_last_force_ms(tis_ctx *c, double *force_ms) {
  if (!c || !force_ms) return TIS_ERR_ARG;
  *force_ms = c->last_force_ms;
  return TIS_OK;
}

int32_t tis_last_timing(tis_ctx *c, double *total_ms, double *step_kernel_ms, double *pre_kernel_ms,
                        int64_t *nlaunch) {
  if (!c) return TIS_ERR_ARG;
  if (total_ms) *total_ms = c->last_total_ms;
  if (step_kernel_ms) *step_kernel_ms = c->last_main_ms;
  if (pre_kernel_ms) *pre_kernel_ms = c->last_pre_ms;
  if (nlaunch) *nlaunch = c->last_nlaunch;
  return TIS_OK;
}

}  // extern "C"
