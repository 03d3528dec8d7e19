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
#include <string>
#include <vector>

#include "../../include/tis_hip.h"
#include "tis_dev.h"
#include "tis_internal.h"

namespace tis {
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


struct HostSys {
  DSys d;
  int32_t bt_begin = 0, bt_count = 0;   // bead tiles
  int32_t qt_begin = 0, qt_count = 0;   // charge tiles
  bool has_task = false;                // hbonds or tertiary stacks
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
  int32_t nsys = 0, ntot = 0, nqtot = 0, nhbtot = 0, ntsttot = 0, nsttot = 0;
  std::vector<HostSys> hs;
  std::vector<void *> allocs;
  size_t dev_bytes = 0;                  // bytes held in allocs
  DevPtrs P{};
  DSys *d_sys = nullptr;
  Tile *d_btiles = nullptr, *d_qtiles = nullptr;
  int32_t nbtile = 0, nqtile = 0;
  int32_t *d_systask = nullptr, *d_sysidx = nullptr;
  int32_t nsystask = 0;
  std::vector<int32_t> systask_of;       // index into d_systask per system or -1
  int32_t *d_ststat = nullptr;           // st_status output of tis_term_status [nsttot]
  double *d_energy = nullptr;
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
  int64_t h2d_bytes = 0, d2h_bytes = 0;  // payload bytes of uploads / state downloads since creation
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
  c->allocs.clear();
  c->dev_bytes = 0;
  c->hs.clear();
  c->P = DevPtrs{};
  c->nsys = c->ntot = 0;
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
  return 0;
}

double *Xc(tis_ctx *c) { return c->P.x[c->cur]; }
double *PXc(tis_ctx *c) { return c->P.px[c->cur]; }

// --- launch helpers on a system subset: isys = -1 all, else one system
struct Range {
  bool all;                 // the whole batch
  int isys;                 // else this system
  int bt0, btn, qt0, qtn, task0, taskn;
  const int32_t *sysidx; int nsel;
  int b0, nb, a0, na, s0, ns;
  int ek0, ekn, ef0, efn;   // Ewald tasks
  int pt0, ptn;             // periodic electrostatic chunk-pair tasks
};

Range range_of(tis_ctx *c, int32_t isys) {
  Range r{};
  if (isys < 0) {
    r.all = true;
    r.bt0 = 0; r.btn = c->nbtile; r.qt0 = 0; r.qtn = c->nqtile; r.task0 = 0; r.taskn = c->nsystask;
    r.sysidx = c->d_sysidx; r.nsel = c->nsys;
    r.b0 = 0; r.nb = c->P.nbd; r.a0 = 0; r.na = c->P.nba; r.s0 = 0; r.ns = c->P.nst;
    r.ek0 = 0; r.ekn = c->n_ewsk; r.ef0 = 0; r.efn = c->n_ewf;
    r.pt0 = 0; r.ptn = c->n_ptask;
  } else {
    const HostSys &h = c->hs[isys];
    r.isys = isys;
    r.bt0 = h.bt_begin; r.btn = h.bt_count; r.qt0 = h.qt_begin; r.qtn = h.qt_count;
    r.task0 = c->systask_of[isys] < 0 ? 0 : c->systask_of[isys];
    r.taskn = c->systask_of[isys] < 0 ? 0 : 1;
    r.sysidx = c->d_sysidx + isys; r.nsel = 1;
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
  c->nsys = nsys;
  c->hs.resize(nsys);
  // host staging
  std::vector<double> rl2;
  std::vector<int8_t> mptype, iclass;
  std::vector<int32_t> unit, qidx, qbead, inc_off;
  std::vector<uint32_t> inc;
  std::vector<double> exv_rad, charge, mass;
  std::vector<uint8_t> lcl;
  std::vector<LangClass> lcls;
  std::vector<float> exv_eps;
  std::vector<int32_t> bd_mp, ba_mp, st_mp, tst_mp, tst2st, tst_excl, hb_mp, hb_site, hb_tert, site_val,
      hb_sys, tst_sys;
  std::vector<double> bd_nat, bd_coef, ba_nat, ba_coef, st_nat, st_coef, tst_nat, tst_coef, hb_nat, hb_coef;
  std::vector<Tile> btiles, qtiles;
  std::vector<int32_t> systask;
  std::vector<double> ewhx, ewhy, ewhz, ewc;
  std::vector<EwTask> ewsk, ewf, ptask, bpair, qpair;
  int64_t psk_off = 0, pf_off = 0;
  // excluded-volume species: distinct (mptype, class, radius, epsilon)
  struct Spec { int8_t t, c; double r, e; };
  std::vector<Spec> specs;
  std::vector<uint8_t> species;
  auto spec_of = [&](int8_t t, int8_t cl, double r, double e) -> int {
    for (size_t k = 0; k < specs.size(); ++k)
      if (specs[k].t == t && specs[k].c == cl && specs[k].r == r && specs[k].e == e) return (int)k;
    specs.push_back(Spec{t, cl, r, e});
    return (int)specs.size() - 1;
  };
  c->systask_of.assign(nsys, -1);
  int64_t moff = 0, mioff = 0, pp_off = 0, lp_off = 0;
  int32_t off = 0, qoff = 0, nbd = 0, nba = 0, nst = 0, ntst = 0, nhb = 0, nsite = 0, uoff = 0;
  c->any_ele = false;
  c->any_periodic = false;
  for (int is = 0; is < nsys; ++is) {
    const tis_system &s = sys[is];
    HostSys &h = c->hs[is];
    if (s.nmp <= 0 || s.nunit <= 0) return fail(c, TIS_ERR_ARG, "system %d has no beads", is);
    DSys &d = h.d;
    std::memset(&d, 0, sizeof(d));
    fill_derived(d, s);
    d.key = keys ? keys[is] : (uint32_t)is;
    d.off = off; d.nmp = s.nmp;
    d.uoff = uoff; d.nunit = s.nunit;
    d.rl2max = 0.0;
    for (int k = 0; k < s.nunit * s.nunit; ++k) {
      rl2.push_back(s.rlist2_unit[k]);
      d.rl2max = std::max(d.rl2max, s.rlist2_unit[k]);
    }
    uoff += s.nunit * s.nunit;
    if (s.p.periodic) c->any_periodic = true;
    if (s.p.ele_form >= 0) c->any_ele = true;
    if (s.p.ele_form == TIS_ELE_EWALD) {
      if (!s.p.periodic || s.p.box[0] != s.p.box[1] || s.p.box[0] != s.p.box[2])
        return fail(c, TIS_ERR_ARG, "system %d: Ewald needs a periodic cubic box", is);
      if (!(s.p.ewld_alpha > 0.0) || s.p.ewld_hmax <= 0)
        return fail(c, TIS_ERR_ARG, "system %d: Ewald needs ewld_alpha > 0 and ewld_hmax > 0", is);
    }
    if (s.p.ele_form != TIS_ELE_NONE && s.p.ele_form != TIS_ELE_DH && s.p.ele_form != TIS_ELE_COULOMB &&
        s.p.ele_form != TIS_ELE_EWALD)
      return fail(c, TIS_ERR_ARG, "system %d: unknown ele_form %d", is, s.p.ele_form);
    // beads
    d.qoff = qoff;
    int nq = 0;
    for (int i = 0; i < s.nmp; ++i) {
      mptype.push_back((int8_t)s.mptype[i]);
      iclass.push_back((int8_t)s.iclass[i]);
      unit.push_back(s.unit[i]);
      exv_rad.push_back(s.exv_rad[i]);
      exv_eps.push_back((float)s.exv_eps[i]);
      species.push_back((uint8_t)spec_of((int8_t)s.mptype[i], (int8_t)s.iclass[i], s.exv_rad[i], s.exv_eps[i]));
      charge.push_back(s.charge[i]);
      mass.push_back(s.mass[i]);
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
      lcl.push_back((uint8_t)cl);
      // charges of the ele list: |q| >= ZERO_JUDGE (neighbor_list_ele.F90:134,147)
      if (s.p.ele_form != TIS_ELE_NONE && std::fabs(s.charge[i]) >= 1.0e-6) {
        qidx.push_back(qoff + nq);
        qbead.push_back(off + i);
        ++nq;
      } else {
        qidx.push_back(-1);
      }
    }
    // its Langevin classes (constants r[] set by tis_langevin_init)
    d.lc0 = (int32_t)lcls.size();
    d.nlc = (int32_t)h.lc_fric.size();
    for (int k = 0; k < d.nlc; ++k)
      lcls.push_back(LangClass{{0.0, 0.0, 0.0, 0.0}, 1.0 / h.lc_mass[k], h.lc_fric[k], h.lc_mass[k], 0.0});
    d.nq = nq;
    d.mwords = (nq + 31) / 32;
    d.moff = moff;
    moff += (int64_t)nq * d.mwords;
    d.mioff = s.p.periodic ? mioff : -1;
    if (s.p.periodic) mioff += (int64_t)nq * nq;
    qoff += nq;
    // terms with global bead indices
    d.boff = nbd; d.nbd = s.nbd;
    for (int b = 0; b < s.nbd; ++b) {
      bd_mp.push_back(off + s.bd_mp[2 * b]); bd_mp.push_back(off + s.bd_mp[2 * b + 1]);
      bd_nat.push_back(s.bd_nat[b]); bd_coef.push_back(s.bd_coef[b]);
    }
    d.aoff = nba; d.nba = s.nba;
    for (int a = 0; a < s.nba; ++a) {
      for (int k = 0; k < 3; ++k) ba_mp.push_back(off + s.ba_mp[3 * a + k]);
      ba_nat.push_back(s.ba_nat[a]); ba_coef.push_back(s.ba_coef[a]);
    }
    d.soff = nst; d.nst = s.nst;
    for (int t = 0; t < s.nst; ++t) {
      for (int k = 0; k < 7; ++k) st_mp.push_back(off + s.st_mp[7 * t + k]);
      for (int k = 0; k < 3; ++k) st_nat.push_back(s.st_nat[3 * t + k]);
      for (int k = 0; k < 4; ++k) st_coef.push_back(s.st_coef[4 * t + k]);
    }
    d.toff = ntst; d.ntst = s.ntst;
    for (int t = 0; t < s.ntst; ++t) {
      for (int k = 0; k < 6; ++k) tst_mp.push_back(off + s.tst_mp[6 * t + k]);
      for (int k = 0; k < 6; ++k) tst_nat.push_back(s.tst_nat[6 * t + k]);
      for (int k = 0; k < 7; ++k) tst_coef.push_back(s.tst_coef[7 * t + k]);
      for (int k = 0; k < 2; ++k) {
        tst2st.push_back(s.tst2st[2 * t + k] >= 0 ? nst + s.tst2st[2 * t + k] : -1);
        tst_excl.push_back(s.tst_excl[2 * t + k]);
      }
      tst_sys.push_back(is);
    }
    d.sitoff = nsite; d.nsite = s.nhbsite;
    for (int k = 0; k < s.nhbsite; ++k) site_val.push_back(s.hbsite_valence[k]);
    d.hoff = nhb; d.nhb = s.nhb;
    std::vector<int> site_use(std::max(s.nhbsite, 1), 0);
    for (int t = 0; t < s.nhb; ++t) {
      for (int k = 0; k < 6; ++k) hb_mp.push_back(off + s.hb_mp[6 * t + k]);
      for (int k = 0; k < 6; ++k) hb_nat.push_back(s.hb_nat[6 * t + k]);
      for (int k = 0; k < 7; ++k) hb_coef.push_back(s.hb_coef[7 * t + k]);
      for (int k = 0; k < 6; ++k) {
        int st = s.hb_site[6 * t + k];
        hb_site.push_back(st >= 0 ? nsite + st : -1);
        if (st >= 0) site_use[st] += 1;
      }
      hb_tert.push_back(s.hb_tertiary ? s.hb_tertiary[t] : 0);
      hb_sys.push_back(is);
    }
    d.hb_excess_possible = 0;
    for (int k = 0; k < s.nhbsite; ++k)
      if (site_use[k] > s.hbsite_valence[k]) d.hb_excess_possible = 1;
    // incidence lists: bonds, angles, stacks, tertiary stacks, hbonds
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
    for (int i = 0; i < s.nmp; ++i) {
      inc_off.push_back((int32_t)inc.size());
      inc.insert(inc.end(), incb[i].begin(), incb[i].end());
    }
    nbd += s.nbd; nba += s.nba; nst += s.nst; ntst += s.ntst; nhb += s.nhb; nsite += s.nhbsite;
    // tiles
    h.bt_begin = (int32_t)btiles.size();
    for (int st0 = 0; st0 < s.nmp; st0 += kTile) btiles.push_back(Tile{is, st0});
    h.bt_count = (int32_t)btiles.size() - h.bt_begin;
    h.qt_begin = (int32_t)qtiles.size();
    for (int st0 = 0; st0 < nq; st0 += kTile) qtiles.push_back(Tile{is, st0});
    h.qt_count = (int32_t)qtiles.size() - h.qt_begin;
    d.bt0 = h.bt_begin; d.nbt = h.bt_count; d.qt0 = h.qt_begin; d.nqt = h.qt_count;
    // electrostatic pairs: one wave per chunk pair (I, J), I <= J
    d.pp_off = pp_off;
    h.pt_begin = (int32_t)ptask.size();
    if (s.p.ele_form != TIS_ELE_NONE && nq > 0) {
      pp_off += (int64_t)d.nqt * (d.nqt + 1) / 2 * 6 * kTile;
      for (int I = 0; I < d.nqt; ++I)
        for (int J = I; J < d.nqt; ++J) ptask.push_back(EwTask{is, I, J});
    }
    h.pt_count = (int32_t)ptask.size() - h.pt_begin;
    // tile pairs of the periodic list builds (k_lpair; periodic systems
    // only: their number and the count buffer grow as nbt^2)
    d.lp_off = lp_off;
    h.bp_begin = (int32_t)bpair.size();
    h.qp_begin = (int32_t)qpair.size();
    if (s.p.periodic) {
      lp_off += (int64_t)d.nbt * d.nbt;
      for (int a = 0; a < d.nbt; ++a)
        for (int b = 0; b < d.nbt; ++b) bpair.push_back(EwTask{is, a, b});
      for (int a = 0; a < d.nqt; ++a)
        for (int b = 0; b < d.nqt; ++b) qpair.push_back(EwTask{is, a, b});
    }
    h.bp_count = (int32_t)bpair.size() - h.bp_begin;
    h.qp_count = (int32_t)qpair.size() - h.qp_begin;
    h.has_task = (s.nhb > 0 || s.ntst > 0);
    if (h.has_task) { c->systask_of[is] = (int32_t)systask.size(); systask.push_back(is); }
    // Ewald reciprocal vectors and work items
    if (s.p.ele_form == TIS_ELE_EWALD && nq > 0) {
      d.ewk0 = (int32_t)ewc.size();
      ewald_kvectors(s.p, ewhx, ewhy, ewhz, ewc);
      d.nk = (int32_t)ewc.size() - d.ewk0;
      d.nqc = (nq + kTile - 1) / kTile;
      d.nkc = (d.nk + kEwKChunk - 1) / kEwKChunk;
      d.psk_off = psk_off; psk_off += 2 * (int64_t)d.nqc * d.nk;
      d.pf_off = pf_off; pf_off += 3 * (int64_t)d.nkc * nq;
      h.esk_begin = (int32_t)ewsk.size();
      for (int kt = 0; kt < (d.nk + kTile - 1) / kTile; ++kt)
        for (int qc = 0; qc < d.nqc; ++qc) ewsk.push_back(EwTask{is, kt, qc});
      h.esk_count = (int32_t)ewsk.size() - h.esk_begin;
      h.ef_begin = (int32_t)ewf.size();
      for (int qt = 0; qt < d.nqc; ++qt)
        for (int kc = 0; kc < d.nkc; ++kc) ewf.push_back(EwTask{is, qt, kc});
      h.ef_count = (int32_t)ewf.size() - h.ef_begin;
    }
    off += s.nmp;
  }
  inc_off.push_back((int32_t)inc.size());
  if (off >= (1 << 27)) return fail(c, TIS_ERR_ARG, "too many beads (%d)", off);
  if (specs.size() > 255) return fail(c, TIS_ERR_ARG, "too many bead species (%zu)", specs.size());
  // incidence index of every (term, slot): the term passes write there
  std::vector<int32_t> bd_pos(2 * (size_t)nbd), ba_pos(3 * (size_t)nba), st_pos(7 * (size_t)nst),
      tst_pos(6 * (size_t)ntst), hb_pos(6 * (size_t)nhb);
  for (size_t k = 0; k < inc.size(); ++k) {
    const uint32_t e = inc[k], type = e >> 29, slot = (e >> 26) & 7u, term = e & ((1u << 26) - 1u);
    switch (type) {
      case INC_BOND: bd_pos[(size_t)slot * nbd + term] = (int32_t)k; break;
      case INC_ANGLE: ba_pos[(size_t)slot * nba + term] = (int32_t)k; break;
      case INC_STACK: st_pos[(size_t)slot * nst + term] = (int32_t)k; break;
      case INC_TSTACK: tst_pos[(size_t)slot * ntst + term] = (int32_t)k; break;
      default: hb_pos[(size_t)slot * nhb + term] = (int32_t)k; break;
    }
  }
  // species-pair tables (force_exv_dt15.F90:60-80): D_ij and eps_i eps_j,
  // one per distinct (exv_dist, exv_dist_PS) of the batch
  const int NS = (int)specs.size();
  std::vector<SpPair> sptab;
  std::vector<std::pair<double, double>> tabkeys;
  for (int is = 0; is < nsys; ++is) {
    DSys &d = c->hs[is].d;
    int found = -1;
    for (size_t k = 0; k < tabkeys.size(); ++k)
      if (tabkeys[k].first == d.exv_dist && tabkeys[k].second == d.exv_dist_PS) found = (int)k;
    if (found < 0) {
      found = (int)tabkeys.size();
      tabkeys.push_back({d.exv_dist, d.exv_dist_PS});
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
          sptab.push_back(SpPair{dij, dij * dij, A.e * B.e});
        }
    }
    d.sptab = found * NS * NS;
  }
  c->ntot = off; c->nqtot = qoff; c->nhbtot = nhb; c->ntsttot = ntst; c->nsttot = nst;
  c->nbtile = (int32_t)btiles.size(); c->nqtile = (int32_t)qtiles.size(); c->nsystask = (int32_t)systask.size();

  // device buffers
  DevPtrs &P = c->P;
  P = DevPtrs{};
  P.nsys = nsys; P.ntot = off; P.nqtot = qoff;
  int rc = 0;
  std::vector<DSys> dsys(nsys);
  for (int is = 0; is < nsys; ++is) dsys[is] = c->hs[is].d;
#define UP(ptr, vec) do { rc = dupload(c, &ptr, vec); if (rc) return rc; } while (0)
  UP(c->d_sys, dsys);
  P.sys = c->d_sys;
  for (int k = 0; k < 2; ++k) {
    rc = dalloc(c, &P.x[k], 3 * (size_t)off); if (rc) return rc;
    rc = dalloc(c, &P.px[k], c->any_periodic ? 3 * (size_t)off : 1); if (rc) return rc;
  }
  rc = dalloc(c, &P.v, 3 * (size_t)off); if (rc) return rc;
  rc = dalloc(c, &P.a, 3 * (size_t)off); if (rc) return rc;
  rc = dalloc(c, &P.f, 3 * (size_t)off); if (rc) return rc;
  uint8_t *d_lcl;
  UP(d_lcl, lcl); P.lcl = d_lcl;
  UP(P.lcls, lcls);
  double *d_mass, *d_rad, *d_charge, *d_rl2;
  int8_t *d_mpt, *d_icl;
  int32_t *d_unit, *d_qidx, *d_qbead, *d_incoff;
  uint32_t *d_inc;
  float *d_eps;
  UP(d_mass, mass); UP(d_rad, exv_rad); UP(d_charge, charge); UP(d_rl2, rl2);
  UP(d_mpt, mptype); UP(d_icl, iclass); UP(d_unit, unit); UP(d_qidx, qidx); UP(d_qbead, qbead);
  UP(d_incoff, inc_off); UP(d_inc, inc); UP(d_eps, exv_eps);
  int32_t *d_pos;
  UP(d_pos, bd_pos); P.bd_pos = d_pos; UP(d_pos, ba_pos); P.ba_pos = d_pos; UP(d_pos, st_pos); P.st_pos = d_pos;
  UP(d_pos, tst_pos); P.tst_pos = d_pos; UP(d_pos, hb_pos); P.hb_pos = d_pos;
  P.ninc = (int32_t)inc.size();
  rc = dalloc(c, &P.fslot, 3 * std::max<size_t>(inc.size(), 1)); if (rc) return rc;
  uint8_t *d_species;
  SpPair *d_sptab;
  UP(d_species, species); UP(d_sptab, sptab);
  P.species = d_species; P.sptab = d_sptab; P.nspecies = NS;
  rc = dalloc(c, &P.bbox, (size_t)c->nbtile); if (rc) return rc;
  rc = dalloc(c, &P.qbbox, (size_t)c->nqtile); if (rc) return rc;
  rc = dalloc(c, &P.sbox, 4 * (size_t)c->nbtile); if (rc) return rc;
  rc = dalloc(c, &P.lcnt, (size_t)std::max<int64_t>(lp_off, 1) * kTile); if (rc) return rc;
  P.mass = d_mass; P.exv_rad = d_rad; P.charge = d_charge; P.rl2 = d_rl2;
  P.mptype = d_mpt; P.iclass = d_icl; P.unit = d_unit; P.qidx = d_qidx; P.qbead = d_qbead;
  P.inc_off = d_incoff; P.inc = d_inc; P.exv_eps = d_eps;
  int32_t *i32; double *f64;
  UP(i32, bd_mp); P.bd_mp = i32; UP(f64, bd_nat); P.bd_nat = f64; UP(f64, bd_coef); P.bd_coef = f64; P.nbd = nbd;
  UP(i32, ba_mp); P.ba_mp = i32; UP(f64, ba_nat); P.ba_nat = f64; UP(f64, ba_coef); P.ba_coef = f64; P.nba = nba;
  UP(i32, st_mp); P.st_mp = i32; UP(f64, st_nat); P.st_nat = f64; UP(f64, st_coef); P.st_coef = f64; P.nst = nst;
  UP(i32, tst_mp); P.tst_mp = i32; UP(f64, tst_nat); P.tst_nat = f64; UP(f64, tst_coef); P.tst_coef = f64;
  UP(i32, tst2st); P.tst2st = i32; UP(i32, tst_excl); P.tst_excl = i32; P.ntst = ntst;
  UP(i32, hb_mp); P.hb_mp = i32; UP(f64, hb_nat); P.hb_nat = f64; UP(f64, hb_coef); P.hb_coef = f64;
  UP(i32, hb_site); P.hb_site = i32; UP(i32, hb_tert); P.hb_tert = i32; P.nhb = nhb;
  UP(i32, hb_sys); P.hb_sys = i32; UP(i32, tst_sys); P.tst_sys = i32;
  UP(i32, site_val); P.site_val = i32;
  {
    // per consecutive stack, the exclusive tertiary stacks that name it
    std::vector<int32_t> cnt(nst + 1, 0), exoff(nst + 1, 0), ex;
    for (int t = 0; t < ntst; ++t)
      for (int k = 0; k < 2; ++k)
        if (tst_excl[2 * t + k] && tst2st[2 * t + k] >= 0) ++cnt[tst2st[2 * t + k]];
    for (int st = 0; st < nst; ++st) exoff[st + 1] = exoff[st] + cnt[st];
    ex.assign(std::max(exoff[nst], 1), 0);
    std::vector<int32_t> fill(exoff.begin(), exoff.end() - 1);
    for (int t = 0; t < ntst; ++t)
      for (int k = 0; k < 2; ++k)
        if (tst_excl[2 * t + k] && tst2st[2 * t + k] >= 0) ex[fill[tst2st[2 * t + k]]++] = t;
    UP(i32, exoff); P.st_exoff = i32; UP(i32, ex); P.st_ex = i32;
  }
  rc = dalloc(c, &P.hb_inlist, (size_t)nhb); if (rc) return rc;
  rc = dalloc(c, &P.hb_status, (size_t)nhb); if (rc) return rc;
  rc = dalloc(c, &P.hb_ene, (size_t)nhb); if (rc) return rc;
  rc = dalloc(c, &P.for_hb, 18 * (size_t)nhb); if (rc) return rc;
  rc = dalloc(c, &P.site_cnt, (size_t)nsite); if (rc) return rc;
  rc = dalloc(c, &P.hb_seq, (size_t)nhb); if (rc) return rc;
  rc = dalloc(c, &c->d_ststat, (size_t)nst); if (rc) return rc;
  rc = dalloc(c, &P.nb_cnt, (size_t)off); if (rc) return rc;
  P.nb_cap = 64;
  rc = dalloc(c, &P.nb, (size_t)P.nb_cap * off); if (rc) return rc;
  rc = dalloc(c, &P.dxyz, 3 * (size_t)off); if (rc) return rc;
  rc = dalloc(c, &P.dlast, 3 * (size_t)off); if (rc) return rc;
  rc = dalloc(c, &P.ttop, 2 * btiles.size()); if (rc) return rc;
  rc = dalloc(c, &P.sys_flag, (size_t)nsys); if (rc) return rc;
  rc = dalloc(c, &P.pp, (size_t)std::max<int64_t>(pp_off, 1)); if (rc) return rc;
  rc = dalloc(c, &P.qmask, (size_t)moff); if (rc) return rc;
  rc = dalloc(c, &P.qmirror, (size_t)std::max<int64_t>(mioff, 1)); if (rc) return rc;
  rc = dalloc(c, &c->d_counters, TIS_CNT_MAX); if (rc) return rc;
  rc = dalloc(c, &c->d_stepbase, 1); if (rc) return rc;
  P.counters = c->d_counters;
  P.stepbase = c->d_stepbase;
  P.seed = seed;
  UP(c->d_btiles, btiles);
  UP(c->d_qtiles, qtiles);
  {
    std::vector<double> h3;
    h3.insert(h3.end(), ewhx.begin(), ewhx.end());
    h3.insert(h3.end(), ewhy.begin(), ewhy.end());
    h3.insert(h3.end(), ewhz.begin(), ewhz.end());
    double *dh = nullptr, *dc = nullptr;
    UP(dh, h3); UP(dc, ewc);
    P.ewh = dh; P.ewc = dc; P.nktot = (int32_t)ewc.size();
    rc = dalloc(c, &P.psk, (size_t)std::max<int64_t>(psk_off, 1)); if (rc) return rc;
    rc = dalloc(c, &P.pf, (size_t)std::max<int64_t>(pf_off, 1)); if (rc) return rc;
    UP(c->d_ewsk, ewsk);
    UP(c->d_ewf, ewf);
    UP(c->d_ptask, ptask);
    c->n_ptask = (int32_t)ptask.size();
    // whole-batch k_force launches take the off-diagonal chunk pairs (64
    // steps) before the diagonal ones (32): the order changes no result
    std::vector<EwTask> lpt(ptask);
    std::stable_partition(lpt.begin(), lpt.end(), [](const EwTask &t) { return t.a != t.b; });
    UP(c->d_ptask_lpt, lpt);
    UP(c->d_bpair, bpair);
    c->n_bpair = (int32_t)bpair.size();
    UP(c->d_qpair, qpair);
    c->n_qpair = (int32_t)qpair.size();
    c->n_ewsk = (int32_t)ewsk.size();
    c->n_ewf = (int32_t)ewf.size();
  }
  UP(c->d_systask, systask);
  std::vector<int32_t> sysidx(nsys);
  for (int is = 0; is < nsys; ++is) sysidx[is] = is;
  UP(c->d_sysidx, sysidx);
  rc = dalloc(c, &c->d_energy, 25 * (size_t)nsys); if (rc) return rc;
#undef UP
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cur = 0;
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
    HIPCHK(c, hipMemcpyAsync(base + (size_t)d * c->ntot + h.d.off, comp.data(), comp.size() * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    c->h2d_bytes += (int64_t)(comp.size() * sizeof(double));
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
    HIPCHK(c, hipMemcpyAsync(comp.data(), base + (size_t)d * c->ntot + h.d.off, comp.size() * sizeof(double),
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
    int rc = dalloc(c, &nb, (size_t)newcap * c->ntot);
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
      HIPCHK(c, hipMemcpy(nb.data() + (size_t)m * d.nmp, c->P.nb + (size_t)m * c->ntot + d.off,
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
  // k_force: per-system tertiary stacks and hbonds, the electrostatic chunk
  // pairs, then the bonded terms in blocks of 256 lanes
  const int nterm = r.nb + r.na + r.ns;
  const int npt = c->any_ele ? r.ptn : 0;
  const int nblk = (r.taskn + kWaves - 1) / kWaves + (npt + kWaves - 1) / kWaves +
                   (nterm + kTile * kWaves - 1) / (kTile * kWaves);
  if (nblk > 0) {
    const BondedRange br{r.b0, r.nb, r.a0, r.na, r.s0, r.ns};
    const int32_t *task = c->d_systask + r.task0;
    const EwTask *pt = r.all ? c->d_ptask_lpt : c->d_ptask + r.pt0;
    if (c->timing)
      hipExtLaunchKernelGGL(k_force, dim3(nblk), dim3(kTile * kWaves), 0, c->stream, e3[1], e3[2], 0, c->P, X, PX,
                            task, r.taskn, pt, npt, br, pre_t);
    else
      hipLaunchKernelGGL(k_force, dim3(nblk), dim3(kTile * kWaves), 0, c->stream, c->P, X, PX, task, r.taskn, pt,
                         npt, br, pre_t);
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
                          PXN, c->d_btiles + r.bt0, r.btn, (int32_t)t_local, mode, ele_on, count_step);
  else if (r.btn > 0)
    hipLaunchKernelGGL(k_main, dim3(r.btn), dim3(kTile * 2), 0, c->stream, c->P, X, PX, XN, PXN,
                       c->d_btiles + r.bt0, r.btn, (int32_t)t_local, mode, ele_on, count_step);
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
  const int nsel = isys < 0 ? c->nsys : 1;
  hipLaunchKernelGGL(k_energy, dim3(nsel), dim3(256), 0, c->stream, c->P, Xc(c), PXc(c),
                     c->d_sysidx + (isys < 0 ? 0 : isys), c->d_energy, 0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(energy_out, c->d_energy, 25 * sizeof(double) * (size_t)nsel, hipMemcpyDeviceToHost,
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
                     c->d_sysidx + isys, c->d_eunit, 1);
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
    if (isys[k] < 0 || isys[k] >= c->nsys) return fail(c, TIS_ERR_ARG, "system %d out of range", isys[k]);
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

int32_t tis_transfer_bytes(tis_ctx *c, int64_t *h2d, int64_t *d2h) {
  if (!c) return TIS_ERR_ARG;
  if (h2d) *h2d = c->h2d_bytes;
  if (d2h) *d2h = c->d2h_bytes;
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
  std::vector<LangClass> lc;
  for (int is = 0; is < c->nsys; ++is) {
    const HostSys &h = c->hs[is];
    for (int k = 0; k < h.d.nlc; ++k) {
      lc.push_back(LangClass{{0.0, 0.0, 0.0, 0.0}, 1.0 / h.lc_mass[k], h.lc_fric[k], h.lc_mass[k], 0.0});
      langevin_consts(tstep, h.d.tempk, h.lc_fric[k], h.lc_mass[k], lc.back().r);
    }
  }
  if (!lc.empty())
    HIPCHK(c, hipMemcpyAsync((void *)c->P.lcls, lc.data(), lc.size() * sizeof(LangClass), hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_stepbase, &step0, sizeof(step0), hipMemcpyHostToDevice, c->stream));
  c->step_host = step0;
  rc = rebuild_lists(c, -1);
  if (rc) return rc;
  // displacements since this rebuild start at zero, no rebuild pending
  HIPCHK(c, hipMemsetAsync(c->P.dxyz, 0, 3 * sizeof(double) * (size_t)std::max(c->ntot, 1), c->stream));
  HIPCHK(c, hipMemsetAsync(c->P.sys_flag, 0, (size_t)std::max(c->nsys, 1), c->stream));
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
    for (; done < nsteps; ++done) {
      for (int is = 0; is < c->nsys; ++is) eager_lists(c, is);
      for (int is = 0; is < c->nsys; ++is) eager_forces(c, is, is == c->nsys - 1);
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, c->stream, c->d_stepbase, (uint64_t)1, c->d_counters);
      c->cur ^= 1; ++c->t_since_init; *nlaunch += 2 * c->nsys;
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
  int rc = dalloc(c, &nb, (size_t)newcap * c->ntot);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(nb, c->P.nb, sizeof(uint32_t) * (size_t)c->P.nb_cap * c->ntot, hipMemcpyDeviceToDevice,
                           c->stream));
  c->P.nb = nb;  // the old buffer stays in allocs until the batch is freed
  c->P.nb_cap = newcap;
  for (int k = 0; k < 2; ++k)
    if (c->graph[k]) { (void)hipGraphExecDestroy(c->graph[k]); c->graph[k] = nullptr; }
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
      std::vector<int32_t> cnt((size_t)c->ntot);
      HIPCHK(c, hipMemcpy(cnt.data(), c->P.nb_cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost));
      for (int32_t v : cnt) need = std::max(need, v);
    }
    if (cap < need)
      return fail(c, TIS_ERR_ARG, "capacity %d below the %d entries of the current lists", cap, need);
    uint32_t *nb = nullptr;
    rc = dalloc(c, &nb, (size_t)cap * c->ntot);
    if (rc) return rc;
    const int keep = std::min(cap, c->P.nb_cap);
    HIPCHK(c, hipMemcpyAsync(nb, c->P.nb, sizeof(uint32_t) * (size_t)keep * c->ntot, hipMemcpyDeviceToDevice,
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
