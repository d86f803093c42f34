// Host side of the C ABI declared in include/adcgpu.h: context set-up
// (host preprocessing of the reference's mesh tables into the device layout),
// state transfer, per-step launch sequence and the drop-in JCG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/adcgpu.h"
#include "jcg_host.h"
#include "kernels.h"
#include "mlp.h"

using namespace adcgpu;

// message of the last failed call on this thread: a fixed buffer, so that
// recording an error allocates nothing
thread_local char g_err[512];

void adcgpu::set_last_error(const char* msg) { std::snprintf(g_err, sizeof(g_err), "%s", msg); }

namespace {

// bound of every wait inside the persistent JCG kernel (adcgpu_set_wait_limit)
double g_wait_seconds = 0.5;

int fail(int code, const std::string& msg) {
  set_last_error(msg.c_str());
  return code;
}

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(ADCGPU_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dmalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(ADCGPU_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

template <class T>
int upload(T** p, const std::vector<T>& v) {
  int rc = dmalloc(p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

inline int nblk(int n, int b) { return (n + b - 1) / b; }

}  // namespace

struct adcgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int np = 0, ne = 0, mnei = 0, neta = 0, nch = 0;
  adcgpu_params prm{};
  double ifnlfa = 1, ifnlct = 1, ifnlcat = 1, iflinbf = 0, ifnlbf = 1, tkmax = 1;
  Hybrid hyb{0.0, 1.0, 10.0, 1.0 / 3.0};  // hybrid friction (NOLIBF = 2)
  double habsmin = 0, hoff = 0, errorelev = 1000.0;
  // mesh
  int *nm = nullptr, *inc_ptr = nullptr, *inc = nullptr, *slot = nullptr, *nneigh = nullptr;
  int *col = nullptr, *slot_cols = nullptr;
  int *csrpos = nullptr, *bnd_index = nullptr, *nbd = nullptr, *obc_q = nullptr, *obc_s = nullptr;
  int nobc = 0;
  int *obcg_ptr = nullptr, *obcg = nullptr, *vbc_ptr = nullptr, *vbc = nullptr, *lbcodei = nullptr;
  double *fdx = nullptr, *fdy = nullptr, *areas = nullptr, *x = nullptr, *y = nullptr, *dp = nullptr;
  double *tau0 = nullptr, *evm = nullptr, *corif = nullptr, *csii = nullptr, *siii = nullptr;
  // state (pointers rotate between time levels)
  double *eta1, *eta2, *etas, *uu1, *vv1, *uu2, *vv2, *qx1, *qy1, *qx2, *qy2, *h1, *h2, *tk;
  double *etamax, *totalarea, *wsx1, *wsy1, *wsx2, *wsy2, *pr1, *pr2, *gwce_lv, *alphal;
  double *coef, *obc, *fric, *etas_prev = nullptr, *umax = nullptr;
  int *nodecode, *nnodecode, *noff, *noffold, *mju, *etamax_it, *umax_it;
  int cur_it = 0;  // step opened by adcgpu_begin_step
  // work
  double *contrib = nullptr, *cx = nullptr, *cy = nullptr, *ep_part = nullptr, *scal = nullptr;
  double *eta_bc = nullptr, *eta_bc_table = nullptr;
  int eta_bc_table_steps = 0;
  WdWork wd{};
  int* flags = nullptr;
  // JCG work
  double *cs = nullptr, *jd = nullptr, *jw = nullptr;
  double *jpart = nullptr, *jsaved = nullptr;
  unsigned* jbar = nullptr;
  int jcg_grid = 0;
  int clock_khz = 0;            // device wall-clock rate: the JCG wait bound in ticks
  int jcg_cap = 0;              // resident k_jcg blocks on this device
  int jcg_halo = -1;            // row tiles each side a tile's rows reference (-1: any)
  int jcg_hrows = -1;           // rows each side staged in LDS with a tile's (-1: none)
  int* jhrange = nullptr;       // per row tile: first and one past the last row it references
  unsigned* jlcol = nullptr;    // staged: the columns as 16-bit LDS positions, two slots a word
  int jcg_limit = 0;            // caller's limit on JCG blocks per launch (0: the device's)
  bool jcg_mem = false;         // more row tiles than blocks: k_jcg_mem
  double *juo = nullptr, *jdpv = nullptr, *jndv = nullptr;  // k_jcg_mem's iterates
  int jmem_n = 0;               // rows those hold
  // tidal potential (NTIP = 1): constituents, ramp and time origin, the
  // species factors L_N [3][np] and longitudes on the device, TIP1/TIP2
  bool ctip = false;
  std::vector<double> tip_tpk, tip_amigt, tip_etrf, tip_fft, tip_facet, tip_pert;
  int tip_nramp = 0;
  double tip_dramp = 0.0, tip_statim = 0.0, tip_reftim = 0.0;
  double *tip1 = nullptr, *tip2 = nullptr, *tip_ln = nullptr, *tip_slam = nullptr;
  // specified normal flux boundaries (NFLUXF = 1): the flow-boundary tables
  // kept on the host, each flux node's edge terms, QN0/QN1/QN2 per entry and
  // the resident run's QN2 rows
  int nvel = 0;
  std::vector<int> h_nbv, h_bnd_index;
  bool nfluxf = false;
  int *fb_ptr = nullptr, *fb_ent = nullptr, *nbv0 = nullptr;
  double *bndlen = nullptr, *qn0 = nullptr, *qn1 = nullptr, *qn2 = nullptr, *qrad = nullptr;
  double* elevdisc = nullptr;  // river radiation's reference elevation per entry
  bool has_rad = false;   // radiating entries (LBCODEI 30, 52)
  bool has_spec = false;  // specified-flux entries (LBCODEI 2, 12, 22)
  double* qn_table = nullptr;
  int qn_table_steps = 0, qn_table_cap = 0;
  int64_t run_iters = 0, run_rebuilds = 0, run_itmax = 0, run_errstep = 0, run_mlused = 0;
  bool coef_scaled = false;  // off-diagonals of COEF left scaled by a deferred solve
  // the GWCE matrix's unchanging terms per incidence slot, and each element's
  // NCEle as the next rebuild sees it (written by the momentum element pass;
  // recomputed when the flags were changed otherwise)
  double *lhs_td = nullptr, *lhs_ta = nullptr, *lhs_tb = nullptr;
  unsigned char* ncele = nullptr;
  bool ncele_fresh = false;
  // mixed precision (ADCGPU_MIXED): FP32 element geometry and cached matrix
  // terms; contrib, cx and cy then hold floats
  bool mixed = false;
  float *fdx32 = nullptr, *fdy32 = nullptr, *areas32 = nullptr;
  float *lhs_td32 = nullptr, *lhs_ta32 = nullptr, *lhs_tb32 = nullptr;
  // the cached terms per off-diagonal entry (k_lhs_pairs; ILump = 0)
  bool lhs_pairs = false;
  double *lp_t1 = nullptr, *lp_t2 = nullptr;
  float *lp_t1f = nullptr, *lp_t2f = nullptr;
  int *lp_e1 = nullptr, *lp_e2 = nullptr;
  // learned JCG start (set up on first use): FP32 master parameters and Adam
  // moments, the kernels' weight layouts, the per-step correction and second
  // start, training samples and the training work arrays
  bool ml_ready = false;
  float *ml_master = nullptr, *ml_m1 = nullptr, *ml_m2 = nullptr;
  MlWeights mlw{};
  double *ml_corr = nullptr, *ml_u = nullptr, *jd2 = nullptr;
  bf16* ml_fx = nullptr;      // the network's inputs, operand layout (k_ml_feat)
  double* ml_sv = nullptr;    // the nodes' input scales
  MlSamples smp{};
  int64_t smp_total = 0;  // samples collected so far (the ring keeps the last cap)
  int ml_collect = 0;     // samples per step while collecting (0: off)
  int ml_step = 0;        // Adam steps taken
  bf16 *tx = nullptr, *txt = nullptr, *ta1 = nullptr, *ta1t = nullptr, *tdz2 = nullptr,
       *tdz2t = nullptr, *tdz1t = nullptr;
  float *tz1 = nullptr, *tz2 = nullptr, *tda1 = nullptr, *tgw1 = nullptr, *tgw2 = nullptr,
        *tpo = nullptr, *tpz = nullptr, *tloss = nullptr, *ty = nullptr, *tw = nullptr,
        *twt = nullptr, *tmeanw = nullptr;
  int* tperm = nullptr;
  // optional timing of the JCG launches with HIP events on the context stream
  int profile = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> jcg_events;  // pool, created ahead
  size_t ev_used = 0;
  std::vector<void*> allocs;

  DevMesh mesh() const {
    DevMesh m;
    m.np = np; m.ne = ne; m.mnei = mnei; m.neta = neta;
    m.dt = prm.dtdp; m.g = prm.g; m.a00 = prm.a00; m.b00 = prm.b00; m.c00 = prm.c00;
    m.nm = nm; m.fdx = fdx; m.fdy = fdy; m.areas = areas; m.x = x; m.y = y; m.dp = dp;
    m.tau0 = tau0; m.evm = evm; m.corif = corif; m.inc_ptr = inc_ptr; m.inc = inc;
    m.slot = slot; m.slot_cols = slot_cols;
    m.nneigh = nneigh; m.col = col; m.csrpos = csrpos; m.bnd_index = bnd_index; m.nbd = nbd;
    m.obc_q = obc_q; m.obc_s = obc_s; m.nobc = nobc; m.obcg_ptr = obcg_ptr; m.obcg = obcg; m.vbc_ptr = vbc_ptr;
    m.vbc = vbc; m.lbcodei = lbcodei; m.csii = csii; m.siii = siii;
    m.fb_ptr = nfluxf ? fb_ptr : nullptr; m.fb_ent = nfluxf ? fb_ent : nullptr;
    m.nbv = nfluxf ? nbv0 : nullptr; m.bndlen = nfluxf ? bndlen : nullptr;
    m.fdx32 = fdx32; m.fdy32 = fdy32; m.areas32 = areas32;
    return m;
  }
  DevState state() const {
    DevState s;
    s.eta1 = eta1; s.eta2 = eta2; s.etas = etas; s.uu1 = uu1; s.vv1 = vv1; s.uu2 = uu2;
    s.vv2 = vv2; s.qx1 = qx1; s.qy1 = qy1; s.qx2 = qx2; s.qy2 = qy2; s.h1 = h1; s.h2 = h2;
    s.tk = tk; s.etamax = etamax; s.totalarea = totalarea; s.wsx1 = wsx1; s.wsy1 = wsy1;
    s.wsx2 = wsx2; s.wsy2 = wsy2; s.pr1 = pr1; s.pr2 = pr2; s.gwce_lv = gwce_lv;
    s.alphal = alphal; s.coef = coef; s.obc = obc; s.fric = fric; s.nodecode = nodecode;
    s.nnodecode = nnodecode; s.noff = noff; s.noffold = noffold; s.mju = mju;
    s.etamax_it = etamax_it; s.umax = umax; s.umax_it = umax_it; s.ncele = ncele;
    s.tip1 = ctip ? tip1 : nullptr;
    s.tip2 = ctip ? tip2 : nullptr;
    s.qn0 = nfluxf ? qn0 : nullptr;
    s.qn1 = nfluxf ? qn1 : nullptr;
    s.qn2 = nfluxf ? qn2 : nullptr;
    s.qrad = nfluxf ? qrad : nullptr;
    s.elevdisc = nfluxf ? elevdisc : nullptr;
    return s;
  }
};

namespace {

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ADCGPU_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

// The sticky error conditions of the device flag words, as a status code.
int flag_errors(const int* f) {
  const int e = f[F_ERROR];
  const std::string at = " (first at step " + std::to_string(f[F_ERRSTEP]) + ")";
  if (e & ERR_BARRIER) return fail(ADCGPU_E_HIP, "JCG wait timed out");
  if (e & ERR_DIAG) return fail(ADCGPU_E_DIAG, "GWCE diagonal <= 0");
  if (e & ERR_NONFINITE) return fail(ADCGPU_E_NONFINITE, "non-finite velocity" + at);
  if (e & ERR_ELEV) return fail(ADCGPU_E_ERRORELEV, "wet-node elevation above ErrorElev" + at);
  return 0;
}

// JCG_CPB chunks of 2048 rows per block plus the eigenvalue block, all
// resident; a system with more row tiles than the blocks allowed (the
// resident ones, or the caller's limit) runs memory-resident (k_jcg_mem) on
// all of them.
int set_jcg_grid(adcgpu_ctx* c) {
  const int need = jcg_grid_for(c->np);
  const int lim = c->jcg_limit > 0 ? std::min(c->jcg_limit, c->jcg_cap) : c->jcg_cap;
  if (need <= lim) {
    c->jcg_grid = need;
    c->jcg_mem = false;
    return 0;
  }
  if (lim < 2)
    return fail(ADCGPU_E_CAPACITY, "JCG needs at least two resident blocks");
  if (c->prm.warm_start == 2)
    return fail(ADCGPU_E_UNSUPPORTED,
                "the learned start needs a system the register-resident JCG holds (" +
                    std::to_string((lim - 1) * JCG_CPB * RED_CHUNK) + " rows in " +
                    std::to_string(lim) + " blocks)");
  if (c->jmem_n < c->np) {
    for (double** p : {&c->juo, &c->jdpv, &c->jndv}) {
      if (*p) {
        c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)*p));
        (void)hipFree(*p);
        *p = nullptr;
      }
      if (int rc = dmalloc(p, (size_t)c->np)) return rc;
      c->allocs.push_back((void*)*p);
    }
    c->jmem_n = c->np;
  }
  c->jcg_grid = lim;
  c->jcg_mem = true;
  return 0;
}

int setup_jcg_grid(adcgpu_ctx* c) {
  HIPCHK(jcg_resident_cap(c->device, &c->jcg_cap));
  int khz = 0;
  HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
  c->clock_khz = std::max(khz, 1000);
  return set_jcg_grid(c);
}

// Event pairs for timing JCG launches, created outside the timed region.
int grow_events(adcgpu_ctx* c, size_t n) {
  while (c->jcg_events.size() < n) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return fail(ADCGPU_E_HIP, "hipEventCreate");
    }
    c->jcg_events.emplace_back(e0, e1);
  }
  return 0;
}

int jcg_barrier_words(const adcgpu_ctx* c) { return jcg_barrier_words_for(c->jcg_grid); }

JcgArgs make_jcg_args(adcgpu_ctx* c, double* coef, const int* col, const int* nneigh,
                      double* rhs, double* u, int n, int itmax, double zeta, double* eta2,
                      const double* eta1, const int* nodecode, bool prescaled,
                      bool defer_unscal, const double* u2) {
  JcgArgs a;
  std::memset(&a, 0, sizeof(a));  // padding too: batches compare whole structs
  a.n = n;
  jcg_tolerances(a, itmax, zeta);
  a.coef = coef;
  a.cs = c->cs;
  a.col = col;
  a.nneigh = nneigh;
  a.rhs = rhs;
  a.u = u;
  a.d = c->jd;
  a.w = c->jw;
  a.part = c->jpart;
  a.saved = c->jsaved;
  a.prescaled = prescaled;
  a.defer_unscal = defer_unscal;
  a.barrier = c->jbar;
  a.halo = c->jcg_halo;
  a.hrows = c->jcg_hrows;
  a.hrange = c->jhrange;
  a.lcol = c->jlcol;
  a.uo = c->juo;
  a.dpv = c->jdpv;
  a.ndv = c->jndv;
  if (c->jcg_mem) {  // gathered products, grid barriers
    a.hrows = -1;
    a.halo = -1;
  }
  a.barrier_ticks = (long long)((double)c->clock_khz * 1000.0 * g_wait_seconds);
  a.flags = c->flags;
  a.eta2 = eta2;
  a.eta1 = eta1;
  a.nodecode = nodecode;
  // the learned second start, formed by the pass before (learned mode)
  a.u2 = u2;
  a.d2 = c->jd2;
  return a;
}

int launch_jcg(adcgpu_ctx* c, double* coef, const int* col, const int* nneigh, double* rhs,
               double* u, int n, int itmax, double zeta, double* eta2, const double* eta1,
               const int* nodecode, bool prescaled, bool defer_unscal,
               const double* u2 = nullptr) {
  JcgArgs a = make_jcg_args(c, coef, col, nneigh, rhs, u, n, itmax, zeta, eta2, eta1, nodecode,
                            prescaled, defer_unscal, u2);
  // a prescaling pass (k_rhs_node) has already cleared the barrier words
  if (!prescaled) {
    const int bar_words = jcg_barrier_words(c);
    hipLaunchKernelGGL(k_fill_u32, dim3(nblk(bar_words, 256)), dim3(256), 0, c->stream, c->jbar,
                       0u, bar_words);
  }
  void* args[] = {&a};
  if (c->profile) {
    if (c->ev_used == c->jcg_events.size())
      if (int rc = grow_events(c, c->jcg_events.size() + 256)) return rc;
    HIPCHK(hipEventRecord(c->jcg_events[c->ev_used].first, c->stream));
  }
  const void* fn =
      c->jcg_mem     ? (const void*)k_jcg_mem
      : a.hrows >= 0 ? (u2 ? (const void*)k_jcg<true, true> : (const void*)k_jcg<false, true>)
                     : (u2 ? (const void*)k_jcg<true, false> : (const void*)k_jcg<false, false>);
  HIPCHK(hipLaunchCooperativeKernel(fn, dim3(c->jcg_grid), dim3(JCG_BLOCK), args,
                                    (unsigned)jcg_stage_bytes(a.hrows), c->stream));
  if (c->profile) HIPCHK(hipEventRecord(c->jcg_events[c->ev_used++].second, c->stream));
  return 0;
}

// ---------------------------------------------------------------- learned start
constexpr int ML_CAP = 1 << 20;  // training samples kept (node snapshots)

// Host conversions to and from bfloat16 (round to nearest even).
uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_to_f32(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Initial parameters: W1, W2 uniform in +-sqrt(6/(fan_in+fan_out)) (Glorot &
// Bengio 2010) from a fixed-seed generator, biases zero, and the output layer
// zero, so the untrained network's correction is zero: the reference's start.
std::vector<float> ml_initial_params() {
  std::vector<float> p(ML_NPARAM, 0.0f);
  uint64_t st = 0x9e3779b97f4a7c15ull;
  auto uni = [&]() {  // splitmix64 -> [0, 1)
    st += 0x9e3779b97f4a7c15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const double a1 = std::sqrt(6.0 / (ML_F + ML_H)), a2 = std::sqrt(6.0 / (ML_H + ML_H));
  for (int k = 0; k < ML_W1; ++k) p[ML_OFF_W1 + k] = (float)((2.0 * uni() - 1.0) * a1);
  for (int k = 0; k < ML_W2; ++k) p[ML_OFF_W2 + k] = (float)((2.0 * uni() - 1.0) * a2);
  return p;
}

int ml_setup(adcgpu_ctx* c) {
  if (c->ml_ready) return 0;
  HIPCHK(hipSetDevice(c->device));
  int rc = 0;
  auto A = [&](auto** p, size_t n) {
    if (!rc) rc = dmalloc(p, n);
    if (!rc) {
      c->allocs.push_back((void*)*p);
      (void)hipMemset(*p, 0, n * sizeof(**p));
    }
  };
  A(&c->ml_master, ML_NPARAM); A(&c->ml_m1, ML_NPARAM); A(&c->ml_m2, ML_NPARAM);
  A(&c->mlw.w1p, ML_W1); A(&c->mlw.w2p, ML_W2); A(&c->mlw.w1t, ML_W1); A(&c->mlw.w2t, ML_W2);
  A(&c->mlw.w2n, ML_W2); A(&c->mlw.b1p, 128); A(&c->mlw.b2p, 128); A(&c->mlw.w3p, 128);
  A(&c->mlw.b3, 1);
  A(&c->ml_corr, c->np); A(&c->ml_u, c->np); A(&c->jd2, c->np);
  A(&c->ml_fx, (size_t)((c->np + 31) / 32) * 3 * 64 * 8); A(&c->ml_sv, c->np);
  c->smp.cap = ML_CAP;
  A(&c->smp.x, (size_t)ML_CAP * ML_F); A(&c->smp.s, ML_CAP); A(&c->smp.e0, ML_CAP);
  A(&c->smp.node, ML_CAP); A(&c->smp.y, ML_CAP);
  const size_t B = ML_BATCH;
  A(&c->tx, B * ML_F); A(&c->txt, B * ML_F); A(&c->ta1, B * ML_H); A(&c->ta1t, B * ML_H);
  A(&c->tdz2, B * ML_H); A(&c->tdz2t, B * ML_H); A(&c->tdz1t, B * ML_H);
  A(&c->tz1, B * ML_H); A(&c->tz2, B * ML_H); A(&c->tda1, B * ML_H); A(&c->tgw1, ML_W1);
  A(&c->tgw2, ML_W2); A(&c->tpo, (size_t)ML_TRAIN_BLOCKS * ML_PART_OUT);
  A(&c->tpz, (size_t)ML_TRAIN_BLOCKS * ML_H); A(&c->tloss, ML_TRAIN_BLOCKS + 1); A(&c->ty, B);
  A(&c->tw, B); A(&c->twt, ML_CAP); A(&c->tmeanw, 1); A(&c->tperm, B);
  if (rc) return rc;
  c->ml_ready = true;
  const std::vector<float> p = ml_initial_params();
  HIPCHK(hipMemcpy(c->ml_master, p.data(), p.size() * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_ml_pack, dim3(nblk(ML_W2 * 2, 256)), dim3(256), 0, c->stream,
                     (const float*)c->ml_master, c->mlw);
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_launch("learned start set-up");
}

// The step's feature inputs; `eta` is eta^k.
MlIn ml_in(const adcgpu_ctx* c, const double* eta) {
  MlIn in;
  in.np = c->np;
  in.nneigh = c->nneigh;
  in.col = c->col;
  in.bnd_index = c->bnd_index;
  in.eta = eta;
  in.etas = c->etas;
  in.etasp = c->etas_prev;
  in.dp = c->dp;
  in.wsx2 = c->wsx2;
  in.wsy2 = c->wsy2;
  in.pr1 = c->pr1;
  in.pr2 = c->pr2;
  in.tk = c->tk;
  in.nodecode = c->nodecode;
  in.dt = c->prm.dtdp;
  return in;
}

// The network's correction for every node into ml_corr: the inputs pass, then
// the network pass (grid-stride; its LDS and registers allow three blocks
// per CU, so three per CU of the 256 are launched).
void ml_infer(adcgpu_ctx* c, const MlIn& in) {
  const int ntile = (c->np + 31) / 32;
  bf16x8* fx = reinterpret_cast<bf16x8*>(c->ml_fx);
  hipLaunchKernelGGL(k_ml_feat, dim3(nblk(ntile, 4)), dim3(256), 0, c->stream, in, fx, c->ml_sv);
  hipLaunchKernelGGL(k_ml_mlp, dim3(std::min(nblk(ntile, 4), 3 * 256)), dim3(256), 0, c->stream,
                     in, c->mlw, (const bf16x8*)fx, (const double*)c->ml_sv, c->ml_corr);
}

// One minibatch through the network and back; `perm` (device) names its
// ML_BATCH sample slots (-1 pads).  With apply, one Adam step follows.
int ml_minibatch(adcgpu_ctx* c, float inv_n, float* gout, bool apply, float lr, float* lossacc) {
  const hipStream_t st = c->stream;
  const int B = ML_BATCH;
  hipLaunchKernelGGL(k_ml_batch, dim3(nblk(B * ML_F, 256)), dim3(256), 0, st, c->smp,
                     (const float*)c->twt, (const int*)c->tperm, c->tx, c->txt, c->ty, c->tw);
  // Z1 = X W1
  hipLaunchKernelGGL(k_ml_gemm_nt, dim3(B / 32, ML_H / 32), dim3(64), 0, st, B, ML_H, ML_F,
                     (const bf16*)c->tx, ML_F, (const bf16*)c->mlw.w1t, ML_F, c->tz1, ML_H);
  hipLaunchKernelGGL(k_ml_act1, dim3(nblk(B * ML_H, 256)), dim3(256), 0, st, c->tz1,
                     (const float*)c->ml_master, c->ta1, c->ta1t);
  // Z2 = A1 W2
  hipLaunchKernelGGL(k_ml_gemm_nt, dim3(B / 32, ML_H / 32), dim3(64), 0, st, B, ML_H, ML_H,
                     (const bf16*)c->ta1, ML_H, (const bf16*)c->mlw.w2t, ML_H, c->tz2, ML_H);
  hipLaunchKernelGGL(k_ml_out, dim3(ML_TRAIN_BLOCKS), dim3(ML_H), 0, st, (const float*)c->tz2,
                     (const float*)c->ml_master, (const float*)c->ty, (const float*)c->tw, inv_n,
                     c->tdz2, c->tdz2t, c->tpo, c->tloss);
  // dW2 = A1^T dZ2 ; dA1 = dZ2 W2^T
  hipLaunchKernelGGL(k_ml_gemm_nt, dim3(ML_H / 32, ML_H / 32), dim3(64), 0, st, ML_H, ML_H, B,
                     (const bf16*)c->ta1t, B, (const bf16*)c->tdz2t, B, c->tgw2, ML_H);
  hipLaunchKernelGGL(k_ml_gemm_nt, dim3(B / 32, ML_H / 32), dim3(64), 0, st, B, ML_H, ML_H,
                     (const bf16*)c->tdz2, ML_H, (const bf16*)c->mlw.w2n, ML_H, c->tda1, ML_H);
  hipLaunchKernelGGL(k_ml_dz1, dim3(ML_TRAIN_BLOCKS), dim3(ML_H), 0, st, (const float*)c->tda1,
                     (const float*)c->tz1, c->tdz1t, c->tpz);
  // dW1 = X^T dZ1 (48 rows)
  hipLaunchKernelGGL(k_ml_gemm_nt, dim3((ML_F + 31) / 32, ML_H / 32), dim3(64), 0, st, ML_F,
                     ML_H, B, (const bf16*)c->txt, B, (const bf16*)c->tdz1t, B, c->tgw1, ML_H);
  if (apply) ++c->ml_step;
  hipLaunchKernelGGL(k_ml_adam, dim3(nblk(ML_NPARAM, 256)), dim3(256), 0, st, c->ml_master,
                     c->ml_m1, c->ml_m2, (const float*)c->tgw1, (const float*)c->tgw2,
                     (const float*)c->tpo, (const float*)c->tpz, (const float*)c->tloss, lossacc,
                     gout, apply ? 1 : 0, c->ml_step, lr);
  if (apply)
    hipLaunchKernelGGL(k_ml_pack, dim3(nblk(ML_W2 * 2, 256)), dim3(256), 0, st,
                       (const float*)c->ml_master, c->mlw);
  return check_launch("training minibatch");
}

}  // namespace

extern "C" {

const char* adcgpu_last_error(void) { return g_err; }

int adcgpu_create(const adcgpu_mesh* M, const adcgpu_params* P, int device, adcgpu_ctx** out) {
  if (!M || !P || !out) return fail(ADCGPU_E_ARG, "null argument");
  *out = nullptr;
  const int np = M->np, ne = M->ne, mnei = M->mnei, neta = M->neta;
  if (np <= 0 || ne <= 0 || mnei < 2 || neta < 0 || M->nvel < 0 || M->nvelme < 0)
    return fail(ADCGPU_E_ARG, "invalid sizes");
  // the kernels index with 32-bit offsets: the widest arrays (MNEI doubles
  // per node, three slots per element) stay within them
  if ((size_t)mnei * (size_t)np * sizeof(double) >= ((size_t)1 << 32) ||
      3 * (size_t)ne >= ((size_t)1 << 31))
    return fail(ADCGPU_E_CAPACITY, "mesh too large for 32-bit offsets");
  if (mnei > LHS_MAXW)
    return fail(ADCGPU_E_UNSUPPORTED, "MNEI above " + std::to_string(LHS_MAXW));
  if (P->nolibf < 0 || P->nolibf > 3) return fail(ADCGPU_E_ARG, "NOLIBF");
  if (P->itmax <= 0 || P->itmax > JCG_MAX_ITMAX) return fail(ADCGPU_E_ARG, "ITMAX out of range");
  if (P->precision != ADCGPU_FP64 && P->precision != ADCGPU_MIXED)
    return fail(ADCGPU_E_ARG, "precision must be ADCGPU_FP64 or ADCGPU_MIXED");
  if (P->warm_start < 0 || P->warm_start > 2) return fail(ADCGPU_E_ARG, "warm_start must be 0, 1 or 2");
  if (P->warm_start == 2 && P->ilump != 0)
    return fail(ADCGPU_E_UNSUPPORTED, "the learned start needs the JCG solve (ILump = 0)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return fail(ADCGPU_E_HIP, "no HIP device " + std::to_string(device));
  HIPCHK(hipSetDevice(device));

  // ---- host preprocessing of the reference's tables
  std::vector<int> nm0(3 * (size_t)ne);
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) {
      int v = M->nm[e + (size_t)ne * k] - 1;
      if (v < 0 || v >= np) return fail(ADCGPU_E_ARG, "NM out of range");
      nm0[3 * (size_t)e + k] = v;
    }
  std::vector<int> cnt(np + 1, 0);
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) cnt[nm0[3 * (size_t)e + k] + 1]++;
  for (int i = 0; i < np; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> inc(3 * (size_t)ne), fillp(cnt.begin(), cnt.end() - 1);
  // ascending element order per node; slot is the inverse map, so element
  // kernels write each (element, vertex) contribution where its node's
  // ordered sum reads it
  std::vector<int> slot(3 * (size_t)ne);
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k) {
      const int p = fillp[nm0[3 * (size_t)e + k]]++;
      inc[p] = 3 * e + k;
      slot[3 * (size_t)e + k] = p;
    }
  std::vector<int> nneigh(M->nneigh, M->nneigh + np);
  std::vector<int> col((size_t)mnei * np, -1);
  int bw = 0;  // matrix bandwidth: the widest |column - row|
  for (int i = 0; i < np; ++i) {
    if (nneigh[i] < 1 || nneigh[i] > mnei) return fail(ADCGPU_E_ARG, "NNEIGH out of range");
    if (M->neitab[i] != i + 1) return fail(ADCGPU_E_ARG, "NEITAB(i,1) must be i");
    for (int j = 0; j < nneigh[i]; ++j) {
      int v = M->neitab[i + (size_t)np * j] - 1;
      if (v < 0 || v >= np) return fail(ADCGPU_E_ARG, "NEITAB out of range");
      col[(size_t)j * np + i] = v;
      bw = std::max(bw, std::abs(v - i));
    }
  }
  // a JCG block's rows reference rows at most this many blocks away
  const int halo_blocks = (bw + JCG_CPB * RED_CHUNK - 1) / (JCG_CPB * RED_CHUNK);
  // per row tile of a JCG block, the rows its rows reference: [lo, hi)
  const int ntile = (np + JCG_CPB * RED_CHUNK - 1) / (JCG_CPB * RED_CHUNK);
  std::vector<int> hrange(2 * (size_t)ntile);
  for (int tl = 0; tl < ntile; ++tl) {
    hrange[2 * tl] = tl * JCG_CPB * RED_CHUNK;
    hrange[2 * tl + 1] = std::min(np, (tl + 1) * JCG_CPB * RED_CHUNK);
  }
  for (int i = 0; i < np; ++i) {
    const int tl = i / (JCG_CPB * RED_CHUNK);
    for (int j = 1; j < nneigh[i]; ++j) {
      const int v = col[(size_t)j * np + i];
      hrange[2 * tl] = std::min(hrange[2 * tl], v);
      hrange[2 * tl + 1] = std::max(hrange[2 * tl + 1], v + 1);
    }
  }
  std::vector<int> csrpos(6 * (size_t)ne);
  for (int e = 0; e < ne; ++e)
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 2; ++l) {
        int J = M->csr_position[e + (size_t)ne * k + (size_t)3 * ne * l];
        int node = nm0[3 * (size_t)e + k];
        if (J < 2 || J > nneigh[node]) return fail(ADCGPU_E_ARG, "CSR_POSITION out of range");
        csrpos[(3 * (size_t)e + k) * 2 + l] = J - 1;
      }
  // per incidence slot, the two CSR_POSITION columns of its row
  std::vector<int> slot_cols(3 * (size_t)ne);
  for (size_t p = 0; p < 3 * (size_t)ne; ++p) {
    const int code = inc[p];
    slot_cols[p] = csrpos[(size_t)code * 2] | (csrpos[(size_t)code * 2 + 1] << 16);
  }
  std::vector<double> fdx(3 * (size_t)ne), fdy(3 * (size_t)ne);
  for (size_t q = 0; q < 3 * (size_t)ne; ++q) {
    fdx[q] = M->fdxe[q];
    fdy[q] = M->fdye[q];
  }
  std::vector<int> bnd_index(np, -1), nbd0(std::max(neta, 1), 0);
  for (int I = 0; I < neta; ++I) {
    int b = M->nbd[I] - 1;
    if (b < 0 || b >= np) return fail(ADCGPU_E_ARG, "NBD out of range");
    if (bnd_index[b] >= 0) return fail(ADCGPU_E_UNSUPPORTED, "repeated elevation-boundary node");
    bnd_index[b] = I;
    nbd0[I] = b;
  }
  std::vector<int> obc_slot((size_t)mnei * np, -1);
  std::vector<std::vector<int>> obcl(np);
  for (int I = 0; I < neta; ++I) {
    int b = nbd0[I];
    for (int J = 2; J <= nneigh[b]; ++J) {
      int nb = col[(size_t)(J - 1) * np + b];
      int found = 0;
      for (int IJ = 2; IJ <= nneigh[nb]; ++IJ)
        if (col[(size_t)(IJ - 1) * np + nb] == b) {
          obc_slot[(size_t)(IJ - 1) * np + nb] = I * mnei + (J - 2);
          ++found;
        }
      if (found != 1) return fail(ADCGPU_E_ARG, "asymmetric neighbour table at a boundary");
      obcl[nb].push_back(I * mnei + (J - 2));
    }
  }
  // the COEF positions that involve a boundary node, for the matrix's boundary pass
  std::vector<int> obc_q, obc_s;
  for (int j = 1; j < mnei; ++j)
    for (int i = 0; i < np; ++i) {
      const int sl = obc_slot[(size_t)j * np + i];
      if (sl < 0) continue;
      obc_q.push_back(j * np + i);
      obc_s.push_back(bnd_index[i] >= 0 ? -1 - sl : sl);
    }
  const int nobc = (int)obc_q.size();
  if (obc_q.empty()) {  // device arrays of at least one element
    obc_q.push_back(0);
    obc_s.push_back(0);
  }
  std::vector<int> obcg_ptr(np + 1, 0), obcg;
  for (int i = 0; i < np; ++i) {
    obcg_ptr[i + 1] = obcg_ptr[i] + (int)obcl[i].size();
    obcg.insert(obcg.end(), obcl[i].begin(), obcl[i].end());
  }
  std::vector<std::vector<int>> vbl(np);
  std::vector<int> lbc(std::max(M->nvel, 1), 0);
  std::vector<double> csii(std::max(M->nvel, 1), 0.0), siii(std::max(M->nvel, 1), 0.0);
  for (int I = 0; I < M->nvel; ++I) {
    lbc[I] = M->lbcodei[I];
    csii[I] = M->csii[I];
    siii[I] = M->siii[I];
    if (lbc[I] < 0 || (lbc[I] > 29 && lbc[I] != 30 && lbc[I] != 52))
      return fail(ADCGPU_E_UNSUPPORTED, "flow boundary type " + std::to_string(lbc[I]));
  }
  for (int J = 0; J < M->nvelme; ++J) {
    int I = M->me2gw[J] - 1;
    if (I < 0 || I >= M->nvel) return fail(ADCGPU_E_ARG, "ME2GW out of range");
    int b = M->nbv[I] - 1;
    if (b < 0 || b >= np) return fail(ADCGPU_E_ARG, "NBV out of range");
    vbl[b].push_back(I);
  }
  std::vector<int> h_nbv(std::max(M->nvel, 1), 0);
  for (int I = 0; I < M->nvel; ++I) h_nbv[I] = M->nbv[I] - 1;
  std::vector<int> vbc_ptr(np + 1, 0), vbc;
  for (int i = 0; i < np; ++i) {
    vbc_ptr[i + 1] = vbc_ptr[i] + (int)vbl[i].size();
    vbc.insert(vbc.end(), vbl[i].begin(), vbl[i].end());
  }

  // NCEle is recovered from NCEle*AREAS/2 in the wet/dry D2 gather
  for (int e = 0; e < ne; ++e)
    if (!(M->areas[e] > 0.0))
      return fail(ADCGPU_E_ARG, "element " + std::to_string(e + 1) + " has a non-positive AREAS");

  adcgpu_ctx* c = new adcgpu_ctx();
  c->device = device;
  c->np = np; c->ne = ne; c->mnei = mnei; c->neta = neta;
  c->nch = (np + RED_CHUNK - 1) / RED_CHUNK;
  c->jcg_halo = halo_blocks <= JCG_HALO_MAX ? halo_blocks : -1;
  // staging needs the point-to-point halo wait (every staged row belongs to a
  // tile within jcg_halo)
  c->jcg_hrows = c->jcg_halo >= 0 ? jcg_stage_rows(bw) : -1;
  // staged: each off-diagonal column as its position in the tile's staged
  // rows (fewer than 2^16 by the LDS bound), two slots per word
  std::vector<uint32_t> lcol;
  if (c->jcg_hrows >= 0) {
    lcol.assign((size_t)std::max(mnei / 2, 1) * np, 0u);
    for (int i = 0; i < np; ++i) {
      const int t0 = i / (JCG_CPB * RED_CHUNK) * (JCG_CPB * RED_CHUNK);
      const int slo = std::max(0, t0 - c->jcg_hrows);
      for (int j = 1; j < nneigh[i]; ++j) {
        const uint32_t v = (uint32_t)(col[(size_t)j * np + i] - slo);
        lcol[(size_t)((j - 1) / 2) * np + i] |= v << (16 * ((j - 1) & 1));
      }
    }
  }
  c->prm = *P;
  c->ifnlfa = P->nolifa == 0 ? 0.0 : 1.0;
  c->ifnlct = P->nolica == 0 ? 0.0 : 1.0;
  c->ifnlcat = P->nolicat == 0 ? 0.0 : 1.0;
  c->iflinbf = P->nolibf == 0 ? 1.0 : 0.0;
  c->ifnlbf = (P->nolibf == 1 || P->nolibf == 3) ? 1.0 : 0.0;
  c->hyb.on = P->nolibf == 2 ? 1.0 : 0.0;  // IFHYBF (src/nodalattr.F:1984-1988)
  c->tkmax = (double)2.0f / P->dtdp;
  c->habsmin = 0.8 * P->h0;
  c->hoff = 1.2 * P->h0;
  if (P->errorelev < 0.0) {
    delete c;
    return fail(ADCGPU_E_ARG, "ErrorElev must not be negative");
  }
  c->errorelev = P->errorelev > 0.0 ? P->errorelev : 1000.0;
  int rc = 0;
  auto U = [&](auto** p, const auto& v) {
    if (!rc) rc = upload(p, v);
    if (!rc) c->allocs.push_back((void*)*p);
  };
  auto A = [&](auto** p, size_t n) {
    if (!rc) rc = dmalloc(p, n);
    if (!rc) {
      c->allocs.push_back((void*)*p);
      (void)hipMemset(*p, 0, n * sizeof(**p));
    }
  };
  std::vector<double> vx(M->x, M->x + np), vy(M->y, M->y + np), vdp(M->dp, M->dp + np);
  std::vector<double> vt(P->tau0var, P->tau0var + np), ve(P->evm, P->evm + np);
  std::vector<double> vcor(P->corif, P->corif + np), vfric(P->fric, P->fric + np);
  std::vector<double> var(M->areas, M->areas + ne);
  U(&c->nm, nm0); U(&c->inc_ptr, cnt); U(&c->inc, inc); U(&c->slot, slot); U(&c->slot_cols, slot_cols); U(&c->nneigh, nneigh); U(&c->col, col);
  U(&c->jhrange, hrange);
  if (!lcol.empty()) U(&c->jlcol, lcol);
  U(&c->csrpos, csrpos); U(&c->bnd_index, bnd_index); U(&c->nbd, nbd0); U(&c->obc_q, obc_q);
  U(&c->obc_s, obc_s); c->nobc = nobc;
  c->nvel = M->nvel;
  for (int I = 0; I < M->nvel; ++I) {
    c->has_rad = c->has_rad || lbc[I] == 30 || lbc[I] == 52;
    c->has_spec = c->has_spec || lbc[I] == 2 || lbc[I] == 12 || lbc[I] == 22 || lbc[I] == 52;
  }
  c->h_nbv = h_nbv;
  c->h_bnd_index = bnd_index;
  U(&c->obcg_ptr, obcg_ptr); U(&c->obcg, obcg); U(&c->vbc_ptr, vbc_ptr); U(&c->vbc, vbc);
  U(&c->lbcodei, lbc); U(&c->fdx, fdx); U(&c->fdy, fdy); U(&c->areas, var); U(&c->x, vx);
  U(&c->y, vy); U(&c->dp, vdp); U(&c->tau0, vt); U(&c->evm, ve); U(&c->corif, vcor);
  U(&c->csii, csii); U(&c->siii, siii); U(&c->fric, vfric);
  double** nodal[] = {&c->eta1, &c->eta2, &c->etas, &c->uu1, &c->vv1, &c->uu2, &c->vv2,
                      &c->qx1, &c->qy1, &c->qx2, &c->qy2, &c->h1, &c->h2, &c->tk,
                      &c->etamax, &c->totalarea, &c->wsx1, &c->wsy1, &c->wsx2, &c->wsy2,
                      &c->pr1, &c->pr2, &c->gwce_lv, &c->jd, &c->jw};
  for (auto p : nodal) A(p, np);
  A(&c->alphal, ne); A(&c->coef, (size_t)mnei * np); A(&c->cs, (size_t)mnei * np);
  A(&c->obc, (size_t)std::max(neta, 1) * mnei);
  A(&c->nodecode, np); A(&c->nnodecode, np); A(&c->noff, ne); A(&c->noffold, ne);
  A(&c->mju, np); A(&c->etamax_it, np); A(&c->etas_prev, np);
  A(&c->umax, np); A(&c->umax_it, np);
  A(&c->contrib, 3 * (size_t)ne); A(&c->cx, 3 * (size_t)ne); A(&c->cy, 3 * (size_t)ne);
  A(&c->lhs_td, 3 * (size_t)ne); A(&c->lhs_ta, 3 * (size_t)ne); A(&c->lhs_tb, 3 * (size_t)ne);
  A(&c->ncele, ne);
  c->mixed = P->precision == ADCGPU_MIXED;
  if (c->mixed) {
    std::vector<float> f32(3 * (size_t)ne);
    for (size_t k = 0; k < f32.size(); ++k) f32[k] = (float)fdx[k];
    U(&c->fdx32, f32);
    for (size_t k = 0; k < f32.size(); ++k) f32[k] = (float)fdy[k];
    U(&c->fdy32, f32);
    f32.resize(ne);
    for (int e = 0; e < ne; ++e) f32[e] = (float)var[e];
    U(&c->areas32, f32);
    A(&c->lhs_td32, 3 * (size_t)ne); A(&c->lhs_ta32, 3 * (size_t)ne);
    A(&c->lhs_tb32, 3 * (size_t)ne);
  }
  A(&c->ep_part, c->nch); A(&c->scal, S_COUNT); A(&c->eta_bc, std::max(neta, 1));
  {
    const size_t nb = (size_t)(ne + 255) / 256;
    A(&c->wd.mask, nb * 4); A(&c->wd.bcnt, nb); A(&c->wd.bofs, nb);
    A(&c->wd.tkc, ne); A(&c->wd.lc, ne); A(&c->wd.m0, ne); A(&c->wd.pred, ne);
    A(&c->wd.wnode, ne); A(&c->wd.M, ne);
    const size_t nwb = (size_t)(ne + WD_BLOCK - 1) / WD_BLOCK;
    A(&c->wd.cv, ne); A(&c->wd.cE, (size_t)ne * WD_XW); A(&c->wd.xtab, nwb * WD_XMAX);
    A(&c->wd.xcnt, nwb);
    A(&c->wd.xval, nwb * WD_XMAX);
  }
  A(&c->flags, F_COUNT); A(&c->jpart, 2 * (size_t)c->nch * 4); A(&c->jsaved, 12);
  A(&c->jbar, JCG_BAR_WORDS);
  if (!rc) {
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) rc = fail(ADCGPU_E_HIP, hipGetErrorString(e));
  }
  if (!rc) rc = setup_jcg_grid(c);
  if (!rc) {
    // the LHS row kernel keeps MNEI x 512 doubles of LDS per block; the
    // wetting propagation stages candidates through LDS
    hipError_t e = hipSuccess;
    const void* rows_fns[] = {(const void*)k_lhs_rows<false, double>,
                              (const void*)k_lhs_rows<true, double>,
                              (const void*)k_lhs_rows<false, float>,
                              (const void*)k_lhs_rows<true, float>};
    for (const void* fn : rows_fns)
      if (e == hipSuccess)
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                mnei * RED_THREADS * (int)sizeof(double));
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)k_wd_link, hipFuncAttributeMaxDynamicSharedMemorySize,
                              WD_LINK_LDS);
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)k_wd_close, hipFuncAttributeMaxDynamicSharedMemorySize,
                              WD_CLOSE_LDS);
    if (e != hipSuccess) rc = fail(ADCGPU_E_UNSUPPORTED, std::string("LDS request: ") +
                                                         hipGetErrorString(e));
  }
  if (!rc) {
    std::vector<double> ones(ne, 1.0), pr(np, P->prbckgrnd_mh2o), emax(np, -99999.0);
    (void)hipMemcpy(c->alphal, ones.data(), ne * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->pr1, pr.data(), np * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->pr2, pr.data(), np * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->etamax, emax.data(), np * sizeof(double), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->umax, emax.data(), np * sizeof(double), hipMemcpyHostToDevice);
    std::vector<int> one_i(std::max(np, ne), 1);
    (void)hipMemcpy(c->nodecode, one_i.data(), np * sizeof(int), hipMemcpyHostToDevice);
    (void)hipMemcpy(c->noff, one_i.data(), ne * sizeof(int), hipMemcpyHostToDevice);
    int f[F_COUNT] = {0};
    f[F_NCCHANGE] = 1;
    (void)hipMemcpy(c->flags, f, sizeof(f), hipMemcpyHostToDevice);
    // the GWCE matrix's unchanging terms, once
    if (P->ilump != 0)
      hipLaunchKernelGGL(k_lhs_terms<true>, dim3(nblk(ne, 256)), dim3(256), 0, c->stream,
                         c->mesh(), c->lhs_td, c->lhs_ta, c->lhs_tb, c->lhs_td32, c->lhs_ta32,
                         c->lhs_tb32);
    else
      hipLaunchKernelGGL(k_lhs_terms<false>, dim3(nblk(ne, 256)), dim3(256), 0, c->stream,
                         c->mesh(), c->lhs_td, c->lhs_ta, c->lhs_tb, c->lhs_td32, c->lhs_ta32,
                         c->lhs_tb32);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = fail(ADCGPU_E_HIP, hipGetErrorString(e));
  }
  if (!rc && P->ilump == 0) {
    // the terms per off-diagonal entry: slot p's ta term goes to column ja,
    // its tb term to jb, in ascending element order; more than two
    // contributions to one entry (a non-manifold edge) keep the row pass
    const size_t nq = (size_t)(mnei - 1) * np;
    std::vector<int> pc1(nq, -1), pc2(nq, -1);
    bool ok = true;
    for (int i = 0; i < np && ok; ++i)
      for (int p = cnt[i]; p < cnt[i + 1] && ok; ++p) {
        const int jj[2] = {slot_cols[p] & 0xffff, slot_cols[p] >> 16};
        for (int w = 0; w < 2; ++w) {
          const size_t q = (size_t)(jj[w] - 1) * np + i;
          const int code = 2 * p + w;
          if (pc1[q] < 0) pc1[q] = code;
          else if (pc2[q] < 0) pc2[q] = code;
          else ok = false;
        }
      }
    if (ok) {
      // the codes are needed only here: uploaded outside the context's list
      // of allocations and freed below
      int *d1 = nullptr, *d2 = nullptr;
      if (!rc) rc = upload(&d1, pc1);
      if (!rc) rc = upload(&d2, pc2);
      if (c->mixed) {
        A(&c->lp_t1f, nq); A(&c->lp_t2f, nq);
      } else {
        A(&c->lp_t1, nq); A(&c->lp_t2, nq);
      }
      A(&c->lp_e1, nq); A(&c->lp_e2, nq);
      if (!rc) {
        const LhsTerms lt{c->lhs_td, c->lhs_ta, c->lhs_tb, c->lhs_td32, c->lhs_ta32, c->lhs_tb32, 0};
        hipLaunchKernelGGL(k_lhs_pair_setup, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                           c->stream, c->mesh(), lt, (const int*)d1, (const int*)d2, c->lp_t1,
                           c->lp_t2, c->lp_t1f, c->lp_t2f, c->lp_e1, c->lp_e2);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(ADCGPU_E_HIP, hipGetErrorString(e));
        c->lhs_pairs = rc == 0;
      }
      if (d1) (void)hipFree(d1);
      if (d2) (void)hipFree(d2);
    }
  }
  if (!rc && P->warm_start == 2) rc = ml_setup(c);
  if (rc) {
    adcgpu_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

void adcgpu_destroy(adcgpu_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& ev : c->jcg_events) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->eta_bc_table) (void)hipFree(c->eta_bc_table);
  for (void* p : {(void*)c->fb_ptr, (void*)c->fb_ent, (void*)c->nbv0, (void*)c->bndlen,
                  (void*)c->qn_table})
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

static double** field_ptr(adcgpu_ctx* c, int f) {
  switch (f) {
    case ADCGPU_ETA1: return &c->eta1;
    case ADCGPU_ETA2: return &c->eta2;
    case ADCGPU_ETAS: return &c->etas;
    case ADCGPU_UU1: return &c->uu1;
    case ADCGPU_VV1: return &c->vv1;
    case ADCGPU_UU2: return &c->uu2;
    case ADCGPU_VV2: return &c->vv2;
    case ADCGPU_QX1: return &c->qx1;
    case ADCGPU_QY1: return &c->qy1;
    case ADCGPU_QX2: return &c->qx2;
    case ADCGPU_QY2: return &c->qy2;
    case ADCGPU_H1: return &c->h1;
    case ADCGPU_H2: return &c->h2;
    case ADCGPU_TK: return &c->tk;
    case ADCGPU_ETAMAX: return &c->etamax;
    case ADCGPU_TOTALAREA: return &c->totalarea;
    case ADCGPU_WSX1: return &c->wsx1;
    case ADCGPU_WSY1: return &c->wsy1;
    case ADCGPU_WSX2: return &c->wsx2;
    case ADCGPU_WSY2: return &c->wsy2;
    case ADCGPU_PR1: return &c->pr1;
    case ADCGPU_PR2: return &c->pr2;
    case ADCGPU_GWCE_LV: return &c->gwce_lv;
    case ADCGPU_ETAS_PREV: return &c->etas_prev;
    case ADCGPU_TIP1: return c->ctip ? &c->tip1 : nullptr;
    case ADCGPU_TIP2: return c->ctip ? &c->tip2 : nullptr;
    case ADCGPU_UMAX: return &c->umax;
    case ADCGPU_COEF: return &c->coef;
    case ADCGPU_ALPHAL: return &c->alphal;
  }
  return nullptr;
}

static int** ifield_ptr(adcgpu_ctx* c, int f) {
  switch (f) {
    case ADCGPU_NODECODE: return &c->nodecode;
    case ADCGPU_NOFF: return &c->noff;
    case ADCGPU_MJU: return &c->mju;
    case ADCGPU_ETAMAX_IT: return &c->etamax_it;
    case ADCGPU_UMAX_IT: return &c->umax_it;
  }
  return nullptr;
}

static int64_t field_len(adcgpu_ctx* c, int f) {
  if (f == ADCGPU_COEF) return (int64_t)c->np * c->mnei;
  if (f == ADCGPU_NOFF || f == ADCGPU_ALPHAL) return c->ne;
  return c->np;
}

int adcgpu_set_field(adcgpu_ctx* c, int f, const void* host, int64_t count) {
  if (!c || !host) return fail(ADCGPU_E_ARG, "null argument");
  if (count != field_len(c, f)) return fail(ADCGPU_E_ARG, "field length");
  HIPCHK(hipSetDevice(c->device));
  if (double** p = field_ptr(c, f)) {
    // COEF(NP,MNEI) column-major is the device layout too; a COEF from the
    // host replaces any scaled off-diagonals a deferred solve left
    if (f == ADCGPU_COEF) c->coef_scaled = false;
    HIPCHK(hipMemcpyAsync(*p, host, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else if (int** q = ifield_ptr(c, f)) {
    HIPCHK(hipMemcpyAsync(*q, host, count * sizeof(int), hipMemcpyHostToDevice, c->stream));
  } else {
    return fail(ADCGPU_E_ARG, "unknown field");
  }
  // new wet/dry flags: the next matrix rebuild forms NCEle afresh; an ALPHAL
  // from the host may differ from 1, so the matrix terms are recomputed from
  // it until the next wet/dry pass
  if (f == ADCGPU_NODECODE || f == ADCGPU_NOFF) c->ncele_fresh = false;
  if (f == ADCGPU_ALPHAL)
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->flags + F_ALPHA, 1, 1);
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_launch("set_field");
}

static int restore_coef(adcgpu_ctx* c);

int adcgpu_get_field(adcgpu_ctx* c, int f, void* host, int64_t count) {
  if (!c || !host) return fail(ADCGPU_E_ARG, "null argument");
  if (count != field_len(c, f)) return fail(ADCGPU_E_ARG, "field length");
  HIPCHK(hipSetDevice(c->device));
  if (f == ADCGPU_COEF)
    if (int rc = restore_coef(c)) return rc;
  if (double** p = field_ptr(c, f)) {
    HIPCHK(hipMemcpyAsync(host, *p, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else if (int** q = ifield_ptr(c, f)) {
    HIPCHK(hipMemcpyAsync(host, *q, count * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  } else {
    return fail(ADCGPU_E_ARG, "unknown field");
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int set_elevdisc(adcgpu_ctx* c);

int adcgpu_cold_start(adcgpu_ctx* c, const double* eta2, const int32_t* nodecode,
                      const int32_t* noff) {
  if (!c || !eta2 || !nodecode || !noff) return fail(ADCGPU_E_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const size_t np = c->np, ne = c->ne;
  // everything else starts cold too, so that a context can be cold-started
  // again: velocities, fluxes, increments, friction, forcing, the running
  // maxima, the JCG start history and ITPACK's saved state
  double* zero_d[] = {c->uu1, c->vv1, c->uu2, c->vv2, c->qx1, c->qy1, c->qx2, c->qy2,
                      c->etas, c->etas_prev, c->tk, c->gwce_lv, c->h1, c->h2,
                      c->wsx1, c->wsy1, c->wsx2, c->wsy2};
  for (double* p : zero_d) HIPCHK(hipMemsetAsync(p, 0, np * 8, c->stream));
  HIPCHK(hipMemsetAsync(c->etamax_it, 0, np * 4, c->stream));
  HIPCHK(hipMemsetAsync(c->umax_it, 0, np * 4, c->stream));
  HIPCHK(hipMemsetAsync(c->jsaved, 0, 12 * sizeof(double), c->stream));
  if (c->nfluxf)
    for (double* p : {c->qn0, c->qn1, c->qn2})
      HIPCHK(hipMemsetAsync(p, 0, (size_t)c->nvel * 8, c->stream));
  {
    std::vector<double> pr(np, c->prm.prbckgrnd_mh2o), low(np, -99999.0), ones(ne, 1.0);
    HIPCHK(hipMemcpyAsync(c->pr1, pr.data(), np * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->pr2, pr.data(), np * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->etamax, low.data(), np * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->umax, low.data(), np * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->alphal, ones.data(), ne * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->coef_scaled = false;
  c->ncele_fresh = false;
  HIPCHK(hipMemcpyAsync(c->eta2, eta2, np * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->eta1, eta2, np * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->nodecode, nodecode, np * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->nnodecode, nodecode, np * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->noff, noff, ne * 4, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_total_area, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream, c->mesh(),
                     c->state());
  int f[F_COUNT] = {0};
  f[F_NCCHANGE] = 1;
  HIPCHK(hipMemcpyAsync(c->flags, f, sizeof(f), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->nfluxf)
    if (int rc = set_elevdisc(c)) return rc;
  return check_launch("cold start");
}

// The time-level shifts of the step's start; qn_next: this step's QN2 on the
// device (null: zero until adcgpu_set_flux).
static int begin_step(adcgpu_ctx* c, int it, const double* qn_next) {
  c->cur_it = it;
  if (c->has_rad && !c->nfluxf)
    return fail(ADCGPU_E_ARG, "radiating boundaries need adcgpu_set_flux_boundary");
  if (it == 1)
    hipLaunchKernelGGL(k_init_h, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream, c->np, c->dp,
                       c->eta2, c->ifnlfa, c->h1, c->h2);
  // QX1 <- QX2, UU1 <- UU2, ... WSX1 <- WSX2, PR1 <- PR2 (src/timestep.F:385-431)
  std::swap(c->qx1, c->qx2); std::swap(c->qy1, c->qy2);
  std::swap(c->uu1, c->uu2); std::swap(c->vv1, c->vv2);
  // the normal flux's time levels (after the velocities' shift: radiation
  // reads the last step's velocities as UU1/VV1)
  if (c->nfluxf)
    hipLaunchKernelGGL(k_flux_shift, dim3(nblk(c->nvel, 256)), dim3(256), 0, c->stream,
                       c->mesh(), c->state(), qn_next, c->nvel);
  if (c->prm.nws != 0) {
    std::swap(c->wsx1, c->wsx2); std::swap(c->wsy1, c->wsy2); std::swap(c->pr1, c->pr2);
  }
  if (c->ctip) {
    // TIP1 <- TIP2 (src/timestep.F:412-413), then this step's TIP2
    // (src/timestep.F:841, 1507-1562): the per-constituent phases and
    // multipliers on the host, the nodal sum on the device
    std::swap(c->tip1, c->tip2);
    const double timeloc = it * c->prm.dtdp + c->tip_statim * 86400.0;
    const double timeh = it * c->prm.dtdp + (c->tip_statim - c->tip_reftim) * 86400.0;
    double ramptip = 1.0;
    if (c->tip_nramp != 0) ramptip = std::tanh((2.0 * timeloc / 86400.0) / c->tip_dramp);
    if (c->tip_dramp < 1.0e-6) ramptip = 1.0;
    TipPrm tp;
    std::memset(&tp, 0, sizeof(tp));
    tp.ntif = (int)c->tip_tpk.size();
    for (int j = 0; j < tp.ntif; ++j) {
      const double pert = c->tip_pert[j];
      const int ncyc = pert == 0.0 ? 0 : (int)(timeh / pert);
      tp.argt[j] = c->tip_amigt[j] * (timeh - ncyc * pert) + c->tip_facet[j];
      tp.tpmul[j] = ramptip * c->tip_etrf[j] * c->tip_tpk[j] * c->tip_fft[j];
      tp.na[j] = std::min((int)std::lround(c->tip_amigt[j] / 7e-5), 2);
    }
    hipLaunchKernelGGL(k_tide_pot, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream, c->np, tp,
                       (const double*)c->tip_ln, (const double*)c->tip_slam, c->tip2);
  }
  return check_launch("begin_step");
}

int adcgpu_begin_step(adcgpu_ctx* c, int it) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  return begin_step(c, it, nullptr);
}

int adcgpu_set_flux(adcgpu_ctx* c, const double* qn2) {
  if (!c || !qn2) return fail(ADCGPU_E_ARG, "null argument");
  if (!c->nfluxf) return fail(ADCGPU_E_ARG, "no flux boundaries set up (adcgpu_set_flux_boundary)");
  HIPCHK(hipMemcpyAsync(c->qn2, qn2, (size_t)c->nvel * 8, hipMemcpyHostToDevice, c->stream));
  return 0;
}

static FricPrm fric_prm(const adcgpu_ctx* c) {
  return FricPrm{c->uu1, c->vv1, c->dp, c->eta2, c->fric, c->ifnlfa, c->iflinbf, c->ifnlbf,
                 c->tkmax, c->hyb};
}

int adcgpu_friction(adcgpu_ctx* c) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  hipLaunchKernelGGL(k_friction, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream, c->np,
                     fric_prm(c), c->tk);
  return check_launch("friction");
}

int adcgpu_set_forcing(adcgpu_ctx* c, const double* wsx2, const double* wsy2, const double* pr2) {
  if (!c || !wsx2 || !wsy2 || !pr2) return fail(ADCGPU_E_ARG, "null argument");
  const size_t n = (size_t)c->np * 8;
  HIPCHK(hipMemcpyAsync(c->wsx2, wsx2, n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->wsy2, wsy2, n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->pr2, pr2, n, hipMemcpyHostToDevice, c->stream));
  return 0;
}

// One GWCE assembly and solve.  With `deferred` the solve leaves the
// off-diagonals of COEF scaled and the next assembly (or adcgpu_run's end)
// restores them; the resident run uses it, since most steps rebuild COEF.
// What the solve and the pass after it need from the assembly.
struct GwceStep {
  bool solve = false;    // a JCG solve follows (ILump = 0)
  bool learned = false;  // with the learned second start
  bool collect = false;  // training samples were recorded at slot0
  int slot0 = 0;
};

// The GWCE assembly up to the solve: matrix, right-hand side, time-level
// rotation, the learned start's inputs, and the node pass that scales the
// rows for JCG (or, with ILump = 1, solves).
static int gwce_assemble(adcgpu_ctx* c, const double* d_eta_bc, GwceStep* g) {
  const int np = c->np, ne = c->ne;
  const bool lumped = c->prm.ilump != 0;
  const size_t rows_lds = (size_t)c->mnei * RED_THREADS * sizeof(double);
  const double* rcs = c->coef_scaled ? (const double*)c->cs : nullptr;
  const double* rw = c->coef_scaled ? (const double*)c->jw : nullptr;
  // a rebuild multiplies the cached terms by each element's NCEle, which the
  // momentum element pass left unless the flags changed since
  if (!c->ncele_fresh) {
    hipLaunchKernelGGL(k_ncele, dim3(nblk(ne, 256)), dim3(256), 0, c->stream, c->mesh(),
                       c->state());
    c->ncele_fresh = true;
  }
  const LhsTerms lt{c->lhs_td,   c->lhs_ta,   c->lhs_tb,       c->lhs_td32,
                    c->lhs_ta32, c->lhs_tb32, c->lhs_pairs ? 1 : 0};
  // contrib, cx and cy are free until the GWCE right-hand side
  if (!lumped)  // the element pass only runs while some ALPHAL differs from 1
    hipLaunchKernelGGL(k_lhs_elem<false>, dim3(nblk(ne, 256)), dim3(256), 0, c->stream,
                       c->mesh(), c->state(), c->flags, c->contrib, c->cx, c->cy);
  const dim3 rg(c->nch), rb(RED_THREADS);
  if (c->mixed) {
    if (lumped)
      hipLaunchKernelGGL((k_lhs_rows<true, float>), rg, rb, rows_lds, c->stream, c->mesh(),
                         c->state(), c->flags, c->contrib, c->cx, c->cy, lt, c->ep_part, rcs, rw);
    else
      hipLaunchKernelGGL((k_lhs_rows<false, float>), rg, rb, rows_lds, c->stream, c->mesh(),
                         c->state(), c->flags, c->contrib, c->cx, c->cy, lt, c->ep_part, rcs, rw);
  } else {
    if (lumped)
      hipLaunchKernelGGL((k_lhs_rows<true, double>), rg, rb, rows_lds, c->stream, c->mesh(),
                         c->state(), c->flags, c->contrib, c->cx, c->cy, lt, c->ep_part, rcs, rw);
    else
      hipLaunchKernelGGL((k_lhs_rows<false, double>), rg, rb, rows_lds, c->stream, c->mesh(),
                         c->state(), c->flags, c->contrib, c->cx, c->cy, lt, c->ep_part, rcs, rw);
  }
  if (c->lhs_pairs) {
    // the cached rebuild (k_lhs_rows above only restored or left at once)
    const LhsPairs lp{c->lp_t1, c->lp_t2, c->lp_t1f, c->lp_t2f, c->lp_e1, c->lp_e2};
    const dim3 pg(nblk(np, 256), c->mnei);
    if (c->mixed)
      hipLaunchKernelGGL(k_lhs_pairs<float>, pg, dim3(256), 0, c->stream, c->mesh(), c->state(),
                         (const int*)c->flags, lt, lp);
    else
      hipLaunchKernelGGL(k_lhs_pairs<double>, pg, dim3(256), 0, c->stream, c->mesh(), c->state(),
                         (const int*)c->flags, lt, lp);
    hipLaunchKernelGGL(k_lhs_ep, rg, rb, 0, c->stream, c->mesh(), c->state(),
                       (const int*)c->flags, c->ep_part);
  }
  c->coef_scaled = false;
  RhsPrm q{c->ifnlfa, c->ifnlct, c->ifnlcat, lumped ? 4.0 : 2.0, lumped ? 0.0 : 1.0,
           lumped ? 1 : 0, c->prm.nws};
  // its last block finishes the matrix (EP, elevation-boundary rows) and
  // clears NCChange
  static_assert(RED_XTHREADS == 256, "the matrix-finishing block is a 256-thread block");
  float* contrib32 = reinterpret_cast<float*>(c->contrib);
  if (c->mixed)
    hipLaunchKernelGGL(k_rhs_elem32, dim3(nblk(ne, 256) + 1), dim3(256), 0, c->stream, c->mesh(),
                       c->state(), q, c->flags, contrib32, c->nch, c->ep_part, c->scal);
  else
    hipLaunchKernelGGL(k_rhs_elem, dim3(nblk(ne, 256) + 1), dim3(256), 0, c->stream, c->mesh(),
                       c->state(), q, c->flags, c->contrib, c->nch, c->ep_part, c->scal);
  // ETA1 <- ETA2, H1 <- H2 (src/gwce.F:1622-1631) by rotation
  std::swap(c->eta1, c->eta2);
  std::swap(c->h1, c->h2);
  // the learned start's correction and the training samples are formed from
  // the state before the solve (eta^k now in ETA1)
  const bool learned = c->prm.warm_start == 2 && !lumped;
  const bool collect = c->ml_collect > 0 && c->ml_ready && !lumped;
  const int slot0 = collect ? (int)(c->smp_total % c->smp.cap) : 0;
  if (learned || collect) {
    const MlIn in = ml_in(c, c->eta1);
    if (learned) ml_infer(c, in);
    if (collect)
      hipLaunchKernelGGL(k_ml_collect, dim3(nblk(c->ml_collect, 128)), dim3(256), 0, c->stream,
                         in, c->smp, slot0, c->ml_collect, c->cur_it);
  }
  JcgPre pre{c->jw, c->jbar, jcg_barrier_words(c), c->flags, c->etas_prev, c->prm.warm_start,
             (c->prm.warm_start != 0 || c->ml_ready) ? 1 : 0, c->ml_corr, c->ml_u};
  const dim3 ng(nblk(np, 256)), nbk(256);
  // with ILump = 1 the explicit solve of the lumped GWCE ends the node pass; no JCG
  if (c->mixed) {
    if (lumped)
      hipLaunchKernelGGL((k_rhs_node<true, float>), ng, nbk, 0, c->stream, c->mesh(), c->state(),
                         (const float*)contrib32, d_eta_bc, c->scal, pre);
    else
      hipLaunchKernelGGL((k_rhs_node<false, float>), ng, nbk, 0, c->stream, c->mesh(), c->state(),
                         (const float*)contrib32, d_eta_bc, c->scal, pre);
  } else {
    if (lumped)
      hipLaunchKernelGGL((k_rhs_node<true, double>), ng, nbk, 0, c->stream, c->mesh(), c->state(),
                         (const double*)c->contrib, d_eta_bc, c->scal, pre);
    else
      hipLaunchKernelGGL((k_rhs_node<false, double>), ng, nbk, 0, c->stream, c->mesh(),
                         c->state(), (const double*)c->contrib, d_eta_bc, c->scal, pre);
  }
  g->solve = !lumped;
  g->learned = learned;
  g->collect = collect;
  g->slot0 = slot0;
  return check_launch(lumped ? "gwce assembly and lumped solve" : "gwce assembly");
}

// The arguments of the step's JCG launch (prescaled rows, ETA2 written at the end).
static JcgArgs gwce_jcg_args(adcgpu_ctx* c, const GwceStep& g, bool deferred) {
  return make_jcg_args(c, c->coef, c->col, c->nneigh, c->gwce_lv, c->etas, c->np, c->prm.itmax,
                       c->prm.convcr, c->eta2, c->eta1, c->nodecode, true, deferred,
                       g.learned ? c->ml_u : nullptr);
}

// After the solve: the off-diagonals left scaled (deferred) and the training
// targets from this step's solution.
static int gwce_finish(adcgpu_ctx* c, const GwceStep& g, bool deferred) {
  if (!g.solve) return 0;
  if (deferred) c->coef_scaled = true;
  if (!g.collect) return 0;
  hipLaunchKernelGGL(k_ml_target, dim3(nblk(c->ml_collect, 256)), dim3(256), 0, c->stream,
                     c->smp, g.slot0, c->ml_collect, (const double*)c->etas);
  c->smp_total += c->ml_collect;
  return check_launch("training samples");
}

static int gwce_device(adcgpu_ctx* c, const double* d_eta_bc, bool deferred) {
  GwceStep g;
  int rc = gwce_assemble(c, d_eta_bc, &g);
  if (!rc && g.solve)
    rc = launch_jcg(c, c->coef, c->col, c->nneigh, c->gwce_lv, c->etas, c->np, c->prm.itmax,
                    c->prm.convcr, c->eta2, c->eta1, c->nodecode, true, deferred,
                    g.learned ? c->ml_u : nullptr);
  if (!rc) rc = gwce_finish(c, g, deferred);
  return rc;
}

// COEF with the off-diagonals of a deferred unscaling restored.
static int restore_coef(adcgpu_ctx* c) {
  if (!c->coef_scaled) return 0;
  hipLaunchKernelGGL(k_coef_restore, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream, c->mesh(),
                     c->state(), (const double*)c->cs, (const double*)c->jw);
  c->coef_scaled = false;
  return check_launch("coef restore");
}

int adcgpu_solve_gwce(adcgpu_ctx* c, int it, const double* eta_bc) {
  (void)it;
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  if (c->neta > 0) {
    if (!eta_bc) return fail(ADCGPU_E_ARG, "eta_bc required");
    HIPCHK(hipMemcpyAsync(c->eta_bc, eta_bc, c->neta * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  }
  return gwce_device(c, c->eta_bc, false);
}

// Wet/dry; without `alpha_pass` the ALPHAL pass is left to the momentum
// element pass (the resident run).
static int wetdry_device(adcgpu_ctx* c, bool alpha_pass) {
  const int np = c->np, ne = c->ne;
  // the flags change here; the momentum element pass records NCEle again
  c->ncele_fresh = false;
  if (c->prm.nolifa != 2) {
    hipLaunchKernelGGL(k_h2, dim3(nblk(np, 256)), dim3(256), 0, c->stream, np, c->dp, c->eta2,
                       c->ifnlfa, c->h2);
    return check_launch("h2");
  }
  hipLaunchKernelGGL(k_wd_d1, dim3(nblk(np, 256)), dim3(256), 0, c->stream, c->mesh(),
                     c->state(), c->flags, c->prm.h0, c->habsmin);
  WdPrm q{c->hoff, c->prm.velmin, c->iflinbf, c->ifnlbf, c->tkmax, c->hyb};
  const int nb = nblk(ne, 256);
  hipLaunchKernelGGL(k_wd_elem, dim3(nb), dim3(256), 0, c->stream, c->mesh(), c->state(), q,
                     c->wd);
  hipLaunchKernelGGL(k_wd_index, dim3(1), dim3(1024), 0, c->stream, c->wd.bcnt, nb, c->wd.bofs,
                     c->flags);
  hipLaunchKernelGGL(k_wd_pred, dim3(nb), dim3(256), 0, c->stream, c->mesh(), c->state(), c->wd);
  const int nwb = std::min((ne + WD_BLOCK - 1) / WD_BLOCK, 256);
  hipLaunchKernelGGL(k_wd_close, dim3(nwb), dim3(WD_BLOCK), WD_CLOSE_LDS, c->stream, c->wd,
                     c->flags);
  hipLaunchKernelGGL(k_wd_link, dim3(1), dim3(256), WD_LINK_LDS, c->stream, c->wd, c->flags,
                     c->tk);
  // cx is free until the momentum step: it holds NCEle*AREAS/2 per incidence slot
  hipLaunchKernelGGL(k_wd_final_area, dim3(nblk(ne, 256)), dim3(256), 0, c->stream, c->mesh(),
                     c->state(), c->wd, c->flags, c->tk, c->cx);
  hipLaunchKernelGGL(k_wd_d2, dim3(nblk(np, 256)), dim3(256), 0, c->stream, c->mesh(),
                     c->state(), c->cx, c->flags, c->ifnlfa);
  if (alpha_pass)
    hipLaunchKernelGGL(k_wd_alpha, dim3(nblk(ne, 256)), dim3(256), 0, c->stream, c->mesh(),
                       c->state(), c->flags, 1.0e9);
  return check_launch("wetdry");
}

int adcgpu_wetdry(adcgpu_ctx* c, int it) {
  (void)it;
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  return wetdry_device(c, true);
}

// Momentum; with `alpha` the element pass forms ALPHAL first.
static int momentum_device(adcgpu_ctx* c, bool alpha) {
  MomPrm q{c->ifnlct, c->prm.nws, c->cur_it, c->errorelev, c->flags};
  const dim3 eg(nblk(c->ne, 256)), ng(nblk(c->np, 256)), bk(256);
  if (c->mixed) {
    float* cx32 = reinterpret_cast<float*>(c->cx);
    float* cy32 = reinterpret_cast<float*>(c->cy);
    if (alpha)
      hipLaunchKernelGGL(k_mom_elem32<true>, eg, bk, 0, c->stream, c->mesh(), c->state(), q, cx32,
                         cy32, c->flags);
    else
      hipLaunchKernelGGL(k_mom_elem32<false>, eg, bk, 0, c->stream, c->mesh(), c->state(), q,
                         cx32, cy32, c->flags);
    hipLaunchKernelGGL(k_mom_node<float>, ng, bk, 0, c->stream, c->mesh(), c->state(), q,
                       (const float*)cx32, (const float*)cy32);
  } else {
    if (alpha)
      hipLaunchKernelGGL(k_mom_elem<true>, eg, bk, 0, c->stream, c->mesh(), c->state(), q, c->cx,
                         c->cy, c->flags);
    else
      hipLaunchKernelGGL(k_mom_elem<false>, eg, bk, 0, c->stream, c->mesh(), c->state(), q,
                         c->cx, c->cy, c->flags);
    hipLaunchKernelGGL(k_mom_node<double>, ng, bk, 0, c->stream, c->mesh(), c->state(), q,
                       (const double*)c->cx, (const double*)c->cy);
  }
  c->ncele_fresh = true;
  return check_launch("momentum");
}

int adcgpu_momentum(adcgpu_ctx* c) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  return momentum_device(c, false);
}

int adcgpu_timestep(adcgpu_ctx* c, int it, const double* eta_bc, const double* wsx2,
                    const double* wsy2, const double* pr2) {
  int rc = adcgpu_begin_step(c, it);
  if (!rc) rc = adcgpu_friction(c);
  if (!rc && c->prm.nws != 0) rc = adcgpu_set_forcing(c, wsx2, wsy2, pr2);
  if (!rc) rc = adcgpu_solve_gwce(c, it, eta_bc);
  if (!rc) rc = adcgpu_wetdry(c, it);
  if (!rc) rc = adcgpu_momentum(c);
  return rc;
}

int adcgpu_last_numitr(adcgpu_ctx* c, int* numitr, int* ier) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  int f[F_COUNT];
  HIPCHK(hipMemcpyAsync(f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (numitr) *numitr = f[F_NUMITR];
  if (ier) *ier = f[F_IER];
  return flag_errors(f);
}

int adcgpu_synchronize(adcgpu_ctx* c) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int adcgpu_holland(adcgpu_ctx* c, double t, const double* prm13) {
  if (!c || !prm13) return fail(ADCGPU_E_ARG, "null argument");
  HollandPrm p;
  for (int k = 0; k < 13; ++k) p.v[k] = prm13[k];
  p.t = t;
  hipLaunchKernelGGL(k_holland, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream, c->np, c->x,
                     c->y, p, c->wsx2, c->wsy2, c->pr2);
  return check_launch("holland");
}

// A resident run's set-up: the elevation-boundary table on the device and
// the run's counters cleared.
static int run_prepare(adcgpu_ctx* c, int nsteps, const double* eta_bc_table,
                       const double* holland) {
  if (c->prm.nws != 0 && !holland) return fail(ADCGPU_E_ARG, "met forcing required");
  if (c->nfluxf && c->has_spec && c->qn_table_steps < nsteps)
    return fail(ADCGPU_E_ARG, "flux boundaries: adcgpu_set_flux_table rows for every step required");
  if (c->neta > 0) {
    if (!eta_bc_table) return fail(ADCGPU_E_ARG, "eta_bc_table required");
    if (c->eta_bc_table_steps < nsteps) {
      if (c->eta_bc_table) (void)hipFree(c->eta_bc_table);
      c->eta_bc_table = nullptr;
      HIPCHK(hipMalloc(&c->eta_bc_table, (size_t)nsteps * c->neta * sizeof(double)));
      c->eta_bc_table_steps = nsteps;
    }
    HIPCHK(hipMemcpyAsync(c->eta_bc_table, eta_bc_table, (size_t)nsteps * c->neta * 8,
                          hipMemcpyHostToDevice, c->stream));
  }
  // F_ITSUM and F_NREBUILD are adjacent: one fill clears both
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->flags + F_ITSUM, 0, 2);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->flags + F_NITMAX, 0, 1);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->flags + F_MLUSED, 0, 1);
  return check_launch("run set-up");
}

// A resident step up to the GWCE solve: begin, friction (with this step's
// Holland forcing in the same pass), the assembly.
static int run_step_assemble(adcgpu_ctx* c, int it, int s, const double* holland, GwceStep* g) {
  const bool rows = c->nfluxf && c->qn_table_steps > s;
  int rc = begin_step(c, it, rows ? c->qn_table + (size_t)s * c->nvel : nullptr);
  if (!rc) {
    if (holland) {
      // friction and this step's forcing in one pass
      HollandPrm hp;
      for (int k = 0; k < 13; ++k) hp.v[k] = holland[k];
      hp.t = it * c->prm.dtdp;
      hipLaunchKernelGGL(k_friction_holland, dim3(nblk(c->np, 256)), dim3(256), 0, c->stream,
                         c->np, fric_prm(c), c->tk, c->x, c->y, hp, c->wsx2, c->wsy2, c->pr2);
      rc = check_launch("friction and forcing");
    } else {
      rc = adcgpu_friction(c);
    }
  }
  if (!rc)
    rc = gwce_assemble(c, c->neta > 0 ? c->eta_bc_table + (size_t)s * c->neta : c->eta_bc, g);
  return rc;
}

// A resident step after the solve: wet/dry and momentum (with wet/dry the
// ALPHAL pass rides on the momentum element pass).
static int run_step_finish(adcgpu_ctx* c, const GwceStep& g) {
  int rc = gwce_finish(c, g, true);
  const bool wd = c->prm.nolifa == 2;
  if (!rc) rc = wetdry_device(c, !wd);
  if (!rc) rc = momentum_device(c, wd);
  return rc;
}

// The run's counters from the device flags; the sticky error, if any.
static int run_read_flags(adcgpu_ctx* c, bool* error) {
  int f[F_COUNT];
  HIPCHK(hipMemcpyAsync(f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->run_iters = f[F_ITSUM];
  c->run_rebuilds = f[F_NREBUILD];
  c->run_itmax = f[F_NITMAX];
  c->run_errstep = f[F_ERRSTEP];
  c->run_mlused = f[F_MLUSED];
  *error = f[F_ERROR] != 0;
  return flag_errors(f);
}

int adcgpu_run(adcgpu_ctx* c, int it0, int nsteps, const double* eta_bc_table,
               const double* holland, int use_graph) {
  if (!c || nsteps < 0) return fail(ADCGPU_E_ARG, "invalid argument");
  if (use_graph != 0)
    return fail(ADCGPU_E_UNSUPPORTED, "graph capture of the step is not implemented");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = run_prepare(c, nsteps, eta_bc_table, holland)) return rc;
  for (int s = 0; s < nsteps; ++s) {
    GwceStep g;
    int rc = run_step_assemble(c, it0 + s, s, holland, &g);
    if (!rc && g.solve)
      rc = launch_jcg(c, c->coef, c->col, c->nneigh, c->gwce_lv, c->etas, c->np, c->prm.itmax,
                      c->prm.convcr, c->eta2, c->eta1, c->nodecode, true, true,
                      g.learned ? c->ml_u : nullptr);
    if (!rc) rc = run_step_finish(c, g);
    if (rc) return rc;
    if ((s & 63) == 63) {
      bool error = false;
      const int frc = run_read_flags(c, &error);
      if (error) return frc;
    }
  }
  if (int rc = restore_coef(c)) return rc;
  bool error = false;
  return run_read_flags(c, &error);
}

}  // extern "C"

// ---- batches of contexts (ensembles, uneven meshes)
struct adcgpu_batch {
  std::vector<adcgpu_ctx*> m;
  int device = 0;
  hipStream_t stream = nullptr;
  // the members' JCG solves, grouped into cooperative launches whose blocks
  // are all resident: group g is members[gm[g] .. gm[g+1]-1] of the solving
  // members `sm`, launched with first-block offsets dfirst[gf[g] ..]
  std::vector<int> sm, gm, gf;
  std::vector<bool> gdual;
  int* dfirst = nullptr;
  JcgArgs* dargs = nullptr;  // [2][n]: the launch arguments by step parity
  JcgArgs* hargs = nullptr;  // pinned host copy of what dargs holds
  bool hvalid[2] = {false, false};
};

namespace {
void batch_free(adcgpu_batch* b) {
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->dfirst) (void)hipFree(b->dfirst);
  if (b->dargs) (void)hipFree(b->dargs);
  if (b->hargs) (void)hipHostFree(b->hargs);
  delete b;
}
}  // namespace

extern "C" {

int adcgpu_batch_create(adcgpu_ctx* const* members, int n, adcgpu_batch** out) {
  if (!members || n <= 0 || !out) return fail(ADCGPU_E_ARG, "invalid argument");
  *out = nullptr;
  for (int k = 0; k < n; ++k) {
    if (!members[k]) return fail(ADCGPU_E_ARG, "null member");
    if (members[k]->device != members[0]->device)
      return fail(ADCGPU_E_ARG, "batch members must share one device");
    if (members[k]->jcg_mem)
      return fail(ADCGPU_E_UNSUPPORTED, "batch members must fit the register-resident JCG");
    for (int l = 0; l < k; ++l)
      if (members[l] == members[k]) return fail(ADCGPU_E_ARG, "a context appears twice");
  }
  adcgpu_batch* b = new adcgpu_batch();
  b->m.assign(members, members + n);
  b->device = members[0]->device;
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
    batch_free(b);
    return fail(ADCGPU_E_HIP, "stream creation");
  }
  // greedy grouping in member order: a group holds members with the same
  // start (one or two) whose JCG grids fit on the device together
  const int cap = members[0]->jcg_cap;
  for (int k = 0; k < n; ++k)
    if (members[k]->prm.ilump == 0) b->sm.push_back(k);
  int sum = 0;
  bool dual = false;
  std::vector<int> first;
  for (size_t q = 0; q < b->sm.size(); ++q) {
    adcgpu_ctx* c = members[b->sm[q]];
    const bool d = c->prm.warm_start == 2;
    if (q == 0 || sum + c->jcg_grid > cap || d != dual) {
      b->gm.push_back((int)q);
      b->gf.push_back((int)first.size());
      b->gdual.push_back(d);
      first.push_back(0);
      sum = 0;
      dual = d;
    }
    sum += c->jcg_grid;
    first.push_back(sum);
  }
  b->gm.push_back((int)b->sm.size());
  int rc = 0;
  if (!first.empty()) rc = upload(&b->dfirst, first);
  if (!rc && !b->sm.empty()) rc = dmalloc(&b->dargs, 2 * b->sm.size());
  if (!rc && !b->sm.empty() &&
      hipHostMalloc((void**)&b->hargs, 2 * b->sm.size() * sizeof(JcgArgs), 0) != hipSuccess)
    rc = fail(ADCGPU_E_HIP, "pinned host buffer");
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return 0;
}

void adcgpu_batch_destroy(adcgpu_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
}

int adcgpu_batch_run(adcgpu_batch* b, int it0, int nsteps, const double* const* eta_bc_tables,
                     const double* const* holland) {
  if (!b || nsteps < 0) return fail(ADCGPU_E_ARG, "invalid argument");
  HIPCHK(hipSetDevice(b->device));
  const int n = (int)b->m.size();
  auto table = [&](int k) { return eta_bc_tables ? eta_bc_tables[k] : nullptr; };
  auto hol = [&](int k) { return holland ? holland[k] : nullptr; };
  // every member's work goes to the batch's stream for the run, so that the
  // shared JCG launches are ordered after and before each member's passes
  std::vector<hipStream_t> own(n);
  for (int k = 0; k < n; ++k) {
    own[k] = b->m[k]->stream;
    HIPCHK(hipStreamSynchronize(own[k]));
    b->m[k]->stream = b->stream;
  }
  auto restore = [&]() {
    (void)hipStreamSynchronize(b->stream);
    for (int k = 0; k < n; ++k) b->m[k]->stream = own[k];
  };
  int rc = 0;
  for (int k = 0; k < n && !rc; ++k) rc = run_prepare(b->m[k], nsteps, table(k), hol(k));
  std::vector<GwceStep> g(n);
  const size_t ns = b->sm.size();
  for (int s = 0; s < nsteps && !rc; ++s) {
    for (int k = 0; k < n && !rc; ++k) rc = run_step_assemble(b->m[k], it0 + s, s, hol(k), &g[k]);
    // the solves: this step's arguments (they alternate with the time-level
    // rotation, so each parity's are uploaded once), then one launch per group
    const int par = s & 1;
    if (!rc && ns) {
      JcgArgs* h = b->hargs + par * ns;
      std::vector<JcgArgs> cur(ns);
      for (size_t q = 0; q < ns; ++q) {
        adcgpu_ctx* c = b->m[b->sm[q]];
        cur[q] = gwce_jcg_args(c, g[b->sm[q]], true);
      }
      for (size_t gi = 0; gi + 1 < b->gm.size(); ++gi) {
        // a group runs staged only if all its members can
        bool st = true;
        for (int q = b->gm[gi]; q < b->gm[gi + 1]; ++q) st = st && cur[q].hrows >= 0;
        if (!st)
          for (int q = b->gm[gi]; q < b->gm[gi + 1]; ++q) cur[q].hrows = -1;
      }
      if (!b->hvalid[par] || std::memcmp(h, cur.data(), ns * sizeof(JcgArgs)) != 0) {
        HIPCHK(hipStreamSynchronize(b->stream));
        std::memcpy(h, cur.data(), ns * sizeof(JcgArgs));
        HIPCHK(hipMemcpyAsync(b->dargs + par * ns, h, ns * sizeof(JcgArgs),
                              hipMemcpyHostToDevice, b->stream));
        b->hvalid[par] = true;
      }
      for (size_t gi = 0; gi + 1 < b->gm.size() && !rc; ++gi) {
        const int q0 = b->gm[gi], q1 = b->gm[gi + 1];
        int grid = 0, hmax = -1;
        for (int q = q0; q < q1; ++q) {
          grid += b->m[b->sm[q]]->jcg_grid;
          hmax = std::max(hmax, cur[q].hrows);
        }
        const bool st = cur[q0].hrows >= 0;
        const void* fn = b->gdual[gi] ? (st ? (const void*)k_jcg_batch<true, true>
                                            : (const void*)k_jcg_batch<true, false>)
                                      : (st ? (const void*)k_jcg_batch<false, true>
                                            : (const void*)k_jcg_batch<false, false>);
        const JcgArgs* pa = b->dargs + par * ns + q0;
        const int* pf = b->dfirst + b->gf[gi];
        int nsys = q1 - q0;
        void* args[] = {(void*)&pa, (void*)&pf, (void*)&nsys};
        const hipError_t e = hipLaunchCooperativeKernel(
            fn, dim3(grid), dim3(JCG_BLOCK), args, (unsigned)jcg_stage_bytes(st ? hmax : -1),
            b->stream);
        if (e != hipSuccess) rc = fail(ADCGPU_E_HIP, std::string("batched JCG: ") +
                                                         hipGetErrorString(e));
      }
    }
    for (int k = 0; k < n && !rc; ++k) rc = run_step_finish(b->m[k], g[k]);
    if (!rc && (s & 63) == 63)
      for (int k = 0; k < n && !rc; ++k) {
        bool error = false;
        const int frc = run_read_flags(b->m[k], &error);
        if (error) rc = frc;
      }
  }
  for (int k = 0; k < n && !rc; ++k) rc = restore_coef(b->m[k]);
  for (int k = 0; k < n && !rc; ++k) {
    bool error = false;
    rc = run_read_flags(b->m[k], &error);
  }
  restore();
  return rc;
}

int adcgpu_run_stats(adcgpu_ctx* c, int64_t* out5) {
  if (!c || !out5) return fail(ADCGPU_E_ARG, "null argument");
  out5[0] = c->run_iters;
  out5[1] = c->run_rebuilds;
  out5[2] = c->run_itmax;
  out5[3] = c->run_errstep;
  out5[4] = c->run_mlused;
  return 0;
}

int adcgpu_last_solve(adcgpu_ctx* c, double* out4) {
  if (!c || !out4) return fail(ADCGPU_E_ARG, "null argument");
  double s[12];
  HIPCHK(hipMemcpyAsync(s, c->jsaved, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  JcgArgs a;
  jcg_tolerances(a, c->prm.itmax, c->prm.convcr);
  out4[0] = s[5];
  out4[1] = a.zeta;
  const bool dual = c->prm.warm_start == 2 && c->prm.ilump == 0;
  out4[2] = dual ? s[10] : 0.0;
  out4[3] = dual ? s[11] : 0.0;
  return 0;
}

int adcgpu_set_jcg_blocks(adcgpu_ctx* c, int blocks) {
  if (!c || blocks < 0 || blocks == 1) return fail(ADCGPU_E_ARG, "invalid argument");
  HIPCHK(hipSetDevice(c->device));
  const int old = c->jcg_limit;
  c->jcg_limit = blocks;
  if (int rc = set_jcg_grid(c)) {
    c->jcg_limit = old;
    (void)set_jcg_grid(c);
    return rc;
  }
  return 0;
}

// River radiation's reference elevation ElevDisc: with no flux-settling time
// the first step takes the cold-start elevation (EtaDisc, src/cstart.F:
// 1188-1192, src/timestep.F:917-925); formed from the current ETA2 at the
// flux set-up and at every cold start.
static int set_elevdisc(adcgpu_ctx* c) {
  const int nvel = c->nvel;
  std::vector<double> eta((size_t)c->np), ed((size_t)nvel, 0.0);
  std::vector<int> lbc((size_t)nvel);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(eta.data(), c->eta2, (size_t)c->np * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(lbc.data(), c->lbcodei, (size_t)nvel * sizeof(int), hipMemcpyDeviceToHost));
  for (int k = 0; k < nvel; ++k)
    if (lbc[k] == 52) ed[k] = eta[c->h_nbv[k]];
  HIPCHK(hipMemcpy(c->elevdisc, ed.data(), (size_t)nvel * 8, hipMemcpyHostToDevice));
  return 0;
}

int adcgpu_set_hybrid_friction(adcgpu_ctx* c, double hbreak, double ftheta, double fgamma) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  if (!(hbreak > 0.0) || !(ftheta > 0.0) || !(fgamma >= 0.0) || !std::isfinite(hbreak) ||
      !std::isfinite(ftheta) || !std::isfinite(fgamma))
    return fail(ADCGPU_E_ARG, "HBREAK and FTHETA must be positive, FGAMMA non-negative");
  c->hyb.hbreak = hbreak;
  c->hyb.ftheta = ftheta;
  c->hyb.fgamma = fgamma;
  return 0;
}

int adcgpu_set_flux_boundary(adcgpu_ctx* c, const double* bndlen2o3) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int np = c->np, nvel = c->nvel;
  for (void** p : {(void**)&c->fb_ptr, (void**)&c->fb_ent, (void**)&c->nbv0, (void**)&c->bndlen})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  c->nfluxf = false;
  c->qn_table_steps = 0;
  // the elevation-boundary indices without flux nodes
  HIPCHK(hipMemcpy(c->bnd_index, c->h_bnd_index.data(), (size_t)np * sizeof(int),
                   hipMemcpyHostToDevice));
  if (!bndlen2o3) return 0;
  if (nvel < 2) return fail(ADCGPU_E_ARG, "flux boundaries need at least two flow-boundary nodes");
  for (int k = 0; k < nvel; ++k) {
    if (c->h_nbv[k] < 0 || c->h_nbv[k] >= np) return fail(ADCGPU_E_ARG, "NBV out of range");
    if (!(bndlen2o3[k] >= 0.0 && bndlen2o3[k] <= 1.7976931348623157e308))
      return fail(ADCGPU_E_ARG, "BNDLEN2O3 must be finite and non-negative");
  }
  // each node's edge terms in the reference's edge order (src/gwce.F:1777-1832):
  // edge J adds to NBV(J-1) first, then to NBV(J).  Zero-length edges (after
  // each segment's last node) add exactly +0 and are left out, and so are the
  // elevation-boundary rows, which the boundary condition overwrites.
  std::vector<std::vector<int>> terms(np);
  for (int J = 1; J < nvel; ++J) {
    if (bndlen2o3[J - 1] == 0.0) continue;
    const int a = c->h_nbv[J - 1], b = c->h_nbv[J];
    if (c->h_bnd_index[a] < 0) terms[a].push_back(2 * J);
    if (c->h_bnd_index[b] < 0) terms[b].push_back(2 * J + 1);
  }
  std::vector<int> bidx = c->h_bnd_index, ptr(1, 0), ent;
  for (int i = 0; i < np; ++i) {
    if (terms[i].empty()) continue;
    bidx[i] = -2 - (int)(ptr.size() - 1);
    ent.insert(ent.end(), terms[i].begin(), terms[i].end());
    ptr.push_back((int)ent.size());
  }
  std::vector<double> len(bndlen2o3, bndlen2o3 + nvel);
  if (int rc = upload(&c->fb_ptr, ptr)) return rc;
  if (int rc = upload(&c->fb_ent, ent)) return rc;
  if (int rc = upload(&c->nbv0, c->h_nbv)) return rc;
  if (int rc = upload(&c->bndlen, len)) return rc;
  if (!c->qn0) {
    for (double** p : {&c->qn0, &c->qn1, &c->qn2, &c->qrad, &c->elevdisc}) {
      if (int rc = dmalloc(p, (size_t)nvel)) return rc;
      c->allocs.push_back((void*)*p);
    }
  }
  // QN0 = QN1 = QN2 = 0, as a cold start leaves them (src/cstart.F)
  for (double* p : {c->qn0, c->qn1, c->qn2}) HIPCHK(hipMemset(p, 0, (size_t)nvel * 8));
  HIPCHK(hipMemcpy(c->bnd_index, bidx.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice));
  c->nfluxf = true;
  return set_elevdisc(c);
}

int adcgpu_set_flux_table(adcgpu_ctx* c, int nsteps, const double* qn2_table) {
  if (!c || nsteps < 0 || (nsteps > 0 && !qn2_table)) return fail(ADCGPU_E_ARG, "invalid argument");
  if (!c->nfluxf) return fail(ADCGPU_E_ARG, "no flux boundaries set up (adcgpu_set_flux_boundary)");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->qn_table_cap < nsteps) {
    if (c->qn_table) (void)hipFree(c->qn_table);
    c->qn_table = nullptr;
    c->qn_table_cap = 0;
    if (int rc = dmalloc(&c->qn_table, (size_t)nsteps * c->nvel)) return rc;
    c->qn_table_cap = nsteps;
  }
  if (nsteps > 0)
    HIPCHK(hipMemcpy(c->qn_table, qn2_table, (size_t)nsteps * c->nvel * 8, hipMemcpyHostToDevice));
  c->qn_table_steps = nsteps;
  return 0;
}

int adcgpu_set_tide_potential(adcgpu_ctx* c, int ntif, const double* tpk, const double* amigt,
                              const double* etrf, const double* fft, const double* facet,
                              const double* slam, const double* sfea, int nramp,
                              double dramptip, double statim, double reftim) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  if (ntif < 0 || ntif > TIP_MAX) return fail(ADCGPU_E_ARG, "NTIF out of range");
  if (ntif > 0 && (!tpk || !amigt || !etrf || !fft || !facet || !slam || !sfea))
    return fail(ADCGPU_E_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ntif == 0) {
    c->ctip = false;
    return 0;
  }
  const int np = c->np;
  c->tip_tpk.assign(tpk, tpk + ntif);
  c->tip_amigt.assign(amigt, amigt + ntif);
  c->tip_etrf.assign(etrf, etrf + ntif);
  c->tip_fft.assign(fft, fft + ntif);
  c->tip_facet.assign(facet, facet + ntif);
  c->tip_pert.resize(ntif);
  for (int j = 0; j < ntif; ++j)  // src/read_input.F:3349-3353
    c->tip_pert[j] = amigt[j] == 0.0 ? 0.0 : 2.0 * 3.141592653589793 / amigt[j];
  c->tip_nramp = nramp;
  c->tip_dramp = dramptip;
  c->tip_statim = statim;
  c->tip_reftim = reftim;
  // the species factors (src/adcirc.F:303-307), on the host with the C
  // library's cos and sin like the reference
  std::vector<double> ln(3 * (size_t)np);
  for (int i = 0; i < np; ++i) {
    ln[i] = 1.5 * std::cos(sfea[i]) * std::cos(sfea[i]) - 1.0;
    ln[np + i] = std::sin(2.0 * sfea[i]);
    ln[2 * (size_t)np + i] = std::cos(sfea[i]) * std::cos(sfea[i]);
  }
  if (!c->tip1) {
    for (double** p : {&c->tip1, &c->tip2, &c->tip_slam}) {
      if (int rc = dmalloc(p, (size_t)np)) return rc;
      c->allocs.push_back((void*)*p);
    }
    if (int rc = dmalloc(&c->tip_ln, 3 * (size_t)np)) return rc;
    c->allocs.push_back((void*)c->tip_ln);
  }
  HIPCHK(hipMemcpy(c->tip_ln, ln.data(), ln.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->tip_slam, slam, (size_t)np * 8, hipMemcpyHostToDevice));
  // the cold start's TIP2 (src/cstart.F:358-361)
  HIPCHK(hipMemset(c->tip1, 0, (size_t)np * 8));
  HIPCHK(hipMemset(c->tip2, 0, (size_t)np * 8));
  c->ctip = true;
  return 0;
}

// ---- learned start
int adcgpu_ml_params(adcgpu_ctx* c, const float* in, float* out) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  if (int rc = ml_setup(c)) return rc;
  if (in) {
    HIPCHK(hipMemcpyAsync(c->ml_master, in, (size_t)ML_NPARAM * 4, hipMemcpyHostToDevice,
                          c->stream));
    // a new starting point for training: the Adam moments restart
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ML_NPARAM, 256)), dim3(256), 0, c->stream,
                       reinterpret_cast<int*>(c->ml_m1), 0, ML_NPARAM);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ML_NPARAM, 256)), dim3(256), 0, c->stream,
                       reinterpret_cast<int*>(c->ml_m2), 0, ML_NPARAM);
    c->ml_step = 0;
    hipLaunchKernelGGL(k_ml_pack, dim3(nblk(ML_W2 * 2, 256)), dim3(256), 0, c->stream,
                       (const float*)c->ml_master, c->mlw);
  }
  if (out)
    HIPCHK(hipMemcpyAsync(out, c->ml_master, (size_t)ML_NPARAM * 4, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_launch("ml params");
}

int adcgpu_ml_collect(adcgpu_ctx* c, int per_step) {
  if (!c || per_step < 0 || per_step > ML_CAP) return fail(ADCGPU_E_ARG, "invalid argument");
  if (c->prm.ilump != 0) return fail(ADCGPU_E_UNSUPPORTED, "training samples need ILump = 0");
  if (int rc = ml_setup(c)) return rc;
  c->ml_collect = per_step;
  return 0;
}

int64_t adcgpu_ml_samples(adcgpu_ctx* c) {
  if (!c) return -1;
  return std::min<int64_t>(c->smp_total, c->ml_ready ? c->smp.cap : 0);
}

int adcgpu_ml_train(adcgpu_ctx* c, int epochs, double lr, uint64_t seed, double* loss_out) {
  if (!c || epochs < 0 || !(lr > 0.0)) return fail(ADCGPU_E_ARG, "invalid argument");
  if (int rc = ml_setup(c)) return rc;
  const int ns = (int)std::min<int64_t>(c->smp_total, c->smp.cap);
  if (ns <= 0) return fail(ADCGPU_E_ARG, "no training samples collected");
  hipLaunchKernelGGL(k_ml_weights, dim3(1), dim3(1024), 0, c->stream, c->smp, ns, c->twt,
                     c->tmeanw);
  const int nb = (ns + ML_BATCH - 1) / ML_BATCH;
  std::vector<int> perm((size_t)nb * ML_BATCH, -1);
  uint64_t st = seed ^ 0x243f6a8885a308d3ull;
  auto next = [&]() {  // splitmix64
    st += 0x9e3779b97f4a7c15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  float* lossacc = c->tloss + ML_TRAIN_BLOCKS;
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int k = 0; k < ns; ++k) perm[k] = k;
    for (int k = ns - 1; k > 0; --k) std::swap(perm[k], perm[next() % (uint64_t)(k + 1)]);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream,
                       reinterpret_cast<int*>(lossacc), 0, 1);
    for (int b = 0; b < nb; ++b) {
      HIPCHK(hipMemcpyAsync(c->tperm, perm.data() + (size_t)b * ML_BATCH, ML_BATCH * 4,
                            hipMemcpyHostToDevice, c->stream));
      if (int rc = ml_minibatch(c, 1.0f / ML_BATCH, nullptr, true, (float)lr, lossacc)) return rc;
    }
    float l = 0.0f;
    HIPCHK(hipMemcpyAsync(&l, lossacc, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    last = (double)l / nb;
  }
  if (loss_out) *loss_out = last;
  return 0;
}

int adcgpu_ml_grad(adcgpu_ctx* c, int n, const float* x, const float* y, const float* w,
                   float* grad, double* loss) {
  if (!c || n <= 0 || n > ML_BATCH || !x || !y || !w || !grad)
    return fail(ADCGPU_E_ARG, "invalid argument");
  if (int rc = ml_setup(c)) return rc;
  std::vector<uint16_t> xb((size_t)n * ML_F);
  for (size_t k = 0; k < xb.size(); ++k) xb[k] = f32_to_bf16(x[k]);
  std::vector<int> node(n, 0), perm(ML_BATCH, -1);
  for (int k = 0; k < n; ++k) perm[k] = k;
  HIPCHK(hipMemcpyAsync(c->smp.x, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->smp.y, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->smp.node, node.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->twt, w, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->tperm, perm.data(), ML_BATCH * 4, hipMemcpyHostToDevice, c->stream));
  float* lossacc = c->tloss + ML_TRAIN_BLOCKS;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<int*>(lossacc),
                     0, 1);
  float* gdev = c->tda1;  // free once the minibatch's dZ1 pass has read it
  if (int rc = ml_minibatch(c, 1.0f / n, gdev, false, 0.0f, lossacc)) return rc;
  float l = 0.0f;
  HIPCHK(hipMemcpyAsync(grad, gdev, (size_t)ML_NPARAM * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&l, lossacc, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (loss) *loss = l;
  return 0;
}

int adcgpu_ml_features(adcgpu_ctx* c, float* x_out, double* s_out) {
  if (!c || !x_out) return fail(ADCGPU_E_ARG, "null argument");
  if (int rc = ml_setup(c)) return rc;
  if (c->np > c->smp.cap) return fail(ADCGPU_E_CAPACITY, "more nodes than the sample ring holds");
  // the state at the start of a step's solve: eta^k is ETA2 between steps
  const MlIn in = ml_in(c, c->eta2);
  hipLaunchKernelGGL(k_ml_collect, dim3(nblk(c->np, 128)), dim3(256), 0, c->stream, in, c->smp,
                     0, c->np, -1);
  std::vector<uint16_t> xb((size_t)c->np * ML_F);
  std::vector<float> s(c->np);
  HIPCHK(hipMemcpyAsync(xb.data(), c->smp.x, xb.size() * 2, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(s.data(), c->smp.s, (size_t)c->np * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < xb.size(); ++k) x_out[k] = bf16_to_f32(xb[k]);
  if (s_out)
    for (int i = 0; i < c->np; ++i) s_out[i] = s[i];
  return check_launch("ml features");
}

int adcgpu_ml_correction(adcgpu_ctx* c, double* corr_out) {
  if (!c || !corr_out) return fail(ADCGPU_E_ARG, "null argument");
  if (int rc = ml_setup(c)) return rc;
  const MlIn in = ml_in(c, c->eta2);
  ml_infer(c, in);
  HIPCHK(hipMemcpyAsync(corr_out, c->ml_corr, (size_t)c->np * 8, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_launch("ml correction");
}

// ---- whole-state snapshot (between steps)
extern "C++" {
namespace {
struct StatePart {
  void* dev;
  size_t bytes;
};
std::vector<StatePart> state_parts(adcgpu_ctx* c) {
  const size_t np = c->np, ne = c->ne, mn = (size_t)c->mnei * np;
  std::vector<StatePart> v;
  double* nodal[] = {c->eta1, c->eta2, c->etas, c->uu1, c->vv1, c->uu2, c->vv2, c->qx1, c->qy1,
                     c->qx2, c->qy2, c->h1, c->h2, c->tk, c->etamax, c->totalarea, c->wsx1,
                     c->wsy1, c->wsx2, c->wsy2, c->pr1, c->pr2, c->gwce_lv, c->etas_prev,
                     c->umax};
  for (double* p : nodal) v.push_back({p, np * 8});
  int* inodal[] = {c->nodecode, c->nnodecode, c->mju, c->etamax_it, c->umax_it};
  for (int* p : inodal) v.push_back({p, np * 4});
  v.push_back({c->noff, ne * 4});
  v.push_back({c->noffold, ne * 4});
  v.push_back({c->alphal, ne * 8});
  v.push_back({c->coef, mn * 8});
  v.push_back({c->obc, (size_t)std::max(c->neta, 1) * c->mnei * 8});
  v.push_back({c->scal, S_COUNT * 8});
  v.push_back({c->jsaved, 12 * 8});
  v.push_back({c->flags, F_COUNT * 4});
  if (c->ctip) {
    v.push_back({c->tip1, np * 8});
    v.push_back({c->tip2, np * 8});
  }
  if (c->nfluxf)
    for (double* p : {c->qn0, c->qn1, c->qn2, c->elevdisc}) v.push_back({p, (size_t)c->nvel * 8});
  return v;
}
}  // namespace
}  // extern "C++"

int64_t adcgpu_state_bytes(adcgpu_ctx* c) {
  if (!c) return -1;
  int64_t n = 16;  // header: NP, NE, MNEI, the step of the snapshot
  for (const StatePart& p : state_parts(c)) n += (int64_t)p.bytes;
  return n;
}

int adcgpu_state_save(adcgpu_ctx* c, void* buf, int64_t bytes) {
  if (!c || !buf) return fail(ADCGPU_E_ARG, "null argument");
  if (bytes != adcgpu_state_bytes(c)) return fail(ADCGPU_E_ARG, "snapshot size");
  HIPCHK(hipSetDevice(c->device));
  // the off-diagonals a deferred solve left scaled are restored first (the
  // same operations the next matrix pass would apply)
  if (int rc = restore_coef(c)) return rc;
  char* o = static_cast<char*>(buf);
  const int32_t hdr[4] = {c->np, c->ne, c->mnei, c->cur_it};
  std::memcpy(o, hdr, 16);
  o += 16;
  for (const StatePart& p : state_parts(c)) {
    HIPCHK(hipMemcpyAsync(o, p.dev, p.bytes, hipMemcpyDeviceToHost, c->stream));
    o += p.bytes;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int adcgpu_state_load(adcgpu_ctx* c, const void* buf, int64_t bytes) {
  if (!c || !buf) return fail(ADCGPU_E_ARG, "null argument");
  if (bytes != adcgpu_state_bytes(c)) return fail(ADCGPU_E_ARG, "snapshot size");
  const char* o = static_cast<const char*>(buf);
  int32_t hdr[4];
  std::memcpy(hdr, o, 16);
  if (hdr[0] != c->np || hdr[1] != c->ne || hdr[2] != c->mnei)
    return fail(ADCGPU_E_ARG, "snapshot of another mesh");
  o += 16;
  HIPCHK(hipSetDevice(c->device));
  for (const StatePart& p : state_parts(c)) {
    HIPCHK(hipMemcpyAsync(p.dev, o, p.bytes, hipMemcpyHostToDevice, c->stream));
    o += p.bytes;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->cur_it = hdr[3];
  c->coef_scaled = false;
  c->ncele_fresh = false;
  return 0;
}

int adcgpu_set_wait_limit(double seconds) {
  if (!(seconds >= 0.0 && seconds <= 3600.0)) return fail(ADCGPU_E_ARG, "wait limit");
  g_wait_seconds = seconds;
  return 0;
}

int64_t adcgpu_run_iterations(adcgpu_ctx* c) { return c ? c->run_iters : -1; }

int64_t adcgpu_run_rebuilds(adcgpu_ctx* c) { return c ? c->run_rebuilds : -1; }

int adcgpu_jcg_shape(adcgpu_ctx* c, int* shape3) {
  if (!c || !shape3) return fail(ADCGPU_E_ARG, "null argument");
  shape3[0] = c->jcg_grid;
  shape3[1] = c->jcg_cap;
  shape3[2] = c->nch;
  return 0;
}

int adcgpu_profile(adcgpu_ctx* c, int enable) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ev_used = 0;
  c->profile = enable != 0;
  if (enable) return grow_events(c, std::max<size_t>(c->jcg_events.size(),
                                                     (size_t)std::max(enable, 1024)));
  return 0;
}

int adcgpu_profile_read(adcgpu_ctx* c, double* jcg_ms, int64_t* launches) {
  if (!c) return fail(ADCGPU_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  for (size_t k = 0; k < c->ev_used; ++k) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->jcg_events[k].first, c->jcg_events[k].second));
    tot += ms;
  }
  if (jcg_ms) *jcg_ms = tot;
  if (launches) *launches = (int64_t)c->ev_used;
  return 0;
}

// Periodic elevation forcing at the NBD nodes (src/gwce.F:1638-1651), on the
// host with the C library's cos like the reference.
int adcgpu_elevation_bc(int neta, int nbfr, const double* amig, const double* ff,
                        const double* face, const double* emo, const double* efa, double timeh,
                        double rampelev, double* out) {
  if (neta < 0 || nbfr < 0 || (neta > 0 && !out)) return fail(ADCGPU_E_ARG, "invalid argument");
  const double pi = 3.141592653589793;
  for (int I = 0; I < neta; ++I) out[I] = 0.0;
  for (int J = 0; J < nbfr; ++J) {
    const double per = amig[J] == 0.0 ? 0.0 : 2.0 * pi / amig[J];
    int ncyc = 0;
    if (per != 0.0) ncyc = (int)(timeh / per);
    const double argj = amig[J] * (timeh - ncyc * per) + face[J];
    const double rff = ff[J] * rampelev;
    for (int I = 0; I < neta; ++I) {
      const double arg = argj - efa[(size_t)I * nbfr + J];
      out[I] = out[I] + emo[(size_t)I * nbfr + J] * rff * std::cos(arg);
    }
  }
  return 0;
}

// ---------------------------------------------------------------- drop-in JCG
int adcgpu_flux_bc(int nvel, const int32_t* lbcodei, int nffr, const double* famig,
                   const double* fff, const double* fface, const double* qnam,
                   const double* qnph, double timeh, double rampextflux, double* out) {
  if (nvel < 0 || nffr < 0 || !out || (nvel > 0 && !lbcodei) ||
      (nffr > 0 && (!famig || !fff || !fface || !qnam || !qnph)))
    return fail(ADCGPU_E_ARG, "invalid argument");
  for (int i = 0; i < nvel; ++i) out[i] = 0.0;
  for (int j = 0; j < nffr; ++j) {
    // FPER (src/read_input.F:3551-3555); the phase reduced by whole periods
    const double fper = famig[j] == 0.0 ? 0.0 : 2.0 * 3.141592653589793 / famig[j];
    const int ncyc = fper == 0.0 ? 0 : (int)(timeh / fper);
    const double argj = famig[j] * (timeh - ncyc * fper) + fface[j];
    const double rff = fff[j] * rampextflux;
    for (int i = 0; i < nvel; ++i) {
      const int code = lbcodei[i];
      if (code == 2 || code == 12 || code == 22 || code == 52) {
        const size_t q = (size_t)i * nffr + j;
        out[i] = out[i] + qnam[q] * rff * std::cos(argj - qnph[q]);
      }
    }
  }
  return 0;
}

int adcgpu_jcg(int n, int ndim, int maxnz, int32_t* jcoef, double* coef, double* rhs, double* u,
               int32_t* iparm, double* rparm) {
  if (n <= 0) return 11;
  if (!jcoef || !coef || !rhs || !u || !iparm || !rparm || ndim < n || maxnz < 1)
    return fail(ADCGPU_E_ARG, "invalid argument");
  if (iparm[0] <= 0) return 0;
  if (iparm[5] != 1 || iparm[8] != -2 || iparm[9] != 0)
    return fail(ADCGPU_E_UNSUPPORTED, "only adaptive JCG without red-black or row removal");
  const int itmax = iparm[0];
  if (itmax > JCG_MAX_ITMAX) return fail(ADCGPU_E_ARG, "ITMAX out of range");
  if ((size_t)maxnz * (size_t)n * sizeof(double) >= ((size_t)1 << 32))
    return fail(ADCGPU_E_CAPACITY, "system too large for 32-bit offsets");
  // compact the ELLPACK rows: diagonal first (scal moves it there,
  // src/itpackv.F:2331-2356), padding (jcoef <= 0) dropped
  int width = 1;
  std::vector<int> nn(n, 0);
  std::vector<std::vector<int>> pos(n);
  for (int i = 0; i < n; ++i) {
    int dj = -1;
    for (int j = 0; j < maxnz; ++j)
      if (jcoef[i + (size_t)ndim * j] == i + 1) dj = j;
    if (dj < 0) return 402;
    pos[i].push_back(dj);
    for (int j = 0; j < maxnz; ++j) {
      int v = jcoef[i + (size_t)ndim * j];
      if (j != dj && v > 0) pos[i].push_back(j);
      if (v > n) return fail(ADCGPU_E_ARG, "jcoef out of range");
    }
    nn[i] = (int)pos[i].size();
    width = std::max(width, nn[i]);
    double d = coef[i + (size_t)ndim * dj];
    if (d == 0.0) return 401;
    if (d < 0.0) return fail(ADCGPU_E_UNSUPPORTED, "negative diagonal (nonsymmetric scaling)");
  }
  std::vector<double> hc((size_t)width * n, 0.0);
  std::vector<int> hcol((size_t)width * n, 0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < nn[i]; ++k) {
      hc[(size_t)k * n + i] = coef[i + (size_t)ndim * pos[i][k]];
      hcol[(size_t)k * n + i] = jcoef[i + (size_t)ndim * pos[i][k]] - 1;
    }
  // one persistent scratch context per device, holding ITPACK's saved state
  static adcgpu_ctx* jc[64] = {nullptr};
  static int jc_n[64] = {0}, jc_w[64] = {0};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(ADCGPU_E_ARG, "device ordinal");
  adcgpu_ctx*& c = jc[dev];
  if (!c || jc_n[dev] < n || jc_w[dev] < width) {
    double saved[8] = {0};
    if (c) {
      (void)hipMemcpy(saved, c->jsaved, sizeof(saved), hipMemcpyDeviceToHost);
      for (void* p : c->allocs) (void)hipFree(p);
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
    }
    c = new adcgpu_ctx();
    c->device = dev;
    c->np = n;
    c->mnei = width;
    c->nch = (n + RED_CHUNK - 1) / RED_CHUNK;
    int rc = 0;
    auto A = [&](auto** p, size_t cnt) {
      if (!rc) rc = dmalloc(p, cnt);
      if (!rc) c->allocs.push_back((void*)*p);
    };
    A(&c->coef, (size_t)width * n); A(&c->cs, (size_t)width * n); A(&c->col, (size_t)width * n);
    A(&c->nneigh, n); A(&c->gwce_lv, n); A(&c->etas, n); A(&c->jd, n); A(&c->jw, n);
    A(&c->jpart, 2 * (size_t)c->nch * 4); A(&c->jsaved, 12); A(&c->jbar, JCG_BAR_WORDS);
    A(&c->flags, F_COUNT);
    if (!rc) HIPCHK(hipStreamCreate(&c->stream));
    if (!rc) rc = setup_jcg_grid(c);
    if (rc) {
      jc[dev] = nullptr;
      return rc;
    }
    HIPCHK(hipMemcpy(c->jsaved, saved, sizeof(saved), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->flags, 0, F_COUNT * sizeof(int)));
    jc_n[dev] = n;
    jc_w[dev] = width;
  }
  c->np = n;
  c->nch = (n + RED_CHUNK - 1) / RED_CHUNK;
  if (int rc = set_jcg_grid(c)) return rc;
  HIPCHK(hipMemcpyAsync(c->coef, hc.data(), hc.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->col, hcol.data(), hcol.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->nneigh, nn.data(), n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->gwce_lv, rhs, n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->etas, u, n * 8, hipMemcpyHostToDevice, c->stream));
  // this call's outcome only: a failed earlier call must not leave its
  // error, iteration count or ier behind
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->flags, 0, F_COUNT);
  // the device matrix is stored with leading dimension n
  int rc = launch_jcg(c, c->coef, c->col, c->nneigh, c->gwce_lv, c->etas, n, itmax, rparm[0],
                      nullptr, nullptr, nullptr, false, false);
  if (rc) return rc;
  {
    int fe[F_COUNT];
    HIPCHK(hipMemcpyAsync(fe, c->flags, sizeof(fe), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // the device copies are left as the failed launch left them; the
    // caller's arrays are untouched
    if (int erc = flag_errors(fe)) return erc;
  }
  double saved[8];
  int f[F_COUNT];
  HIPCHK(hipMemcpyAsync(hc.data(), c->coef, hc.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(rhs, c->gwce_lv, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(u, c->etas, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(saved, c->jsaved, sizeof(saved), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(f, c->flags, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < nn[i]; ++k) coef[i + (size_t)ndim * pos[i][k]] = hc[(size_t)k * n + i];
  const int in = f[F_NUMITR];
  const int ier = f[F_IER];
  iparm[7] = 4 * n + 4 * in;
  if (ier == 13 && iparm[2] == 0) rparm[0] = saved[5];
  if (iparm[2] == 0) {
    iparm[0] = in;
    iparm[8] = -2;
    rparm[1] = saved[4];
    rparm[2] = rparm[2] == 0.0 ? -1.0 : rparm[2];
    rparm[10] = saved[5] > 0.0 ? -std::log10(std::fabs(saved[5])) : -std::log10(std::ldexp(1.0, -52));
  }
  return ier;
}

}  // extern "C"
