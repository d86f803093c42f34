// Device-side data layout shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace adcgpu {

// fixed blocked reduction: 2048-row chunks of 512 threads x 4 rows; chunk
// values combined by 256 threads
constexpr int RED_THREADS = 512;
constexpr int RED_PER_THREAD = 4;
constexpr int RED_CHUNK = RED_THREADS * RED_PER_THREAD;
constexpr int RED_XTHREADS = 256;

// persistent JCG launch: one block per chunk plus one eigenvalue block
constexpr int JCG_CPB = 2;                      // 2048-row chunks per block
constexpr int JCG_BLOCK = RED_THREADS * JCG_CPB;  // threads per block
constexpr int JCG_BAR_GROUP = 16;        // blocks per first-level barrier counter
constexpr int JCG_BAR_STRIDE = 64;       // words between barrier counters (256 B)
constexpr int JCG_BAR_WORDS = 10240;     // barrier words: (groups + 2) * JCG_BAR_STRIDE
                                          // + blocks * JCG_HALO_STRIDE
constexpr int JCG_HALO_STRIDE = 16;      // words between per-tile halo flags (64 B)
constexpr int JCG_HALO_MAX = 2;          // widest halo (tiles each side) waited on by flags
constexpr int JCG_MAX_ITMAX = 256;       // tridiagonal history length (itmax)
// dynamic LDS for the staged pseudo-residual of a k_jcg block (JcgArgs.hrows),
// at most: the block's rows and 5,632 rows each side
constexpr int JCG_STAGE_MAX_BYTES = 120 * 1024;
static_assert(JCG_STAGE_MAX_BYTES / 8 <= 65536, "staged positions are 16-bit");
constexpr int NODE_STAGE = 2048;   // incidence slots a 256-node block stages in LDS
constexpr int LHS_MAXW = 32;         // widest COEF row (MNEI): LDS of the row kernel
// ordered wetting updates: candidates are closed in blocks of WD_BLOCK, each
// block naming at most WD_XMAX earlier candidates it depends on
constexpr int WD_BLOCK = 1024;
constexpr int WD_XMAX = 256;                 // four 64-bit mask words per candidate
constexpr int WD_XW = WD_XMAX / 64;
constexpr int WD_HASH = 2 * WD_XMAX;         // LDS hash slots for a block's external names
constexpr int WD_INC_MAX = 12;               // incidence lists a wetting candidate scans at once
constexpr int WD_CLOSE_LDS = 2 * WD_BLOCK * (8 + 8 * WD_XW + 4) + 2 * WD_HASH * 4;
constexpr int WD_LINK_BLOCKS = 64;           // blocks whose link values fit the LDS of k_wd_link
// serial fallback of k_wd_link: candidates staged through LDS in chunks, the
// running maxima of the most recent WD_RING candidates kept in an LDS ring
constexpr int WD_CHUNK = 2048;
constexpr int WD_RING = 8192;   // power of two
constexpr int WD_SERIAL_LDS = WD_RING * 8 + WD_CHUNK * (8 + 8 + 8);
constexpr int WD_LINK_LDS = WD_LINK_BLOCKS * WD_XMAX * 8 > WD_SERIAL_LDS
                                ? WD_LINK_BLOCKS * WD_XMAX * 8 : WD_SERIAL_LDS;

// device flag slots
enum { F_NCCHANGE = 0, F_NUMITR = 2, F_IER, F_ERROR, F_NCAND, F_ITSUM, F_NREBUILD, F_WDSERIAL,
       F_ERRSTEP,  // first step that raised ERR_ELEV or ERR_NONFINITE (0: none)
       F_NITMAX,   // solves that reached ITMAX (ier = 13)
       F_ALPHA,    // some element's ALPHAL differs from 1 (the cached matrix
                   // terms assume ALPHAL = 1; the element pass recomputes them)
       F_MLUSED,   // solves that started from the learned guess
       F_COUNT = 16 };
enum { ERR_DIAG = 1, ERR_BARRIER = 4, ERR_ELEV = 8, ERR_NONFINITE = 16 };
// device scalar slots
enum { S_EP = 0, S_COUNT = 8 };

struct DevMesh {
  int np, ne, mnei, neta;
  double dt, g, a00, b00, c00;
  const int* nm;        // [ne*3] 0-based
  const double* fdx;    // [ne*3]
  const double* fdy;    // [ne*3]
  const double* areas;  // [ne]
  const double* x;
  const double* y;
  const double* dp;
  const double* tau0;
  const double* evm;
  const double* corif;
  const int* inc_ptr;   // [np+1] node -> incident (element, local) codes 3e+k
  const int* inc;       //        sorted by element id
  const int* slot;      // [ne*3] position p of code 3e+k in inc (inc[slot[c]] == c)
  const int* slot_cols; // [ne*3] per slot p: CSR_POSITION slots ja | jb << 16 (0-based)
  const int* nneigh;    // [np]
  const int* col;       // [mnei*np] column-major, 0-based (slot 0 = self)
  const int* csrpos;    // [ne*3*2] 0-based column slots
  const int* bnd_index; // [np] elevation-boundary index, -1, or -2-k for the
                        //      k-th node that takes specified-flux edge terms
  const int* nbd;       // [neta] 0-based
  const int* obc_q;     // [nobc] COEF positions j*np+i that couple row i to an
  const int* obc_s;     //        elevation-boundary node, and their OBCCOEF slot
                        //        I*mnei+(J-2); -1-slot when row i is itself a boundary row
  int nobc;
  const int* obcg_ptr;  // [np+1] node -> boundary-column slots, (I,J) order
  const int* obcg;
  const int* vbc_ptr;   // [np+1] node -> flow-boundary entries in ME2GW order
  const int* vbc;
  const int* lbcodei;
  const double* csii;
  const double* siii;
  // specified normal flux boundaries (null without): per flux node k its
  // edge terms fb_ent[fb_ptr[k] .. fb_ptr[k+1]) in the reference's edge
  // order, each 2*J + r for the edge NBV(J-1)-NBV(J) (0-based J >= 1) with
  // r = 0 when the node is NBV(J-1), 1 when it is NBV(J)
  const int* fb_ptr;
  const int* fb_ent;
  const int* nbv;       // [nvel] 0-based
  const double* bndlen; // [nvel] BNDLEN2O3
  // mixed precision: FP32 copies of the element geometry (null in FP64 runs)
  const float* fdx32;
  const float* fdy32;
  const float* areas32;
};

struct DevState {
  double *eta1, *eta2, *etas, *uu1, *vv1, *uu2, *vv2, *qx1, *qy1, *qx2, *qy2;
  double *h1, *h2, *tk, *etamax, *totalarea, *wsx1, *wsy1, *wsx2, *wsy2, *pr1, *pr2;
  double *gwce_lv, *alphal, *coef, *obc, *fric;
  int *nodecode, *nnodecode, *noff, *noffold, *mju;
  int *etamax_it;  // step of the last ETAMAX rise (0: never)
  double *umax;    // peak depth-averaged speed at wet nodes
  int *umax_it;    // step of the last UMAX rise (0: never)
  unsigned char* ncele;  // [ne] NCEle of each element for the next GWCE matrix
  double *tip1, *tip2;   // tidal potential at time levels s and s+1 (null: none)
  double *qn0, *qn1, *qn2;  // normal flux per flow-boundary entry at three time
                            // levels (null: no flux boundaries)
  double* qrad;             // per entry: the step's QFORCE of an outward-radiating
                            // entry (LBCODEI 30), or the radiation part of a
                            // river-radiation entry's (52), formed at the step's start
  double* elevdisc;         // per entry: river radiation's reference elevation
};

__global__ void k_flux_shift(DevMesh m, DevState s, const double* next, int nvel);

// The tidal potential of one step (src/timestep.F:1527-1557, NTIP = 1), per
// constituent: its species, ARGT (the phase at the step, reduced by whole
// periods) and TPMUL = RampTip*ETRF*TPK*FFT, formed on the host.
constexpr int TIP_MAX = 32;
struct TipPrm {
  int ntif;
  int na[TIP_MAX];
  double argt[TIP_MAX], tpmul[TIP_MAX];
};

// The GWCE matrix terms of every (element, vertex) slot that do not change
// from step to step (geometry, depth, Tau0, ALPHAL = 1), kept in slot order;
// a rebuild multiplies them by the element's NCEle.
struct LhsTerms {
  const double *td, *ta, *tb;  // [3 ne] diagonal and the two off-diagonal terms
  const float *td32, *ta32, *tb32;  // the same rounded to FP32 (mixed precision)
  int pairs;  // 1: a cached rebuild is k_lhs_pairs' and k_lhs_ep's (LhsPairs)
};

// The same terms gathered per off-diagonal entry: entry (i, j >= 1) of COEF
// receives, in ascending element order, the terms of the (at most two)
// elements that hold the edge i-j, so a rebuild is
// COEF(i, j) = (0 + t1 NCEle(e1)) + t2 NCEle(e2), one thread per entry with
// no accumulation in between.  Index (j - 1) NP + i; an absent second term is
// a zero term of element 0.
struct LhsPairs {
  const double *t1, *t2;
  const float *t1f, *t2f;  // mixed precision: the FP32 terms
  const int *e1, *e2;
};

// Compensated (Kahan 1965) summation in FP32: the node sums of FP32
// element contributions in the mixed-precision path.  The result carries the
// compensation into FP64.
struct Kahan32 {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

struct RhsPrm {
  double ifnlfa, ifnlct, ifnlcat;
  double ondiag, offdiag;  // (1+ILump)*2 and 1-ILump (src/gwce.F:419-420)
  int lumped;              // ILump = 1: no ETA1 gradient term (src/gwce.F:1233-1238)
  int nws;
};

struct MomPrm {
  double ifnlct;
  int nws;
  int it;            // this step, recorded where ETAMAX rises
  double errorelev;  // ErrorElev: |ETA2| above it at a wet node stops the run
  int* flags;        // run-time error flags (ERR_ELEV, ERR_NONFINITE, F_ERRSTEP)
};

// Hybrid bottom friction (NOLIBF = 2): IFHYBF and fort.15's HBREAK, FTHETA,
// FGAMMA; `on` is 0 for the other friction laws.
struct Hybrid {
  double on, hbreak, ftheta, fgamma;
};

// IFHYBF*(1 + (HBREAK/H)**FTHETA)**(FGAMMA/FTHETA) (src/nodalattr.F:2406-2410,
// src/wetdry.F:415-418, 447-451): the depth factor of the hybrid law, zero
// without it.  The device's pow differs from the C library's in the last bit
// for some arguments.
__device__ __forceinline__ double hybrid_term(const Hybrid& h, double H) {
  return h.on != 0.0 ? h.on * pow(1.0 + pow(h.hbreak / H, h.ftheta), h.fgamma / h.ftheta) : 0.0;
}

struct WdPrm {
  double hoff, velmin, iflinbf, ifnlbf, tkmax;
  Hybrid hyb;
};

// Wetting candidates of one step (src/wetdry.F:368-475 and its two twins).
// Candidate k is the k-th candidate element in element order; every array is
// sized for the worst case of one candidate per element.
struct WdWork {
  unsigned long long* mask;  // [ceil(ne/64)] bit e%64 of word e/64: element e is a candidate
  int* bcnt;                 // [ceil(ne/256)] candidates per 256-element tile
  int* bofs;                 // [ceil(ne/256)] candidates in earlier tiles
  double* tkc;               // [ne] TK_TMP(NM123) of candidate element e
  int* lc;                   // [ne] local vertex of NM123 | local vertex of the partner << 2
  double* m0;                // [ne] candidate k: TK_TMP and the entry TK of nodes it opens
  int2* pred;                // [ne] candidate k: previous candidate on NM123 / partner, or -1
  int2* wnode;               // [ne] candidate k: NM123 / partner if k is its last candidate, or -1
  double* M;                 // [ne] candidate k: TK of both nodes after k (serial fallback)
  // block closure (k_wd_close): candidate k's value is
  // max(cv[k], X values of the bits of cE[k]), the X values being those of the
  // earlier candidates xtab[b*WD_XMAX + i] of its block b = k / WD_BLOCK
  double* cv;                // [ne] maximum over k's ancestors inside its block
  unsigned long long* cE;    // [ne*WD_XW] external names reached inside the block
  int* xtab;                 // [nblk*WD_XMAX] block b's external names (candidate indices)
  int* xcnt;                 // [nblk] number of external names of block b
  double* xval;              // [nblk*WD_XMAX] final value of each external name
};

struct HollandPrm {
  double v[13];
  double t;
};

struct FricPrm {
  const double *uu1, *vv1, *dp, *eta2, *fric;
  double ifnlfa, iflinbf, ifnlbf, tkmax;
  Hybrid hyb;
};

struct JcgArgs {
  int n, itmax, maxfn;
  double zeta, tz;
  double* coef;       // [mnei*n] unscaled in/out (diag in slot 0)
  double* cs;         // [mnei*n] scaled off-diagonals
  const int* col;
  const int* nneigh;
  double* rhs;        // scaled in place, restored on exit
  double* u;          // initial guess in (scaled in place), ETAS out
  double* d;          // pseudo-residual of the current iteration
  double* w;          // 1/sqrt(diag)
  double* part;       // [2][nch*4] chunk partials (d.d, d.Nd, u.u), by iteration parity
  double* saved;      // [12] ITPACK SAVE state q1, gamold, rhoold, cmold, cme,
                      // stptst; [6..9] = cme, cmold of the running iteration, by parity
  int prescaled;      // rows already scaled (w, u, rhs) by k_rhs_node
  int defer_unscal;   // leave the off-diagonals scaled (k_lhs_rows restores them)
  unsigned* barrier;  // [JCG_BAR_WORDS] grid barrier counters, then per-tile halo flags
  int hrows;          // >= 0: the pseudo-residual of the tile's rows and of hrows
                      // rows each side (at least the matrix bandwidth) is kept
                      // in LDS for the products (dynamic shared memory of
                      // 8 (JCG_CPB RED_CHUNK + 2 hrows) bytes); -1: gathered
                      // from memory
  const int* hrange;  // staged: per row tile, the first and one past the last row
                      // its rows reference (within hrows of the tile)
  const unsigned* lcol;  // staged: [(mnei/2)][n] per row, off-diagonal slots 2p+1
                         // and 2p+2 as positions in the tile's staged rows, 16
                         // bits each (low half first)
  int halo;           // row tiles each side whose pseudo-residual a tile's rows read;
                      // -1: wait for all tiles (a grid barrier) after each update
  long long barrier_ticks;  // longest grid-barrier wait, in wall-clock ticks
  int* flags;
  double* eta2;       // optional: ETA2 = NODECODE*(ETAS+0)+ETA1
  const double* eta1;
  const int* nodecode;
  double* uo;         // k_jcg_mem: [n] the previous u
  double* dpv;        //             [n] the previous d
  double* ndv;        //             [n] N d
  const double* u2;   // k_jcg<true>: the second (learned) start, scaled like u
  double* d2;         //              and its initial pseudo-residual
};

__global__ void k_init_h(int np, const double* dp, const double* eta2, double ifnlfa, double* h1,
                         double* h2);
// TIP2 = sum over constituents of TPMUL*L_N(NA, i)*cos(ARGT + NA*SLAM(i)); ln is
// [3][np] (long-period, diurnal, semidiurnal species)
__global__ void k_tide_pot(int np, TipPrm p, const double* ln, const double* slam, double* tip2);
__global__ void k_friction(int np, FricPrm f, double* tk);
__global__ void k_holland(int np, const double* x, const double* y, HollandPrm p, double* wsx,
                          double* wsy, double* pr);
__global__ void k_friction_holland(int np, FricPrm f, double* tk, const double* x, const double* y,
                                   HollandPrm p, double* wsx, double* wsy, double* pr);
// LUMPED: the lumped GWCE (ILump = 1), whose LHS is the diagonal COEFD only
// (kept in COEF's diagonal column) and whose solve is explicit
template <bool LUMPED>
__global__ void k_lhs_elem(DevMesh m, DevState s, const int* flags, double* sd, double* sa,
                           double* sb);
// CT: the type of the matrix terms summed per row (double; float in the
// mixed path, which sums FP32 terms with compensation)
template <bool LUMPED, typename CT>
__global__ void k_lhs_rows(DevMesh m, DevState s, const int* flags, const double* sd,
                           const double* sa, const double* sb, LhsTerms lt, double* ep_part,
                           const double* restore_cs, const double* restore_w);
// the cached terms (once per context): k_lhs_elem's terms with NCEle = 1 and
// ALPHAL = 1, and their FP32 roundings when td32 != null
template <bool LUMPED>
__global__ void k_lhs_terms(DevMesh m, double* td, double* ta, double* tb, float* td32,
                            float* ta32, float* tb32);
// NCEle of every element from the current NODECODE and NOFF
__global__ void k_ncele(DevMesh m, DevState s);
// per-entry terms (once per context) from the cached terms and the host's
// entry codes (2 p + 0/1 for slot p's ta/tb term, -1 for none)
__global__ void k_lhs_pair_setup(DevMesh m, LhsTerms lt, const int* c1, const int* c2,
                                 double* t1, double* t2, float* t1f, float* t2f, int* e1,
                                 int* e2);
// A rebuild of the consistent GWCE matrix from the cached terms (the common
// case: NCChange set, every ALPHAL 1): grid (row blocks, MNEI), column j of
// COEF by blockIdx.y; then the EP chunk partials of the diagonal in the
// fixed blocked tree.  Both leave at once otherwise (k_lhs_rows does it).
template <typename CT>
__global__ void k_lhs_pairs(DevMesh m, DevState s, const int* flags, LhsTerms lt, LhsPairs lp);
__global__ void k_lhs_ep(DevMesh m, DevState s, const int* flags, double* ep_part);
__global__ void k_coef_restore(DevMesh m, DevState s, const double* cs, const double* w);
// one block more than the elements need: the last finishes the GWCE matrix
__global__ void k_rhs_elem(DevMesh m, DevState s, RhsPrm q, int* flags, double* contrib, int nch,
                           const double* ep_part, double* scal);
// the same in FP32 (the mixed path): FP32 contributions
__global__ void k_rhs_elem32(DevMesh m, DevState s, RhsPrm q, int* flags, float* contrib,
                             int nch, const double* ep_part, double* scal);
// Jacobi scaling of the GWCE rows ahead of k_jcg (its first scaling pass)
struct JcgPre {
  double* w;          // 1/sqrt(diag)
  unsigned* barrier;  // grid barrier words of the coming k_jcg launch, cleared here
  int barrier_words;
  int* flags;
  double* etas_prev;  // the previous step's ETAS (kept when keep_prev != 0)
  int warm;           // 1: start JCG from 2*ETAS - ETAS_prev instead of ETAS;
                      // 2: also form the learned start ETAS + ml_corr into ml_u
  int keep_prev;      // record this step's starting ETAS in etas_prev
  const double* ml_corr;  // warm == 2: the network's correction per node
  double* ml_u;           // warm == 2: the learned start, scaled
};
// CT: the type of the element contributions (float: compensated FP32 sums)
template <bool LUMPED, typename CT>
__global__ void k_rhs_node(DevMesh m, DevState s, const CT* contrib, const double* eta_bc,
                           const double* scal, JcgPre pre);
// DUAL: start from the better (smaller initial pseudo-residual) of a.u and a.u2;
// STAGED: the products read the pseudo-residual from LDS (a.hrows >= 0).
// Instantiated for all four combinations.
template <bool DUAL, bool STAGED>
__global__ void k_jcg(JcgArgs a);
// nsys systems in one launch: blocks first[m] .. first[m+1]-1 solve args[m]
template <bool DUAL, bool STAGED>
__global__ void k_jcg_batch(const JcgArgs* args, const int* first, int nsys);
// more row tiles than blocks: the iterates in memory (a.uo, a.dpv, a.ndv)
__global__ void k_jcg_mem(JcgArgs a);
__global__ void k_wd_d1(DevMesh m, DevState s, int* flags, double h0, double habsmin);
__global__ void k_wd_elem(DevMesh m, DevState s, WdPrm q, WdWork w);
__global__ void k_wd_index(const int* bcnt, int nb, int* bofs, int* flags);
__global__ void k_wd_pred(DevMesh m, DevState s, WdWork w);
__global__ void k_wd_close(WdWork w, int* flags);
__global__ void k_wd_link(WdWork w, int* flags, double* tk);
__global__ void k_wd_final_area(DevMesh m, DevState s, WdWork w, const int* flags, double* tk,
                                double* area_slot);
__global__ void k_wd_d2(DevMesh m, DevState s, const double* area_slot, int* flags, double ifnlfa);
__global__ void k_wd_alpha(DevMesh m, DevState s, int* flags, double slim);
__global__ void k_h2(int np, const double* dp, const double* eta2, double ifnlfa, double* h2);
template <bool ALPHA>
__global__ void k_mom_elem(DevMesh m, DevState s, MomPrm q, double* cx, double* cy, int* flags);
// the same in FP32 (the mixed path)
template <bool ALPHA>
__global__ void k_mom_elem32(DevMesh m, DevState s, MomPrm q, float* cx, float* cy, int* flags);
template <typename CT>
__global__ void k_mom_node(DevMesh m, DevState s, MomPrm q, const CT* cx, const CT* cy);
__global__ void k_fill_u32(unsigned* p, unsigned v, int n);
__global__ void k_fill_i32(int* p, int v, int n);
__global__ void k_total_area(DevMesh m, DevState s);

}  // namespace adcgpu
