// Device kernels of the 2DDI per-timestep hot path for gfx950 (CDNA4).
//
// Parity design: every per-node sum the reference forms by scattering element
// contributions in ascending element order (GWCE matrix, GWCE load vector,
// momentum load vectors, TotalArea) is formed here by a node-centric gather
// over the node's incident elements in ascending element id, which is the same
// order, so the result is bit-identical.  Inner products use one fixed blocked
// tree (2048-row chunks, 8 rows per thread, xor butterfly over the 64 lanes,
// (w0+w1)+(w2+w3) over the 4 waves; chunk sums reduced with the same pattern),
// identical on every launch shape.  Built with -ffp-contract=off so no
// multiply-add is fused.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "reduce.h"

namespace adcgpu {

// ------------------------------------------------------------ time levels
__global__ __launch_bounds__(256) void k_init_h(int np, const double* __restrict__ dp, const double* __restrict__ eta2,
                         double ifnlfa, double* __restrict__ h1, double* __restrict__ h2) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double h = dp[i] + ifnlfa * eta2[i];
  h2[i] = h;
  h1[i] = h;
}

// applyTidePotentialForcing's constituent sum (src/timestep.F:1527-1557): TIP2
// starts at zero and takes each constituent's term in order; the
// self-attraction term of NTIP = 2 is absent (its amplitudes are zero for
// NTIP = 1, src/read_input.F:6460, and adding a zero to this sum changes no bit).
__global__ __launch_bounds__(256) void k_tide_pot(int np, TipPrm p, const double* __restrict__ ln,
                                                  const double* __restrict__ slam,
                                                  double* __restrict__ tip2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double v = 0.0;
  for (int j = 0; j < p.ntif; ++j) {
    const double argtp = p.argt[j] + p.na[j] * slam[i];
    v = v + p.tpmul[j] * ln[(size_t)p.na[j] * np + i] * cos(argtp);
  }
  tip2[i] = v;
}

// Apply2DBottomFriction (src/nodalattr.F:2400-2411), and the TKM = (TK, TK, 0)
// of Apply2DInternalWaveDrag without wave drag; TKM is represented by TK alone.
__device__ __forceinline__ void friction_node(int i, const FricPrm& f, double* __restrict__ tk) {
  double u = f.uu1[i], v = f.vv1[i];
  double uv1 = __builtin_sqrt(u * u + v * v);
  double h1 = f.dp[i] + f.ifnlfa * f.eta2[i];
  const double hyb = hybrid_term(f.hyb, h1);
  double a = f.fric[i] * (f.iflinbf + (uv1 / h1) * (f.ifnlbf + hyb));
  tk[i] = fmin(a, f.tkmax);
}

__global__ __launch_bounds__(256) void k_friction(int np, FricPrm f, double* __restrict__ tk) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  friction_node(i, f, tk);
}

// Holland (1980) vortex forcing on the device (synthetic storms for the
// resident benchmark run).  Same formulas as the oracle's host version; device
// exp/pow/sqrt differ from host libm by a few ulp.
__device__ __forceinline__ void holland_node(int i, const double* __restrict__ x,
                                             const double* __restrict__ y, const HollandPrm& p,
                                             double* __restrict__ wsx, double* __restrict__ wsy,
                                             double* __restrict__ pr) {
  const double t = p.t;
  const double pc = p.v[0] * 100.0, pn = p.v[1] * 100.0, rmax = p.v[2], B = p.v[3];
  const double xc = p.v[4] + p.v[6] * t, yc = p.v[5] + p.v[7] * t, f = p.v[8];
  const double rampd = p.v[9], rhoa = p.v[10], wfac = p.v[11];
  const double beta = p.v[12] * 3.141592653589793 / 180.0;
  const double rho0 = 1000.0, g = 9.80665;
  const double ramp = rampd > 0.0 ? tanh((2.0 * t / 86400.0) / rampd) : 1.0;
  const double dpr = pn - pc, cb = cos(beta), sb = sin(beta);
  double dx = x[i] - xc, dy = y[i] - yc;
  double r = __builtin_sqrt(dx * dx + dy * dy);
  if (r < 1.0) r = 1.0;
  double a = pow(rmax / r, B);
  double e = exp(-a);
  double pp = pc + dpr * e;
  double rf2 = 0.5 * r * f;
  double vg = __builtin_sqrt(B / rhoa * a * dpr * e + rf2 * rf2) - rf2;
  double w = wfac * vg;
  double tx = -dy / r, ty = dx / r, rx = dx / r, ry = dy / r;
  double wx = w * (cb * tx - sb * rx), wy = w * (cb * ty - sb * ry);
  double wm = __builtin_sqrt(wx * wx + wy * wy);
  double cd = 0.001 * (0.75 + 0.067 * wm);
  if (cd > 0.0035) cd = 0.0035;
  double sc = ramp * cd * (rhoa / rho0) * wm;
  wsx[i] = sc * wx;
  wsy[i] = sc * wy;
  pr[i] = (pn + ramp * (pp - pn)) / (rho0 * g);
}

__global__ __launch_bounds__(256) void k_holland(int np, const double* __restrict__ x, const double* __restrict__ y,
                          HollandPrm p, double* __restrict__ wsx, double* __restrict__ wsy,
                          double* __restrict__ pr) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  holland_node(i, x, y, p, wsx, wsy, pr);
}

// Friction and the step's Holland forcing in one pass (the resident run).
__global__ __launch_bounds__(256) void k_friction_holland(int np, FricPrm f, double* __restrict__ tk,
                                                          const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          HollandPrm p, double* __restrict__ wsx,
                                                          double* __restrict__ wsy,
                                                          double* __restrict__ pr) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  friction_node(i, f, tk);
  holland_node(i, x, y, p, wsx, wsy, pr);
}

// --------------------------------------------------------------- GWCE LHS
// The matrix is rebuilt when NCChange > 0 and NCChange is then cleared
// (src/gwce.F:427-428): the LHS kernels test F_NCCHANGE, and the last block
// of the right-hand-side element kernel, which runs after them and before
// wet/dry sets the flag again, finishes the matrix and clears it.

// GWCE matrix, element part (src/gwce.F:456-592): for element e and its local
// vertex k, the diagonal term and the two off-diagonal terms of row NM(e,k),
// written to the row's incidence slot p = slot[3e+k] (sd, sa, sb).
// The three terms of element e's vertex k before the NCEle factor: diagonal
// (MsOn + GA*(FDXk*FDXk + FDYk*FDYk)) and the off-diagonals in the
// reference's local order (k+1, k+2 mod 3 map to CSR_POSITION(e,k,1) and
// (e,k,2)); the lumped LHS has the diagonal MsFacLOnDiag*SPM with OnDiag = 4
// only (src/gwce.F:596-635).  `al` is the element's ALPHAL.
template <bool LUMPED>
__device__ __forceinline__ void lhs_elem_terms(const DevMesh& m, int e, double al, double td[3],
                                               double ta[3], double tb[3]) {
  const double dt = m.dt;
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  if (LUMPED) {
    const double area = m.areas[e] * 0.5;
    const double t0avg = (m.tau0[n1] + m.tau0[n2] + m.tau0[n3]) / 3.0;
    const double mson = (4.0 * area * (1.0 / dt + t0avg / 2.0)) / dt / 12.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      td[k] = mson * 1.0;
      ta[k] = 0.0;
      tb[k] = 0.0;
    }
    return;
  }
  const double fx[3] = {m.fdx[3 * e], m.fdx[3 * e + 1], m.fdx[3 * e + 2]};
  const double fy[3] = {m.fdy[3 * e], m.fdy[3 * e + 1], m.fdy[3 * e + 2]};
  const double area2 = m.areas[e];
  const double area = area2 * 0.5;
  const double area4 = area2 * 2.0;
  const double dpavg = (m.dp[n1] + m.dp[n2] + m.dp[n3]) / 3.0;
  const double ga = al * m.a00 * m.g * dpavg * 1.0 / area4;
  const double t0avg = (m.tau0[n1] + m.tau0[n2] + m.tau0[n3]) / 3.0;
  const double mson = (2.0 * area * (1.0 / dt + t0avg / 2.0)) / dt / 12.0;
  const double msoff = (1.0 * area * (1.0 / dt + t0avg / 2.0)) / dt / 12.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ka = (k == 0) ? 1 : 0, kb = (k == 2) ? 1 : 2;
    td[k] = mson * 1.0 + ga * (1.0 * fx[k] * fx[k] + 1.0 * fy[k] * fy[k]);
    ta[k] = msoff * 1.0 + ga * (1.0 * fx[k] * fx[ka] + 1.0 * fy[k] * fy[ka]);
    tb[k] = msoff * 1.0 + ga * (1.0 * fx[k] * fx[kb] + 1.0 * fy[k] * fy[kb]);
  }
}

// Element pass of a rebuild, used only while some ALPHAL differs from 1
// (F_ALPHA); otherwise the row pass multiplies the cached terms by NCEle.
template <bool LUMPED>
__global__ __launch_bounds__(256) void k_lhs_elem(DevMesh m, DevState s,
                                                  const int* __restrict__ flags,
                                                  double* __restrict__ sd, double* __restrict__ sa,
                                                  double* __restrict__ sb) {
  if (flags[F_NCCHANGE] == 0 || flags[F_ALPHA] == 0) return;
  const int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  const int ncele = s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e];
  const double NCELE = (double)ncele;
  double td[3], ta[3], tb[3];
  lhs_elem_terms<LUMPED>(m, e, s.alphal[e], td, ta, tb);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int p = m.slot[3 * e + k];
    sd[p] = td[k] * NCELE;
    if (!LUMPED) {
      sa[p] = ta[k] * NCELE;
      sb[p] = tb[k] * NCELE;
    }
  }
}
template __global__ void k_lhs_elem<false>(DevMesh, DevState, const int*, double*, double*, double*);
template __global__ void k_lhs_elem<true>(DevMesh, DevState, const int*, double*, double*, double*);

// The terms with NCEle = 1 and ALPHAL = 1, in slot order, once per context.
// Multiplying one by NCEle (0 or 1) later gives exactly the element pass's
// value: x * 1 = x, and x * 0 is a zero that leaves every node sum unchanged
// (the sums start at +0).
template <bool LUMPED>
__global__ __launch_bounds__(256) void k_lhs_terms(DevMesh m, double* __restrict__ td,
                                                   double* __restrict__ ta,
                                                   double* __restrict__ tb,
                                                   float* __restrict__ td32,
                                                   float* __restrict__ ta32,
                                                   float* __restrict__ tb32) {
  const int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  double d[3], a[3], b[3];
  lhs_elem_terms<LUMPED>(m, e, 1.0, d, a, b);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int p = m.slot[3 * e + k];
    td[p] = d[k];
    ta[p] = a[k];
    tb[p] = b[k];
    if (td32) {
      td32[p] = (float)d[k];
      ta32[p] = (float)a[k];
      tb32[p] = (float)b[k];
    }
  }
}
template __global__ void k_lhs_terms<false>(DevMesh, double*, double*, double*, float*, float*,
                                            float*);
template __global__ void k_lhs_terms<true>(DevMesh, double*, double*, double*, float*, float*,
                                           float*);

__global__ __launch_bounds__(256) void k_ncele(DevMesh m, DevState s) {
  const int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  s.ncele[e] = (unsigned char)(s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e]);
}

// GWCE matrix, row part: row i of COEF is the sum of its incidence slots in
// ascending element order, each row held in LDS while it is formed (the
// reference adds into COEF(i,j) element by element, src/gwce.F:538-592);
// also the EP chunk partials of sum COEF(i,1)^2 (src/gwce.F:646-650) in the
// fixed blocked tree.  Dynamic LDS: MNEI x RED_THREADS doubles, slot-major.
template <bool LUMPED, typename CT>
__global__ __launch_bounds__(RED_THREADS) void k_lhs_rows(DevMesh m, DevState s,
                                                          const int* __restrict__ flags,
                                                          const double* __restrict__ sd,
                                                          const double* __restrict__ sa,
                                                          const double* __restrict__ sb,
                                                          LhsTerms lt,
                                                          double* __restrict__ ep_part,
                                                          const double* __restrict__ restore_cs,
                                                          const double* __restrict__ restore_w) {
  constexpr bool MIXED = sizeof(CT) == 4;
  __shared__ double lds[8];
  extern __shared__ double racc_lds[];
  double(*racc)[RED_THREADS] = reinterpret_cast<double(*)[RED_THREADS]>(racc_lds);
  // mixed path: the same LDS words hold a compensated FP32 pair per column
  float2(*racc2)[RED_THREADS] = reinterpret_cast<float2(*)[RED_THREADS]>(racc_lds);
  const int np = m.np;
  const int t = threadIdx.x;
  const int chunk = xcd_tile(blockIdx.x, gridDim.x);
  const int base = chunk * RED_CHUNK + t;
  if (flags[F_NCCHANGE] == 0) {
    // no rebuild: the previous solve left the off-diagonals scaled; restore
    // them as unscal does (src/itpackv.F:2460-2479), same operations
    if (restore_cs)
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        const int i = base + RED_THREADS * r;
        if (i >= np) break;
        const int nn = m.nneigh[i];
        const double ri = 1.0 / restore_w[i];
        for (int j = 1; j < nn; ++j) {
          const size_t q = (size_t)j * np + i;
          s.coef[q] = restore_cs[q] * ri * (1.0 / restore_w[m.col[q]]);
        }
      }
    return;
  }
  // the cached terms times NCEle, unless some ALPHAL differs from 1 (then
  // the element pass has written the slot values); the lumped matrix has no
  // ALPHAL term
  const bool cached = LUMPED || flags[F_ALPHA] == 0;
  if (!LUMPED && cached && lt.pairs) return;  // k_lhs_pairs and k_lhs_ep do it
  double acc = 0.0;
  for (int r = 0; r < RED_PER_THREAD; ++r) {
    const int i = base + RED_THREADS * r;
    if (i >= np) break;
    const int nn = m.nneigh[i];
    if (LUMPED) {
      // COEFD only; the off-diagonal columns stay zero
      double diag = 0.0;
      if (MIXED) {
        Kahan32 k;
        for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p)
          k.add(lt.td32[p] * (float)s.ncele[(unsigned)m.inc[p] / 3u]);
        diag = k.value();
      } else {
        for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p)
          diag = diag + lt.td[p] * (double)s.ncele[(unsigned)m.inc[p] / 3u];
      }
      s.coef[i] = diag;
      for (int j = 1; j < nn; ++j) s.coef[(size_t)j * np + i] = 0.0;
      acc = acc + diag * diag;
      continue;
    }
    if (MIXED) {
      // FP32 terms, compensated sums per column
      for (int j = 1; j < nn; ++j) racc2[j][t] = make_float2(0.0f, 0.0f);
      Kahan32 kd;
      for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) {
        const int jab = m.slot_cols[p];
        const int ja = jab & 0xffff, jb = jab >> 16;
        float vd, va, vb;
        if (cached) {
          const float nce = (float)s.ncele[(unsigned)m.inc[p] / 3u];
          vd = lt.td32[p] * nce;
          va = lt.ta32[p] * nce;
          vb = lt.tb32[p] * nce;
        } else {
          vd = (float)sd[p];
          va = (float)sa[p];
          vb = (float)sb[p];
        }
        kd.add(vd);
        Kahan32 ka{racc2[ja][t].x, racc2[ja][t].y};
        ka.add(va);
        racc2[ja][t] = make_float2(ka.s, ka.c);
        Kahan32 kb{racc2[jb][t].x, racc2[jb][t].y};
        kb.add(vb);
        racc2[jb][t] = make_float2(kb.s, kb.c);
      }
      const double diag = kd.value();
      s.coef[i] = diag;
      for (int j = 1; j < nn; ++j) {
        const Kahan32 kj{racc2[j][t].x, racc2[j][t].y};
        s.coef[(size_t)j * np + i] = kj.value();
      }
      acc = acc + diag * diag;
      continue;
    }
    for (int j = 1; j < nn; ++j) racc[j][t] = 0.0;
    double diag = 0.0;
    for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) {
      const int jab = m.slot_cols[p];
      const int ja = jab & 0xffff, jb = jab >> 16;
      double vd, va, vb;
      if (cached) {
        const double nce = (double)s.ncele[(unsigned)m.inc[p] / 3u];
        vd = lt.td[p] * nce;
        va = lt.ta[p] * nce;
        vb = lt.tb[p] * nce;
      } else {
        vd = sd[p];
        va = sa[p];
        vb = sb[p];
      }
      diag = diag + vd;
      racc[ja][t] = racc[ja][t] + va;
      racc[jb][t] = racc[jb][t] + vb;
    }
    s.coef[i] = diag;
    for (int j = 1; j < nn; ++j) s.coef[(size_t)j * np + i] = racc[j][t];
    acc = acc + diag * diag;
  }
  double tot = chunk_tree(acc, lds);
  if (t == 0) ep_part[chunk] = tot;
}
template __global__ void k_lhs_rows<false, double>(DevMesh, DevState, const int*, const double*, const double*, const double*, LhsTerms, double*, const double*, const double*);
template __global__ void k_lhs_rows<true, double>(DevMesh, DevState, const int*, const double*, const double*, const double*, LhsTerms, double*, const double*, const double*);
template __global__ void k_lhs_rows<false, float>(DevMesh, DevState, const int*, const double*, const double*, const double*, LhsTerms, double*, const double*, const double*);
template __global__ void k_lhs_rows<true, float>(DevMesh, DevState, const int*, const double*, const double*, const double*, LhsTerms, double*, const double*, const double*);

__global__ __launch_bounds__(256) void k_lhs_pair_setup(DevMesh m, LhsTerms lt,
                                                        const int* __restrict__ c1,
                                                        const int* __restrict__ c2,
                                                        double* __restrict__ t1,
                                                        double* __restrict__ t2,
                                                        float* __restrict__ t1f,
                                                        float* __restrict__ t2f,
                                                        int* __restrict__ e1, int* __restrict__ e2) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t)(m.mnei - 1) * m.np) return;
  const int a = c1[q], b = c2[q];
  // the term of code c: slot c/2's ta (c even) or tb (c odd) term
  if (t1) {
    t1[q] = a < 0 ? 0.0 : ((a & 1) ? lt.tb : lt.ta)[a >> 1];
    t2[q] = b < 0 ? 0.0 : ((b & 1) ? lt.tb : lt.ta)[b >> 1];
  }
  if (t1f) {
    t1f[q] = a < 0 ? 0.0f : ((a & 1) ? lt.tb32 : lt.ta32)[a >> 1];
    t2f[q] = b < 0 ? 0.0f : ((b & 1) ? lt.tb32 : lt.ta32)[b >> 1];
  }
  e1[q] = a < 0 ? 0 : (int)((unsigned)m.inc[a >> 1] / 3u);
  e2[q] = b < 0 ? 0 : (int)((unsigned)m.inc[b >> 1] / 3u);
}

// COEF(i, j) of a cached rebuild.  Column 0 (the diagonal) sums the row's
// slots in ascending element order from +0, as k_lhs_rows does; an
// off-diagonal is (+0 + t1 NCEle) + t2 NCEle, the same operations as
// k_lhs_rows' accumulation of its one or two contributions (an absent second
// term adds +0, which changes no sum that started from +0).  The mixed path
// (CT = float) takes FP32 terms and sums in FP64.
template <typename CT>
__global__ __launch_bounds__(256) void k_lhs_pairs(DevMesh m, DevState s,
                                                   const int* __restrict__ flags, LhsTerms lt,
                                                   LhsPairs lp) {
  constexpr bool MIXED = sizeof(CT) == 4;
  if (flags[F_NCCHANGE] == 0 || flags[F_ALPHA] != 0) return;
  const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int np = m.np;
  if (i >= np || j >= m.nneigh[i]) return;
  if (j == 0) {
    double diag = 0.0;
    for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) {
      const double nce = (double)s.ncele[(unsigned)m.inc[p] / 3u];
      diag = diag + (MIXED ? (double)lt.td32[p] : lt.td[p]) * nce;
    }
    s.coef[i] = diag;
    return;
  }
  const size_t q = (size_t)(j - 1) * np + i;
  const double n1 = (double)s.ncele[lp.e1[q]], n2 = (double)s.ncele[lp.e2[q]];
  const double a = MIXED ? (double)lp.t1f[q] : lp.t1[q];
  const double b = MIXED ? (double)lp.t2f[q] : lp.t2[q];
  s.coef[(size_t)j * np + i] = (0.0 + a * n1) + b * n2;
}
template __global__ void k_lhs_pairs<double>(DevMesh, DevState, const int*, LhsTerms, LhsPairs);
template __global__ void k_lhs_pairs<float>(DevMesh, DevState, const int*, LhsTerms, LhsPairs);

// The EP chunk partials of a cached rebuild: sum COEF(i,1)^2 in k_lhs_rows'
// fixed blocked tree (thread t of chunk c adds rows c 2048 + t + 512 r,
// r = 0..3, in order; then the chunk tree).
__global__ __launch_bounds__(RED_THREADS) void k_lhs_ep(DevMesh m, DevState s,
                                                        const int* __restrict__ flags,
                                                        double* __restrict__ ep_part) {
  __shared__ double lds[8];
  if (flags[F_NCCHANGE] == 0 || flags[F_ALPHA] != 0) return;
  const int chunk = xcd_tile(blockIdx.x, gridDim.x);
  const int base = chunk * RED_CHUNK + threadIdx.x;
  double acc = 0.0;
  for (int r = 0; r < RED_PER_THREAD; ++r) {
    const int i = base + RED_THREADS * r;
    if (i >= m.np) break;
    const double d = s.coef[i];
    acc = acc + d * d;
  }
  const double tot = chunk_tree(acc, lds);
  if (threadIdx.x == 0) ep_part[chunk] = tot;
}

// EP = sqrt(sum of squared diagonals / NP) from the chunk partials, and the
// elevation-boundary rows (src/gwce.F:700-743): a boundary row becomes EP on
// the diagonal with its off-diagonals and OBCCOEF entries zeroed; an entry
// that couples another row to a boundary node moves to OBCCOEF.  Only the
// listed positions are touched: one block.  Zero diagonals of dry rows are
// set to EP by k_rhs_node.
template <bool LUMPED>
__device__ void lhs_bc_block(const DevMesh& m, DevState s, int nch,
                             const double* __restrict__ ep_part, double* __restrict__ scal) {
  __shared__ double lds[4];
  __shared__ double ep_s;
  const int np = m.np, t = threadIdx.x;
  double acc = 0.0;
  for (int c = t; c < nch; c += RED_XTHREADS) acc = acc + ep_part[c];
  const double tot = block_tree(acc, lds);
  if (t == 0) {
    ep_s = __builtin_sqrt(tot / (double)np);
    scal[S_EP] = ep_s;
  }
  __syncthreads();
  const double ep = ep_s;
  double* c = s.coef;
  for (int k = t; k < m.neta; k += blockDim.x) {
    const int i = m.nbd[k];
    c[i] = ep;
    const int nn = m.nneigh[i];
    for (int j = 1; j < nn; ++j) c[(size_t)j * np + i] = 0.0;
  }
  if (LUMPED) return;  // a lumped LHS has no boundary columns
  for (int k = t; k < m.nobc; k += blockDim.x) {
    const int sl = m.obc_s[k];
    if (sl < 0) {
      s.obc[-1 - sl] = 0.0;
    } else {
      const int q = m.obc_q[k];
      s.obc[sl] = c[q];
      c[q] = 0.0;
    }
  }
}

// The off-diagonals of a solve whose unscaling was deferred, restored as
// unscal does (src/itpackv.F:2460-2479).
__global__ __launch_bounds__(256) void k_coef_restore(DevMesh m, DevState s,
                                                      const double* __restrict__ cs,
                                                      const double* __restrict__ w) {
  const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= m.np) return;
  const int np = m.np, nn = m.nneigh[i];
  const double ri = 1.0 / w[i];
  for (int j = 1; j < nn; ++j) {
    const size_t q = (size_t)j * np + i;
    s.coef[q] = cs[q] * ri * (1.0 / w[m.col[q]]);
  }
}

// --------------------------------------------------------------- GWCE RHS
// Element contributions Temp_LV_A1..3 * NCEle (src/gwce.F:1000-1615).  Terms
// the reference multiplies by identically zero quantities in this
// configuration (sponge layer, dispersion, baroclinic, tidal potential, water
// level offset) are omitted; they only add signed zeros.
__global__ __launch_bounds__(256) void k_rhs_elem(DevMesh m, DevState s, RhsPrm q, int* flags,
                                                  double* __restrict__ contrib, int nch,
                                                  const double* __restrict__ ep_part,
                                                  double* __restrict__ scal) {
  // The last block finishes the GWCE matrix instead (EP and the elevation-
  // boundary rows, when it was rebuilt), beside the element blocks, which read
  // neither; it then clears NCChange (src/gwce.F:427-428), after every thread
  // of the block has read it.
  const int nbe = (int)gridDim.x - 1;
  if ((int)blockIdx.x == nbe) {
    const int ncc = flags[F_NCCHANGE];
    if (ncc > 0) {
      if (q.lumped)
        lhs_bc_block<true>(m, s, nch, ep_part, scal);
      else
        lhs_bc_block<false>(m, s, nch, ep_part, scal);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (ncc > 0) flags[F_NREBUILD] += 1;
      flags[F_NCCHANGE] = 0;
    }
    return;
  }
  int e = xcd_tile(blockIdx.x, nbe) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  const int ncele = s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e];
  // a dry element's contributions are its (finite) terms x 0: zeros, and a
  // node sum that starts at +0 never holds -0, so a zero of either sign
  // leaves it unchanged; the terms are skipped, and all lanes meet again for
  // one set of stores
  double R1 = 0.0, R2 = 0.0, R3 = 0.0;
  if (ncele != 0) {
    const double DT = m.dt, G = m.g, A00 = m.a00, B00 = m.b00, C00 = m.c00;
    const double E0N1 = s.eta1[n1], E0N2 = s.eta1[n2], E0N3 = s.eta1[n3];
    const double E1N1 = s.eta2[n1], E1N2 = s.eta2[n2], E1N3 = s.eta2[n3];
    const double E1N1SQ = E1N1 * E1N1, E1N2SQ = E1N2 * E1N2, E1N3SQ = E1N3 * E1N3;
    const double ESN1 = s.etas[n1], ESN2 = s.etas[n2], ESN3 = s.etas[n3];
    const double U1N1 = s.uu1[n1], U1N2 = s.uu1[n2], U1N3 = s.uu1[n3];
    const double V1N1 = s.vv1[n1], V1N2 = s.vv1[n2], V1N3 = s.vv1[n3];
    const double QX1N1 = s.qx1[n1], QX1N2 = s.qx1[n2], QX1N3 = s.qx1[n3];
    const double QY1N1 = s.qy1[n1], QY1N2 = s.qy1[n2], QY1N3 = s.qy1[n3];
    const double H1N1 = s.h2[n1], H1N2 = s.h2[n2], H1N3 = s.h2[n3];
    const double EVMN1 = m.evm[n1], EVMN2 = m.evm[n2], EVMN3 = m.evm[n3];
    const double T0N1 = m.tau0[n1], T0N2 = m.tau0[n2], T0N3 = m.tau0[n3];
    double PR1N1 = 0.0, PR1N2 = 0.0, PR1N3 = 0.0, WSXAvg = 0.0, WSYAvg = 0.0;
    if (q.nws) {
      PR1N1 = s.pr1[n1]; PR1N2 = s.pr1[n2]; PR1N3 = s.pr1[n3];
      WSXAvg = (1.0 * s.wsx1[n1] + 1.0 * s.wsx1[n2] + 1.0 * s.wsx1[n3]) / 3.0;
      WSYAvg = (1.0 * s.wsy1[n1] + 1.0 * s.wsy1[n2] + 1.0 * s.wsy1[n3]) / 3.0;
    }
    // tidal potential at time level s (src/gwce.F:1181-1185); PR - 0 is PR
    double TiPN1 = 0.0, TiPN2 = 0.0, TiPN3 = 0.0;
    if (s.tip1) {
      TiPN1 = s.tip1[n1]; TiPN2 = s.tip1[n2]; TiPN3 = s.tip1[n3];
    }
    // TKM(1)=TKM(2)=TK, TKM(3)=0
    const double tk1 = s.tk[n1], tk2 = s.tk[n2], tk3 = s.tk[n3];
    const double BSXN1 = tk1 * QX1N1 + 0.0 * QY1N1, BSYN1 = 0.0 * QX1N1 + tk1 * QY1N1;
    const double BSXN2 = tk2 * QX1N2 + 0.0 * QY1N2, BSYN2 = 0.0 * QX1N2 + tk2 * QY1N2;
    const double BSXN3 = tk3 * QX1N3 + 0.0 * QY1N3, BSYN3 = 0.0 * QX1N3 + tk3 * QY1N3;
    const double AreaIE2 = m.areas[e];
    const double AreaIE = AreaIE2 / 2.0;
    const double AreaIE4 = 2.0 * AreaIE2;
    const double FDX1 = m.fdx[3 * e], FDX2 = m.fdx[3 * e + 1], FDX3 = m.fdx[3 * e + 2];
    const double FDY1 = m.fdy[3 * e], FDY2 = m.fdy[3 * e + 1], FDY3 = m.fdy[3 * e + 2];
    const double E0XGrad2A = q.lumped ? 0.0 : E0N1 * FDX1 + E0N2 * FDX2 + E0N3 * FDX3;
    const double E0YGrad2A = q.lumped ? 0.0 : E0N1 * FDY1 + E0N2 * FDY2 + E0N3 * FDY3;
    const double E1XGrad2A = E1N1 * FDX1 + E1N2 * FDX2 + E1N3 * FDX3;
    const double E1YGrad2A = E1N1 * FDY1 + E1N2 * FDY2 + E1N3 * FDY3;
    const double Tau0XGrad2A = 1.0 * (T0N1 * FDX1 + T0N2 * FDX2 + T0N3 * FDX3);
    const double Tau0YGrad2A = 1.0 * (T0N1 * FDY1 + T0N2 * FDY2 + T0N3 * FDY3);
    const double EVMXGrad = (EVMN1 * FDX1 + EVMN2 * FDX2 + EVMN3 * FDX3) / AreaIE2;
    const double EVMYGrad = (EVMN1 * FDY1 + EVMN2 * FDY2 + EVMN3 * FDY3) / AreaIE2;
    const double EVMAvgODT = ((EVMN1 + EVMN2 + EVMN3) / 3.0) / DT;
    const double MX = (EVMXGrad * (QX1N1 * FDX1 + QX1N2 * FDX2 + QX1N3 * FDX3) +
                       EVMYGrad * (QY1N1 * FDX1 + QY1N2 * FDX2 + QY1N3 * FDX3) -
                       EVMAvgODT * (ESN1 * FDX1 + ESN2 * FDX2 + ESN3 * FDX3)) / AreaIE2;
    const double MY = (EVMXGrad * (QX1N1 * FDY1 + QX1N2 * FDY2 + QX1N3 * FDY3) +
                       EVMYGrad * (QY1N1 * FDY1 + QY1N2 * FDY2 + QY1N3 * FDY3) -
                       EVMAvgODT * (ESN1 * FDY1 + ESN2 * FDY2 + ESN3 * FDY3)) / AreaIE2;
    const double CorifAvg = (m.corif[n1] + m.corif[n2] + m.corif[n3]) / 3.0;
    const double Tau0Avg = (T0N1 + T0N2 + T0N3) / 3.0;
    const double Tau0QXAvg = (T0N1 * QX1N1 + T0N2 * QX1N2 + T0N3 * QX1N3) / 3.0;
    const double Tau0QYAvg = (T0N1 * QY1N1 + T0N2 * QY1N2 + T0N3 * QY1N3) / 3.0;
    const double U1Avg = (U1N1 + U1N2 + U1N3) / 3.0;
    const double V1Avg = (V1N1 + V1N2 + V1N3) / 3.0;
    const double QX1Avg = (QX1N1 + QX1N2 + QX1N3) / 3.0;
    const double QY1Avg = (QY1N1 + QY1N2 + QY1N3) / 3.0;
    const double ESAvg = (ESN1 + ESN2 + ESN3) / 3.0;
    const double DPAvg = (m.dp[n1] + m.dp[n2] + m.dp[n3]) / 3.0;
    const double al = s.alphal[e];
    const double GDP = al * G * DPAvg * 1.0 / AreaIE4;   // both time levels (IFNL_HDP = 0)
    const double HAvg = (H1N1 + H1N2 + H1N3) / 3.0;
    const double GHAvgOAreaIE2 = (al * G * HAvg) / AreaIE2;
    const double BSXAvg = (BSXN1 + BSXN2 + BSXN3) / 3.0;
    const double BSYAvg = (BSYN1 + BSYN2 + BSYN3) / 3.0;
    const double MsFacR = AreaIE * (1.0 / DT - Tau0Avg / 2.0) / DT / 12.0;
    const double GOAreaIE4 = al * G / AreaIE4;
    const double Tau0SpaVar = (QX1Avg * Tau0XGrad2A + QY1Avg * Tau0YGrad2A) / 6.0;
    const double A00pB00 = A00 + B00;
    double JXAvg = CorifAvg * QY1Avg -
                   q.ifnlfa * GOAreaIE4 * 1.0 * (E1N1SQ * FDX1 + E1N2SQ * FDX2 + E1N3SQ * FDX3) -
                   GHAvgOAreaIE2 * ((PR1N1 - TiPN1) * FDX1 + (PR1N2 - TiPN2) * FDX2 +
                                    (PR1N3 - TiPN3) * FDX3) +
                   WSXAvg - BSXAvg + MX + Tau0QXAvg;
    double JYAvg = -CorifAvg * QX1Avg -
                   q.ifnlfa * GOAreaIE4 * 1.0 * (E1N1SQ * FDY1 + E1N2SQ * FDY2 + E1N3SQ * FDY3) -
                   GHAvgOAreaIE2 * ((PR1N1 - TiPN1) * FDY1 + (PR1N2 - TiPN2) * FDY2 +
                                    (PR1N3 - TiPN3) * FDY3) +
                   WSYAvg - BSYAvg + MY + Tau0QYAvg;
    JXAvg = JXAvg -
            q.ifnlct * (QX1Avg * (U1N1 * FDX1 + U1N2 * FDX2 + U1N3 * FDX3) +
                        QY1Avg * (U1N1 * FDY1 + U1N2 * FDY2 + U1N3 * FDY3)) / AreaIE2 +
            q.ifnlcat * U1Avg * ESAvg / DT;
    JYAvg = JYAvg -
            q.ifnlct * (QX1Avg * (V1N1 * FDX1 + V1N2 * FDX2 + V1N3 * FDX3) +
                        QY1Avg * (V1N1 * FDY1 + V1N2 * FDY2 + V1N3 * FDY3)) / AreaIE2 +
            q.ifnlcat * V1Avg * ESAvg / DT;
    const double A1 = (q.ondiag * ESN1 * 1.0 + q.offdiag * (ESN2 * 1.0 + ESN3 * 1.0)) * MsFacR -
                      GDP * (C00 * (1.0 * FDX1 * E0XGrad2A + 1.0 * FDY1 * E0YGrad2A)) -
                      GDP * (A00pB00 * (1.0 * FDX1 * E1XGrad2A + 1.0 * FDY1 * E1YGrad2A)) +
                      (1.0 * JXAvg * FDX1 + 1.0 * JYAvg * FDY1) * 0.5 + Tau0SpaVar;
    const double A2 = (q.ondiag * ESN2 * 1.0 + q.offdiag * (ESN1 * 1.0 + ESN3 * 1.0)) * MsFacR -
                      GDP * (C00 * (1.0 * FDX2 * E0XGrad2A + 1.0 * FDY2 * E0YGrad2A)) -
                      GDP * (A00pB00 * (1.0 * FDX2 * E1XGrad2A + 1.0 * FDY2 * E1YGrad2A)) +
                      (1.0 * JXAvg * FDX2 + 1.0 * JYAvg * FDY2) * 0.5 + Tau0SpaVar;
    const double A3 = (q.ondiag * ESN3 * 1.0 + q.offdiag * (ESN1 * 1.0 + ESN2 * 1.0)) * MsFacR -
                      GDP * (C00 * (1.0 * FDX3 * E0XGrad2A + 1.0 * FDY3 * E0YGrad2A)) -
                      GDP * (A00pB00 * (1.0 * FDX3 * E1XGrad2A + 1.0 * FDY3 * E1YGrad2A)) +
                      (1.0 * JXAvg * FDX3 + 1.0 * JYAvg * FDY3) * 0.5 + Tau0SpaVar;
    R1 = A1;
    R2 = A2;
    R3 = A3;
  }
  const double nce = (double)ncele;
  contrib[m.slot[3 * e]] = R1 * nce;
  contrib[m.slot[3 * e + 1]] = R2 * nce;
  contrib[m.slot[3 * e + 2]] = R3 * nce;
}

// The gradient sum f1*D1 + f2*D2 + f3*D3 of FP64 nodal values in FP32: with
// D1 + D2 + D3 = 0 (the shape-function derivatives of a triangle) it is
// (f2 - f1) D2 + (f3 - f1) D3, and the differences are taken in FP64 before
// rounding, so a field's mean (an elevation of metres, a pressure head of ten
// metres) costs no FP32 digits of its variation.
__device__ __forceinline__ float grad32(double f1, double f2, double f3, float d2, float d3) {
  return (float)(f2 - f1) * d2 + (float)(f3 - f1) * d3;
}

// The GWCE right-hand side element pass in FP32 (the mixed-precision path):
// the terms of k_rhs_elem with FP32 arithmetic, FP32 geometry and gradients
// formed from FP64 differences (grad32); FP32 contributions, summed with
// compensation by the node pass.
__global__ __launch_bounds__(256) void k_rhs_elem32(DevMesh m, DevState s, RhsPrm q, int* flags,
                                                    float* __restrict__ contrib, int nch,
                                                    const double* __restrict__ ep_part,
                                                    double* __restrict__ scal) {
  const int nbe = (int)gridDim.x - 1;
  if ((int)blockIdx.x == nbe) {
    const int ncc = flags[F_NCCHANGE];
    if (ncc > 0) {
      if (q.lumped)
        lhs_bc_block<true>(m, s, nch, ep_part, scal);
      else
        lhs_bc_block<false>(m, s, nch, ep_part, scal);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (ncc > 0) flags[F_NREBUILD] += 1;
      flags[F_NCCHANGE] = 0;
    }
    return;
  }
  int e = xcd_tile(blockIdx.x, nbe) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  const int ncele = s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e];
  float R1 = 0.0f, R2 = 0.0f, R3 = 0.0f;
  if (ncele != 0) {
    const float DT = (float)m.dt, G = (float)m.g, A00 = (float)m.a00, B00 = (float)m.b00,
                C00 = (float)m.c00;
    const float FDX1 = m.fdx32[3 * e], FDX2 = m.fdx32[3 * e + 1], FDX3 = m.fdx32[3 * e + 2];
    const float FDY1 = m.fdy32[3 * e], FDY2 = m.fdy32[3 * e + 1], FDY3 = m.fdy32[3 * e + 2];
    const float AreaIE2 = m.areas32[e];
    const float AreaIE = AreaIE2 / 2.0f;
    const float AreaIE4 = 2.0f * AreaIE2;
    const double e0a = s.eta1[n1], e0b = s.eta1[n2], e0c = s.eta1[n3];
    const double e1a = s.eta2[n1], e1b = s.eta2[n2], e1c = s.eta2[n3];
    const float ESN1 = (float)s.etas[n1], ESN2 = (float)s.etas[n2], ESN3 = (float)s.etas[n3];
    const double ua = s.uu1[n1], ub = s.uu1[n2], uc = s.uu1[n3];
    const double va = s.vv1[n1], vb = s.vv1[n2], vc = s.vv1[n3];
    const double qxa = s.qx1[n1], qxb = s.qx1[n2], qxc = s.qx1[n3];
    const double qya = s.qy1[n1], qyb = s.qy1[n2], qyc = s.qy1[n3];
    const float H1N1 = (float)s.h2[n1], H1N2 = (float)s.h2[n2], H1N3 = (float)s.h2[n3];
    const double eva = m.evm[n1], evb = m.evm[n2], evc = m.evm[n3];
    const double t0a = m.tau0[n1], t0b = m.tau0[n2], t0c = m.tau0[n3];
    float PR1X = 0.0f, PR1Y = 0.0f, WSXAvg = 0.0f, WSYAvg = 0.0f;
    if (q.nws || s.tip1) {
      // the pressure head less the tidal potential, in FP64 before the gradient
      double pa = 0.0, pb = 0.0, pc = 0.0;
      if (q.nws) {
        pa = s.pr1[n1]; pb = s.pr1[n2]; pc = s.pr1[n3];
      }
      if (s.tip1) {
        pa = pa - s.tip1[n1]; pb = pb - s.tip1[n2]; pc = pc - s.tip1[n3];
      }
      PR1X = grad32(pa, pb, pc, FDX2, FDX3);
      PR1Y = grad32(pa, pb, pc, FDY2, FDY3);
    }
    if (q.nws) {
      WSXAvg = ((float)s.wsx1[n1] + (float)s.wsx1[n2] + (float)s.wsx1[n3]) / 3.0f;
      WSYAvg = ((float)s.wsy1[n1] + (float)s.wsy1[n2] + (float)s.wsy1[n3]) / 3.0f;
    }
    const float U1N1 = (float)ua, U1N2 = (float)ub, U1N3 = (float)uc;
    const float V1N1 = (float)va, V1N2 = (float)vb, V1N3 = (float)vc;
    const float QX1N1 = (float)qxa, QX1N2 = (float)qxb, QX1N3 = (float)qxc;
    const float QY1N1 = (float)qya, QY1N2 = (float)qyb, QY1N3 = (float)qyc;
    const float T0N1 = (float)t0a, T0N2 = (float)t0b, T0N3 = (float)t0c;
    const float tk1 = (float)s.tk[n1], tk2 = (float)s.tk[n2], tk3 = (float)s.tk[n3];
    const float BSXAvg = (tk1 * QX1N1 + tk2 * QX1N2 + tk3 * QX1N3) / 3.0f;
    const float BSYAvg = (tk1 * QY1N1 + tk2 * QY1N2 + tk3 * QY1N3) / 3.0f;
    const float E0XGrad2A = q.lumped ? 0.0f : grad32(e0a, e0b, e0c, FDX2, FDX3);
    const float E0YGrad2A = q.lumped ? 0.0f : grad32(e0a, e0b, e0c, FDY2, FDY3);
    const float E1XGrad2A = grad32(e1a, e1b, e1c, FDX2, FDX3);
    const float E1YGrad2A = grad32(e1a, e1b, e1c, FDY2, FDY3);
    const float E1SQX = grad32(e1a * e1a, e1b * e1b, e1c * e1c, FDX2, FDX3);
    const float E1SQY = grad32(e1a * e1a, e1b * e1b, e1c * e1c, FDY2, FDY3);
    const float Tau0XGrad2A = grad32(t0a, t0b, t0c, FDX2, FDX3);
    const float Tau0YGrad2A = grad32(t0a, t0b, t0c, FDY2, FDY3);
    const float EVMXGrad = grad32(eva, evb, evc, FDX2, FDX3) / AreaIE2;
    const float EVMYGrad = grad32(eva, evb, evc, FDY2, FDY3) / AreaIE2;
    const float EVMAvgODT = (((float)eva + (float)evb + (float)evc) / 3.0f) / DT;
    const float ESX = ESN1 * FDX1 + ESN2 * FDX2 + ESN3 * FDX3;
    const float ESY = ESN1 * FDY1 + ESN2 * FDY2 + ESN3 * FDY3;
    const float MX = (EVMXGrad * grad32(qxa, qxb, qxc, FDX2, FDX3) +
                      EVMYGrad * grad32(qya, qyb, qyc, FDX2, FDX3) - EVMAvgODT * ESX) / AreaIE2;
    const float MY = (EVMXGrad * grad32(qxa, qxb, qxc, FDY2, FDY3) +
                      EVMYGrad * grad32(qya, qyb, qyc, FDY2, FDY3) - EVMAvgODT * ESY) / AreaIE2;
    const float CorifAvg = ((float)m.corif[n1] + (float)m.corif[n2] + (float)m.corif[n3]) / 3.0f;
    const float Tau0Avg = (T0N1 + T0N2 + T0N3) / 3.0f;
    const float Tau0QXAvg = (T0N1 * QX1N1 + T0N2 * QX1N2 + T0N3 * QX1N3) / 3.0f;
    const float Tau0QYAvg = (T0N1 * QY1N1 + T0N2 * QY1N2 + T0N3 * QY1N3) / 3.0f;
    const float U1Avg = (U1N1 + U1N2 + U1N3) / 3.0f;
    const float V1Avg = (V1N1 + V1N2 + V1N3) / 3.0f;
    const float QX1Avg = (QX1N1 + QX1N2 + QX1N3) / 3.0f;
    const float QY1Avg = (QY1N1 + QY1N2 + QY1N3) / 3.0f;
    const float ESAvg = (ESN1 + ESN2 + ESN3) / 3.0f;
    const float DPAvg = (float)((m.dp[n1] + m.dp[n2] + m.dp[n3]) / 3.0);
    const float al = (float)s.alphal[e];
    const float GDP = al * G * DPAvg / AreaIE4;
    const float HAvg = (H1N1 + H1N2 + H1N3) / 3.0f;
    const float GHAvgOAreaIE2 = (al * G * HAvg) / AreaIE2;
    const float MsFacR = AreaIE * (1.0f / DT - Tau0Avg / 2.0f) / DT / 12.0f;
    const float GOAreaIE4 = al * G / AreaIE4;
    const float Tau0SpaVar = (QX1Avg * Tau0XGrad2A + QY1Avg * Tau0YGrad2A) / 6.0f;
    const float A00pB00 = A00 + B00;
    const float fnlfa = (float)q.ifnlfa, fnlct = (float)q.ifnlct, fnlcat = (float)q.ifnlcat;
    float JXAvg = CorifAvg * QY1Avg - fnlfa * GOAreaIE4 * E1SQX - GHAvgOAreaIE2 * PR1X + WSXAvg -
                  BSXAvg + MX + Tau0QXAvg;
    float JYAvg = -CorifAvg * QX1Avg - fnlfa * GOAreaIE4 * E1SQY - GHAvgOAreaIE2 * PR1Y + WSYAvg -
                  BSYAvg + MY + Tau0QYAvg;
    JXAvg = JXAvg -
            fnlct * (QX1Avg * grad32(ua, ub, uc, FDX2, FDX3) +
                     QY1Avg * grad32(ua, ub, uc, FDY2, FDY3)) / AreaIE2 +
            fnlcat * U1Avg * ESAvg / DT;
    JYAvg = JYAvg -
            fnlct * (QX1Avg * grad32(va, vb, vc, FDX2, FDX3) +
                     QY1Avg * grad32(va, vb, vc, FDY2, FDY3)) / AreaIE2 +
            fnlcat * V1Avg * ESAvg / DT;
    const float ond = (float)q.ondiag, offd = (float)q.offdiag;
    const float FX[3] = {FDX1, FDX2, FDX3}, FY[3] = {FDY1, FDY2, FDY3};
    const float ES[3] = {ESN1, ESN2, ESN3};
    float Rk[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float others = ES[(k + 1) % 3] + ES[(k + 2) % 3];
      Rk[k] = (ond * ES[k] + offd * others) * MsFacR -
              GDP * (C00 * (FX[k] * E0XGrad2A + FY[k] * E0YGrad2A)) -
              GDP * (A00pB00 * (FX[k] * E1XGrad2A + FY[k] * E1YGrad2A)) +
              (JXAvg * FX[k] + JYAvg * FY[k]) * 0.5f + Tau0SpaVar;
    }
    R1 = Rk[0];
    R2 = Rk[1];
    R3 = Rk[2];
  }
  const float nce = (float)ncele;
  contrib[m.slot[3 * e]] = R1 * nce;
  contrib[m.slot[3 * e + 1]] = R2 * nce;
  contrib[m.slot[3 * e + 2]] = R3 * nce;
}

// QFORCE of flow-boundary entry K: for a specified (or zero) flux, codes up
// to 29, (QN2 - QN0)/DT2 + Tau0Var*QN1 (src/gwce.F:1741, 1783); for outward
// radiation (30) the value k_flux_shift formed at the step's start.
__device__ __forceinline__ double qforce(const DevMesh& m, const DevState& s, int K) {
  if (m.lbcodei[K] == 30) return s.qrad[K];
  const double q = (s.qn2[K] - s.qn0[K]) / (m.dt * 2.0) + m.tau0[m.nbv[K]] * s.qn1[K];
  // river radiation (52) past the flux-settling steps (src/gwce.F:1767-1775)
  return m.lbcodei[K] == 52 ? q - s.qrad[K] : q;
}

// The specified-flux line integrals a node receives (src/gwce.F:1777-1832):
// for each boundary edge NBV(J-1)-NBV(J) it lies on, in edge order,
// NCI*NCJ*BNDLEN2O3(J-1)/4 * (2 QFORCE(own) + QFORCE(other)) added to its
// GWCE_LV.  The edges whose terms are exactly +0 (zero length) are left out.
__device__ __forceinline__ double flux_terms(const DevMesh& m, const DevState& s, int k, double acc) {
  for (int p = m.fb_ptr[k]; p < m.fb_ptr[k + 1]; ++p) {
    const int e = m.fb_ent[p];
    const int J = e >> 1;
    const int ni = m.nbv[J - 1], nj = m.nbv[J];
    const double qi = qforce(m, s, J - 1), qj = qforce(m, s, J);
    const double w = (double)(s.nodecode[ni] * s.nodecode[nj]) * m.bndlen[J - 1] / 4.0;
    acc = acc + ((e & 1) ? w * (2.0 * qj + qi) : w * (2.0 * qi + qj));
  }
  return acc;
}

// The shift of the normal flux to the previous time levels and the step's
// new QN2 (src/timestep.F:605-612, then 861-903): next == null leaves QN2
// zero, as the reference's shift does.  An outward-radiating entry (30) takes
// QN1 from the last velocities and depth (src/timestep.F:929-940) and its
// QFORCE, -sqrt(G H1) ETAS/DT - Tau0Var QN1 with H1 the depth the GWCE will
// hold (the last step's H2) and ETAS the last increment (src/gwce.F:1750-1754,
// 1785-1789), is formed here, before the assembly rewrites ETAS.
__global__ __launch_bounds__(256) void k_flux_shift(DevMesh m, DevState s,
                                                    const double* __restrict__ next, int nvel) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nvel) return;
  s.qn0[k] = s.qn1[k];
  double q1 = s.qn2[k];
  s.qn2[k] = next ? next[k] : 0.0;
  if (m.lbcodei[k] == 30) {
    const int b = m.nbv[k];
    const double un1 = s.uu1[b] * m.csii[k] + s.vv1[b] * m.siii[k];
    q1 = s.h2[b] * un1;
    const double celerity = __builtin_sqrt(m.g * s.h2[b]);
    s.qrad[k] = -celerity * s.etas[b] / m.dt - m.tau0[b] * q1;
  } else if (m.lbcodei[k] == 52) {
    // Celerity*(ETAS/DT + Tau0Var*(ETA1 - ElevDisc)): ETA2 now is the GWCE's ETA1
    const int b = m.nbv[k];
    const double celerity = __builtin_sqrt(m.g * s.h2[b]);
    s.qrad[k] = celerity * (s.etas[b] / m.dt + m.tau0[b] * (s.eta2[b] - s.elevdisc[k]));
  }
  s.qn1[k] = q1;
}

// Node gather of GWCE_LV, the time-level shift of ETA and the elevation
// boundary conditions (src/gwce.F:1613-1631, 1638-1651, 1846-1855).  `eta1`
// already points at the previous ETA2; `eta2` is rewritten.
template <bool LUMPED, typename CT>
__global__ __launch_bounds__(256) void k_rhs_node(DevMesh m, DevState s, const CT* __restrict__ contrib,
                           const double* __restrict__ eta_bc, const double* __restrict__ scal,
                           JcgPre pre) {
  // the block's slot range staged in LDS, as in k_mom_node
  __shared__ CT sc[NODE_STAGE];
  const int r0 = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
  const int i = r0 + threadIdx.x;
  if (i < pre.barrier_words) pre.barrier[i] = 0u;
  const int p0 = m.inc_ptr[r0];
  const int n_st = m.inc_ptr[min(r0 + (int)blockDim.x, m.np)] - p0;
  const bool staged = n_st <= NODE_STAGE;
  if (staged)
    for (int k = threadIdx.x; k < n_st; k += blockDim.x) sc[k] = contrib[p0 + k];
  __syncthreads();
  if (i >= m.np) return;
  const int bi = m.bnd_index[i];
  double rhs, u;
  if (bi >= 0) {
    const double e2 = eta_bc[bi];
    const double es = e2 - s.eta1[i];
    if (pre.keep_prev) pre.etas_prev[i] = s.etas[i];
    s.eta2[i] = e2;
    u = es;
    // lumped: the boundary row's own COEFD (src/gwce.F:1856-1862)
    rhs = es * (double)s.nodecode[i] * (LUMPED ? s.coef[i] : scal[S_EP]);
  } else {
    double acc = 0.0;
    if (sizeof(CT) == 4) {
      // FP32 contributions: compensated sum
      Kahan32 k;
      if (staged)
        for (int p = m.inc_ptr[i] - p0; p < m.inc_ptr[i + 1] - p0; ++p) k.add((float)sc[p]);
      else
        for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) k.add((float)contrib[p]);
      acc = k.value();
    } else if (staged) {
      for (int p = m.inc_ptr[i] - p0; p < m.inc_ptr[i + 1] - p0; ++p) acc = acc + sc[p];
    } else {
      for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) acc = acc + contrib[p];
    }
    if (bi < -1) acc = flux_terms(m, s, -2 - bi, acc);
    for (int p = m.obcg_ptr[i]; p < (LUMPED ? m.obcg_ptr[i] : m.obcg_ptr[i + 1]); ++p) {
      const int slot = m.obcg[p];
      const int I = slot / m.mnei;
      const int b = m.nbd[I];
      const double es = eta_bc[I] - s.eta1[b];
      acc = acc - es * s.obc[slot];
    }
    rhs = acc;
    u = s.etas[i];
    const double u_ref = u;
    if (pre.warm == 1) {
      // the increment extrapolated from the last two: ETA2 extrapolated to
      // second order instead of the first order the previous increment gives
      u = 2.0 * u - pre.etas_prev[i];
    }
    if (pre.keep_prev) pre.etas_prev[i] = u_ref;
    s.eta2[i] = 0.0;
  }
  // a zero diagonal (a dry row) becomes EP (src/gwce.F:737-741); then the
  // first scaling pass of JCG (src/itpackv.F:2374-2393): w = 1/sqrt(diag),
  // u and rhs scaled in place
  double dg = s.coef[i];
  if (dg == 0.0) {
    dg = scal[S_EP];
    s.coef[i] = dg;
  }
  if (LUMPED) {
    // explicit solve (src/gwce.F:2009-2022); no JCG follows
    if (dg < 0.0) atomicOr(&pre.flags[F_ERROR], ERR_DIAG);
    if (i == 0) {
      pre.flags[F_NUMITR] = 0;
      pre.flags[F_IER] = 0;
    }
    s.gwce_lv[i] = rhs;
    const double rdiag = dg == 0.0 ? 0.0 : 1.0 / dg;
    const double es = rhs * rdiag;
    s.etas[i] = es;
    s.eta2[i] = (double)s.nodecode[i] * (es + 0.0) + s.eta1[i];
    return;
  }
  if (!(dg > 0.0)) atomicOr(&pre.flags[F_ERROR], ERR_DIAG);
  const double sq = __builtin_sqrt(dg);
  const double w = 1.0 / sq;
  s.etas[i] = u * sq;
  // the learned start (zero correction on boundary and dry rows), scaled
  if (pre.warm == 2) pre.ml_u[i] = (u + pre.ml_corr[i]) * sq;
  pre.w[i] = w;
  s.gwce_lv[i] = rhs * w;
}
template __global__ void k_rhs_node<false, double>(DevMesh, DevState, const double*, const double*, const double*, JcgPre);
template __global__ void k_rhs_node<true, double>(DevMesh, DevState, const double*, const double*, const double*, JcgPre);
template __global__ void k_rhs_node<false, float>(DevMesh, DevState, const float*, const double*, const double*, JcgPre);
template __global__ void k_rhs_node<true, float>(DevMesh, DevState, const float*, const double*, const double*, JcgPre);

// ---------------------------------------------------------------- wet/dry
// D1 (src/wetdry.F:234-285); also NNODECODE := NODECODE for the W1 writes.
__global__ __launch_bounds__(256) void k_wd_d1(DevMesh m, DevState s, int* flags, double h0, double habsmin) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  // this step's ALPHAL pass sets it again where needed
  if (i == 0) flags[F_ALPHA] = 0;
  if (i >= m.np) return;
  int nc = s.nodecode[i];
  if (nc == 1) {
    const double htot = m.dp[i] + s.eta2[i];
    if (htot <= h0) {
      if (htot < habsmin) s.eta2[i] = habsmin - m.dp[i];
      nc = 0;
      s.nodecode[i] = 0;
      atomicOr(&flags[F_NCCHANGE], 1);
    }
  }
  s.nnodecode[i] = nc;
}

// W1 candidates (src/wetdry.F:332-818) and DE1 (src/wetdry.F:1110-1233).
// Candidate elements whose steady-flow velocity exceeds VELMIN wet their dry
// node here (an idempotent store).  Their friction updates depend on element
// order: each candidate leaves its TK_TMP value and local vertices at its
// element and sets its bit in the candidate mask (one word per wave), and the
// block's candidate count goes to bcnt[tile] for the compaction below.
__global__ __launch_bounds__(256) void k_wd_elem(DevMesh m, DevState s, WdPrm q, WdWork w) {
  __shared__ int wcount[4];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int e = tile * 256 + threadIdx.x;
  bool cand = false;
  if (e < m.ne) {
    const int n[3] = {m.nm[3 * e], m.nm[3 * e + 1], m.nm[3 * e + 2]};
    s.noffold[e] = s.noff[e];
    int noff = 1;
    const double e1 = s.eta2[n[0]], e2 = s.eta2[n[1]], e3 = s.eta2[n[2]];
    const double h1 = m.dp[n[0]] + e1, h2 = m.dp[n[1]] + e2, h3 = m.dp[n[2]] + e3;
    const int c0 = s.nodecode[n[0]], c1 = s.nodecode[n[1]], c2 = s.nodecode[n[2]];
    if (c0 + c1 + c2 == 2) {
      int a = -1, b = -1, c = -1;
      if (c0 == 1 && c1 == 1) { a = 0; b = 1; c = 2; }
      else if (c1 == 1 && c2 == 1) { a = 1; b = 2; c = 0; }
      else if (c2 == 1 && c0 == 1) { a = 2; b = 0; c = 1; }
      const double ht[3] = {h1, h2, h3};
      if (a >= 0 && ht[a] >= q.hoff && ht[b] >= q.hoff) {
        const int na = n[a], nb = n[b], nc = n[c];
        int l123 = a;
        if (s.eta2[nb] > s.eta2[na]) l123 = b;
        const int nm123 = n[l123];
        const double dy = m.y[nc] - m.y[nm123], dx = m.x[nc] - m.x[nm123];
        const double deldist = __builtin_sqrt(dy * dy + dx * dx);
        double deleta = s.eta2[nm123] - s.eta2[nc];
        if (deleta < 0.0) deleta = 0.0;
        const double H1 = s.eta2[nm123] + m.dp[nm123];
        const double fr = s.fric[nm123];
        const double hyb = hybrid_term(q.hyb, H1);
        const double tkwet = fmax(fr * (q.iflinbf + (q.velmin / H1) * (q.ifnlbf + hyb)), 0.0001);
        const double vel = m.g * (deleta / deldist) / tkwet;
        if (vel > q.velmin) {
          s.nnodecode[nc] = 1;
          w.tkc[e] = fmin(fr * (q.iflinbf + (1.0 * vel / H1) * (q.ifnlbf + hyb)), q.tkmax);
          w.lc[e] = l123 | ((l123 == b ? a : b) << 2);
          cand = true;
        }
      }
    }
    const double HOFF = q.hoff;
    if ((e1 >= e2) && (e2 > e3)) if ((h1 < HOFF) || (h2 < HOFF)) noff = 0;
    if ((e2 >= e3) && (e3 > e1)) if ((h2 < HOFF) || (h3 < HOFF)) noff = 0;
    if ((e3 >= e1) && (e1 > e2)) if ((h3 < HOFF) || (h1 < HOFF)) noff = 0;
    if ((e1 >= e3) && (e3 > e2)) if ((h1 < HOFF) || (h3 < HOFF)) noff = 0;
    if ((e2 >= e1) && (e1 > e3)) if ((h2 < HOFF) || (h1 < HOFF)) noff = 0;
    if ((e3 >= e2) && (e2 > e1)) if ((h3 < HOFF) || (h2 < HOFF)) noff = 0;
    s.noff[e] = noff;
  }
  // every lane of the block reaches the ballot; a wave covers elements
  // 64*word .. 64*word+63
  const unsigned long long bits = __ballot(cand);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    if (tile * 256 + wave * 64 < m.ne) w.mask[tile * 4 + wave] = bits;
    wcount[wave] = __popcll(bits);
  }
  __syncthreads();
  if (threadIdx.x == 0) w.bcnt[tile] = (wcount[0] + wcount[1]) + (wcount[2] + wcount[3]);
}

// Exclusive prefix of the per-tile candidate counts (one block of 1024) and
// the step's candidate count.
__global__ __launch_bounds__(1024) void k_wd_index(const int* __restrict__ bcnt, int nb,
                                                   int* __restrict__ bofs, int* flags) {
  __shared__ int wtot[16];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int lo = min(t * per, nb), hi = min(lo + per, nb);
  int sum = 0;
#pragma unroll 8
  for (int i = lo; i < hi; ++i) sum += bcnt[i];
  int incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(incl, off, 64);
    if ((t & 63) >= off) incl += v;
  }
  if ((t & 63) == 63) wtot[t >> 6] = incl;
  __syncthreads();
  int base = 0;
  for (int i = 0; i < (t >> 6); ++i) base += wtot[i];
  int run = base + incl - sum;
#pragma unroll 8
  for (int i = lo; i < hi; ++i) {
    bofs[i] = run;
    run += bcnt[i];
  }
  if (t == 1023) {
    flags[F_NCAND] = base + incl;
    flags[F_WDSERIAL] = 0;
  }
}

// Candidate index of candidate element e (element order); `word` = mask[e/64].
__device__ __forceinline__ int wd_index(const WdWork& w, int e, unsigned long long word) {
  const int tile = e >> 8, wi = (e >> 6) & 3;
  int k = w.bofs[tile];
  for (int j = 0; j < wi; ++j) k += __popcll(w.mask[tile * 4 + j]);
  return k + __popcll(word & ((1ull << (e & 63)) - 1ull));
}

// Does candidate code 3*e2+k2 (element e2, local vertex k2) update the TK of
// that vertex, i.e. is e2 a candidate and the vertex one of its two wet nodes?
__device__ __forceinline__ bool wd_touches(const WdWork& w, int e2, int k2) {
  if (!((w.mask[e2 >> 6] >> (e2 & 63)) & 1ull)) return false;
  const int lc = w.lc[e2];
  return k2 == (lc & 3) || k2 == ((lc >> 2) & 3);
}

// Ordered W1 friction updates (src/wetdry.F:447-472), links.  The reference
// visits the candidates in ascending element order; candidate k sets
// TK(NM123) = MAX(TK_TMP, TK(NM123)), gives the partner the larger of the two
// and takes it back if the partner was larger, so afterwards both of its
// nodes hold M_k = max(TK_TMP_k, their values before k).  A node's value
// before k is M_j of the nearest earlier candidate j on that node, or its
// value on entry.  The node incidence lists are sorted by element, so j and
// "k is the node's last candidate" are found by walking the node's list
// backward and forward from k's own slot.  Entry values fold into m0.
__global__ __launch_bounds__(256) void k_wd_pred(DevMesh m, DevState s, WdWork w) {
  const int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const unsigned long long word = w.mask[e >> 6];
  if (!((word >> (e & 63)) & 1ull)) return;
  const int k = wd_index(w, e, word);
  const int lc = w.lc[e];
  double v = w.tkc[e];
  int pred[2], wn[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int l = (lc >> (2 * j)) & 3;
    const int x = m.nm[3 * e + l];
    const int pos = m.slot[3 * e + l];
    const int p0 = m.inc_ptr[x], p1 = m.inc_ptr[x + 1];
    int pk = -1;
    bool last = true;
    if (p1 - p0 <= WD_INC_MAX) {
      // the node's whole list at once: its codes, then their mask words,
      // then the candidates' vertex codes, each round of loads independent
      int code[WD_INC_MAX];
#pragma unroll
      for (int q = 0; q < WD_INC_MAX; ++q) code[q] = p0 + q < p1 ? m.inc[p0 + q] : -1;
      unsigned long long mw[WD_INC_MAX];
#pragma unroll
      for (int q = 0; q < WD_INC_MAX; ++q) mw[q] = code[q] >= 0 ? w.mask[(code[q] / 3) >> 6] : 0ull;
      unsigned tch = 0;  // bit q: entry q's element is a candidate that updates node x
      int lq[WD_INC_MAX];
#pragma unroll
      for (int q = 0; q < WD_INC_MAX; ++q) {
        const int e2 = code[q] / 3;
        const bool cand = code[q] >= 0 && ((mw[q] >> (e2 & 63)) & 1ull);
        lq[q] = cand ? w.lc[e2] : -1;
      }
#pragma unroll
      for (int q = 0; q < WD_INC_MAX; ++q) {
        const int k2 = code[q] - 3 * (code[q] / 3);
        if (lq[q] >= 0 && (k2 == (lq[q] & 3) || k2 == ((lq[q] >> 2) & 3))) tch |= 1u << q;
      }
      const int me = pos - p0;
      const unsigned before = tch & ((1u << me) - 1u), after = tch >> (me + 1);
      if (before) {
        const int q = 31 - __clz((int)before);  // the nearest earlier candidate
        const int e2 = m.inc[p0 + q] / 3;
        pk = wd_index(w, e2, w.mask[e2 >> 6]);
      }
      last = after == 0u;
    } else {
      for (int r = pos - 1; r >= p0; --r) {
        const int code = m.inc[r], e2 = code / 3;
        if (wd_touches(w, e2, code - 3 * e2)) {
          pk = wd_index(w, e2, w.mask[e2 >> 6]);
          break;
        }
      }
      for (int r = pos + 1; r < p1; ++r) {
        const int code = m.inc[r], e2 = code / 3;
        if (wd_touches(w, e2, code - 3 * e2)) {
          last = false;
          break;
        }
      }
    }
    if (pk < 0) v = fmax(v, s.tk[x]);
    pred[j] = pk;
    wn[j] = last ? x : -1;
  }
  w.m0[k] = v;
  w.pred[k] = make_int2(pred[0], pred[1]);
  w.wnode[k] = make_int2(wn[0], wn[1]);
}

// Ordered W1 friction updates, values: M_k = max(m0_k, M of k's predecessors)
// in candidate order, then each node takes the M of its last candidate.  TK
// values are finite, so max is exact and order-free, and M_k is the maximum of
// m0 over k and all of its ancestors in the predecessor graph (edges run from
// lower to higher candidate index).
//
// Block closure.  Candidates are cut into blocks of WD_BLOCK.  Inside a block,
// M_k splits into the maximum over k's ancestors inside the block (cv) and the
// set of earlier-block candidates those ancestors name as predecessors (cE,
// bits into the block's table of external names xtab).  Pointer doubling
// forms both: each candidate keeps a frontier of at most two ancestors not
// yet folded in; a round replaces frontier members by their own frontiers
// (folding in their values and name sets) whenever the result still has at
// most two members.  After every such step ancestors(k) = folded(k) plus
// ancestors(frontier(k)), and the lowest candidate with a non-empty frontier
// can always expand it, so the rounds end; along a wetting front, where
// candidates chain through shared nodes, they take about log2 of the chain
// length.  A block naming more than WD_XMAX external candidates sets
// F_WDSERIAL, and k_wd_link then forms every value serially instead.
__device__ __forceinline__ void wd_add(int* u, int& nu, int f) {
  if (f == 0xffff) return;
  for (int i = 0; i < nu; ++i)
    if (u[i] == f) return;
  u[nu++] = f;
}

__global__ __launch_bounds__(1024) void k_wd_close(WdWork w, int* flags) {
  extern __shared__ double wc_lds[];
  double* sv = wc_lds;                                                       // [2][WD_BLOCK]
  unsigned long long* sE = reinterpret_cast<unsigned long long*>(sv + 2 * WD_BLOCK);  // [2][WD_XW][WD_BLOCK]
  int* sF = reinterpret_cast<int*>(sE + 2 * WD_XW * WD_BLOCK);               // [2][WD_BLOCK]
  int* hkey = sF + 2 * WD_BLOCK;                                             // [WD_HASH]
  int* hidx = hkey + WD_HASH;                                                // [WD_HASH]
  __shared__ int xn;
  const int n = flags[F_NCAND];
  const int nblk = (n + WD_BLOCK - 1) / WD_BLOCK;
  const int l = threadIdx.x;
  for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
    const int base = b * WD_BLOCK, k = base + l;
    for (int i = l; i < WD_HASH; i += WD_BLOCK) hkey[i] = -1;
    if (l == 0) xn = 0;
    __syncthreads();
    int qq[2] = {-1, -1};
    double v = 0.0;
    if (k < n) {
      const int2 q = w.pred[k];
      qq[0] = q.x;
      qq[1] = q.y;
      v = w.m0[k];
    }
    // the block's external names: one hash slot per distinct candidate
    int slot[2] = {-1, -1};
    bool won[2] = {false, false};
    bool ovf = false;
    for (int j = 0; j < 2; ++j) {
      const int x = qq[j];
      if (x < 0 || x >= base) continue;
      int h = (int)(((unsigned)x * 2654435761u) >> 23);  // 9 bits: WD_HASH slots
      for (int probes = 0; probes < WD_HASH; ++probes) {
        const int old = atomicCAS(&hkey[h], -1, x);
        if (old == -1 || old == x) {
          won[j] = old == -1;
          slot[j] = h;
          break;
        }
        h = (h + 1) & (WD_HASH - 1);
      }
      if (slot[j] < 0) ovf = true;
    }
    __syncthreads();
    for (int j = 0; j < 2; ++j)
      if (won[j]) {
        const int i = atomicAdd(&xn, 1);
        hidx[slot[j]] = i;
        if (i < WD_XMAX) w.xtab[b * WD_XMAX + i] = qq[j];
      }
    __syncthreads();
    unsigned long long E[WD_XW] = {};
    int f0 = 0xffff, f1 = 0xffff;
    for (int j = 0; j < 2; ++j) {
      const int x = qq[j];
      if (x < 0) continue;
      if (x >= base) {
        const int li = x - base;
        if (f0 == 0xffff) f0 = li;
        else if (li != f0) f1 = li;
      } else if (slot[j] >= 0) {
        const int i = hidx[slot[j]];
        if (i < WD_XMAX) E[i >> 6] |= 1ull << (i & 63);
        else ovf = true;
      }
    }
    if (__syncthreads_or(ovf)) {
      if (l == 0) atomicOr(&flags[F_WDSERIAL], 1);
      continue;
    }
    if (l == 0) w.xcnt[b] = xn;
    int cur = 0;
    sv[l] = v;
#pragma unroll
    for (int i = 0; i < WD_XW; ++i) sE[i * WD_BLOCK + l] = E[i];
    sF[l] = f0 | (f1 << 16);
    __syncthreads();
    while (true) {
      const int nx = cur ^ 1;
      const double* cv_ = sv + cur * WD_BLOCK;
      const unsigned long long* cE_ = sE + cur * WD_XW * WD_BLOCK;
      const int* cF_ = sF + cur * WD_BLOCK;
      int F = cF_[l];
      const int a = F & 0xffff, c = (F >> 16) & 0xffff;
      // members to fold in this round: bit 0 = a, bit 1 = c
      int fold = 0;
      if (a != 0xffff) {
        const int Fa = cF_[a];
        const int Fc = c != 0xffff ? cF_[c] : -1;
        int u[4], nu = 0;
        wd_add(u, nu, Fa & 0xffff);
        wd_add(u, nu, (Fa >> 16) & 0xffff);
        if (c != 0xffff) {
          wd_add(u, nu, Fc & 0xffff);
          wd_add(u, nu, (Fc >> 16) & 0xffff);
        }
        if (nu <= 2) {
          fold = c != 0xffff ? 3 : 1;
          F = (nu > 0 ? u[0] : 0xffff) | ((nu > 1 ? u[1] : 0xffff) << 16);
        } else {
          // expand one member alone
          int ua[3], na = 0;
          wd_add(ua, na, c);
          wd_add(ua, na, Fa & 0xffff);
          wd_add(ua, na, (Fa >> 16) & 0xffff);
          if (na <= 2) {
            fold = 1;
            F = (na > 0 ? ua[0] : 0xffff) | ((na > 1 ? ua[1] : 0xffff) << 16);
          } else if (c != 0xffff) {
            int uc[3], nc = 0;
            wd_add(uc, nc, a);
            wd_add(uc, nc, Fc & 0xffff);
            wd_add(uc, nc, (Fc >> 16) & 0xffff);
            if (nc <= 2) {
              fold = 2;
              F = (nc > 0 ? uc[0] : 0xffff) | ((nc > 1 ? uc[1] : 0xffff) << 16);
            }
          }
        }
      }
      if (fold & 1) {
        v = fmax(v, cv_[a]);
#pragma unroll
        for (int i = 0; i < WD_XW; ++i) E[i] |= cE_[i * WD_BLOCK + a];
      }
      if (fold & 2) {
        v = fmax(v, cv_[c]);
#pragma unroll
        for (int i = 0; i < WD_XW; ++i) E[i] |= cE_[i * WD_BLOCK + c];
      }
      sv[nx * WD_BLOCK + l] = v;
#pragma unroll
      for (int i = 0; i < WD_XW; ++i) sE[(nx * WD_XW + i) * WD_BLOCK + l] = E[i];
      sF[nx * WD_BLOCK + l] = F;
      cur = nx;
      if (!__syncthreads_or((F & 0xffff) != 0xffff)) break;
    }
    if (k < n) {
      w.cv[k] = v;
#pragma unroll
      for (int i = 0; i < WD_XW; ++i) w.cE[(size_t)k * WD_XW + i] = E[i];
    }
  }
}

// Serial form of the ordered updates (blocks with too many external names):
// the candidates pass through LDS in chunks of WD_CHUNK; one lane forms M
// along the chunk (the previous candidate's M in a register, the last WD_RING
// values in an LDS ring, older ones read back from w.M) and the block writes
// the chunk's TK values.  `lds` holds WD_SERIAL_LDS bytes.
__device__ void wd_serial(const WdWork& w, int n, double* __restrict__ tk, double* lds) {
  double* ring = lds;                                             // [WD_RING]
  double* cm0 = ring + WD_RING;                                   // [WD_CHUNK]
  int2* cpred = reinterpret_cast<int2*>(cm0 + WD_CHUNK);          // [WD_CHUNK]
  int2* cwn = cpred + WD_CHUNK;                                   // [WD_CHUNK]
  const int t = threadIdx.x;
  double mprev = 0.0;  // M of the previous candidate (lane 0)
  for (int base = 0; base < n; base += WD_CHUNK) {
    const int cn = min(WD_CHUNK, n - base);
    for (int i = t; i < cn; i += blockDim.x) {
      cm0[i] = w.m0[base + i];
      cpred[i] = w.pred[base + i];
      cwn[i] = w.wnode[base + i];
    }
    __syncthreads();
    if (t == 0) {
      for (int i = 0; i < cn; ++i) {
        const int k = base + i;
        double v = cm0[i];
        const int2 q = cpred[i];
        if (q.x >= 0)
          v = fmax(v, q.x == k - 1 ? mprev
                                   : (q.x >= k - WD_RING ? ring[q.x & (WD_RING - 1)] : w.M[q.x]));
        if (q.y >= 0)
          v = fmax(v, q.y == k - 1 ? mprev
                                   : (q.y >= k - WD_RING ? ring[q.y & (WD_RING - 1)] : w.M[q.y]));
        ring[k & (WD_RING - 1)] = v;
        w.M[k] = v;
        mprev = v;
      }
    }
    __syncthreads();
    for (int i = t; i < cn; i += blockDim.x) {
      const double v = ring[(base + i) & (WD_RING - 1)];
      const int2 x = cwn[i];
      if (x.x >= 0) tk[x.x] = v;
      if (x.y >= 0) tk[x.y] = v;
    }
    __syncthreads();
  }
}

// Values of the external names, block after block: name i of block c is an
// earlier candidate x of block b < c, so its value is max(cv[x], values of
// the names of block b set in cE[x]), all known by then.  One block of 256
// (one lane per name); dynamic LDS of WD_LINK_LDS bytes (the name values of
// up to WD_LINK_BLOCKS blocks; beyond that they stay in w.xval).
__global__ __launch_bounds__(256) void k_wd_link(WdWork w, int* flags, double* __restrict__ tk) {
  extern __shared__ double lk_lds[];
  const int n = flags[F_NCAND];
  if (n == 0) return;
  if (flags[F_WDSERIAL]) {
    wd_serial(w, n, tk, lk_lds);
    return;
  }
  const int nblk = (n + WD_BLOCK - 1) / WD_BLOCK;
  double* X = nblk <= WD_LINK_BLOCKS ? lk_lds : w.xval;
  const int t = threadIdx.x;
  for (int c = 1; c < nblk; ++c) {
    if (t < w.xcnt[c]) {
      const int x = w.xtab[c * WD_XMAX + t];
      double v = w.cv[x];
      const double* Xb = X + (x / WD_BLOCK) * WD_XMAX;
#pragma unroll
      for (int i = 0; i < WD_XW; ++i) {
        unsigned long long E = w.cE[(size_t)x * WD_XW + i];
        while (E) {
          v = fmax(v, Xb[64 * i + __ffsll((long long)E) - 1]);
          E &= E - 1;
        }
      }
      X[c * WD_XMAX + t] = v;
    }
    __syncthreads();
  }
  if (X != w.xval)
    for (int i = t; i < nblk * WD_XMAX; i += blockDim.x) w.xval[i] = X[i];
}

// Final values and the TK of each node from its last candidate.
__global__ __launch_bounds__(256) void k_wd_final_area(DevMesh m, DevState s, WdWork w,
                                                       const int* flags,
                                                       double* __restrict__ tk,
                                                       double* __restrict__ area_slot) {
  // (1) each candidate's M_k, written as TK of its two wet nodes (skipped
  // when the serial pass already did it)
  const int n = flags[F_NCAND];
  if (!flags[F_WDSERIAL]) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
      double v = w.cv[k];
      const double* Xb = w.xval + (k / WD_BLOCK) * WD_XMAX;
#pragma unroll
      for (int i = 0; i < WD_XW; ++i) {
        unsigned long long E = w.cE[(size_t)k * WD_XW + i];
        while (E) {
          v = fmax(v, Xb[64 * i + __ffsll((long long)E) - 1]);
          E &= E - 1;
        }
      }
      const int2 x = w.wnode[k];
      if (x.x >= 0) tk[x.x] = v;
      if (x.y >= 0) tk[x.y] = v;
    }
  }
  // (2) D2, element part: NCEle (NNODECODE of the three nodes times NOFF,
  // both final since the candidate pass) and its area term NCEle*AREAS/2,
  // written to the incidence slots of the element's three nodes
  // (src/wetdry.F:1361-1380).  Independent of (1): one launch for both.
  const int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int ncele = s.nnodecode[m.nm[3 * e]] * s.nnodecode[m.nm[3 * e + 1]] *
                    s.nnodecode[m.nm[3 * e + 2]] * s.noff[e];
  const double a = (double)ncele * m.areas[e] / 2.0;
  area_slot[m.slot[3 * e]] = a;
  area_slot[m.slot[3 * e + 1]] = a;
  area_slot[m.slot[3 * e + 2]] = a;
}

// D2: MJU, TotalArea, landlocked nodes, NODECODE update, NCCHANGE, and
// H2 = DP + IFNLFA*ETA2 (src/wetdry.F:1361-1430, src/timestep.F:1097-1099).
// MJU counts the wet incident elements: AREAS > 0, so NCEle = 1 exactly when
// the slot's area term is non-zero.
__global__ __launch_bounds__(256) void k_wd_d2(DevMesh m, DevState s,
                                               const double* __restrict__ area_slot, int* flags,
                                               double ifnlfa) {
  // the block's slot range staged in LDS, as in k_mom_node
  __shared__ double sa[NODE_STAGE];
  const int r0 = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
  const int i = r0 + threadIdx.x;
  const int p0 = m.inc_ptr[r0];
  const int n_st = m.inc_ptr[min(r0 + (int)blockDim.x, m.np)] - p0;
  const bool staged = n_st <= NODE_STAGE;
  if (staged)
    for (int k = threadIdx.x; k < n_st; k += blockDim.x) sa[k] = area_slot[p0 + k];
  __syncthreads();
  if (i >= m.np) return;
  int mju = 0;
  double ta = 0.0;
  const int q0 = staged ? p0 : 0;
  const double* src = staged ? sa : area_slot;
  for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) {
    const double a = src[p - q0];
    mju += a != 0.0 ? 1 : 0;
    ta = ta + a;
  }
  int nn = s.nnodecode[i];
  if (nn == 1 && mju == 0) nn = 0;
  if (mju == 0) mju = 1;
  s.mju[i] = mju;
  s.totalarea[i] = ta;
  if (nn != s.nodecode[i]) {
    s.nodecode[i] = nn;
    atomicOr(&flags[F_NCCHANGE], 1);
  }
  s.h2[i] = m.dp[i] + ifnlfa * s.eta2[i];
}

// NOFF changes count toward NCCHANGE (src/wetdry.F:1428-1430), and
// computeAlpha (src/wetdry.F:1561-1652): ALPHAL of element e.
__device__ __forceinline__ double wd_alpha(const DevMesh& m, const DevState& s, int e, int* flags,
                                           double slim) {
  if (s.noff[e] != s.noffold[e]) atomicOr(&flags[F_NCCHANGE], 1);
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  double al = 1.0;
  const int ncele = s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e];
  if (ncele == 1) {
    const double hn1 = s.eta2[n1] + m.dp[n1], hn2 = s.eta2[n2] + m.dp[n2],
                 hn3 = s.eta2[n3] + m.dp[n3];
    if ((hn1 <= 5.0e-1) || (hn2 <= 5.0e-1) || (hn3 <= 5.0e-1)) {
      const double FDX1 = m.fdx[3 * e], FDX2 = m.fdx[3 * e + 1], FDX3 = m.fdx[3 * e + 2];
      const double FDY1 = m.fdy[3 * e], FDY2 = m.fdy[3 * e + 1], FDY3 = m.fdy[3 * e + 2];
      const double DPDXA = m.dp[n1] * FDX1 + m.dp[n2] * FDX2 + m.dp[n3] * FDX3;
      const double DPDYA = m.dp[n1] * FDY1 + m.dp[n2] * FDY2 + m.dp[n3] * FDY3;
      const double DETADXA = s.eta2[n1] * FDX1 + s.eta2[n2] * FDX2 + s.eta2[n3] * FDX3;
      const double DETADYA = s.eta2[n1] * FDY1 + s.eta2[n2] * FDY2 + s.eta2[n3] * FDY3;
      const double area = m.areas[e];
      const double DPM = __builtin_sqrt(DPDXA * DPDXA + DPDYA * DPDYA) / area;
      if (DPM > 1e-5) {
        const double DDTMP = DETADXA * (-DPDXA) + DETADYA * (-DPDYA);
        if (fabs(DDTMP) > 1e-12)
          al = slim * (__builtin_sqrt(DPDXA * DPDXA + DPDYA * DPDYA) / ((1.0 / area) * (DDTMP)));
      } else {
        const double DETAM = __builtin_sqrt(DETADXA * DETADXA + DETADYA * DETADYA) / area;
        if (fabs(DETAM) > 1e-12) al = slim / DETAM;
      }
      if (al <= 0.0) al = 1.0;
      if (al >= 1.0) al = 1.0;
    }
  }
  s.alphal[e] = al;
  // the cached GWCE matrix terms hold for ALPHAL = 1 only
  if (al != 1.0) atomicOr(&flags[F_ALPHA], 1);
  return al;
}

__global__ __launch_bounds__(256) void k_wd_alpha(DevMesh m, DevState s, int* flags, double slim) {
  int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  wd_alpha(m, s, e, flags, slim);
}

__global__ __launch_bounds__(256) void k_h2(int np, const double* __restrict__ dp, const double* __restrict__ eta2,
                     double ifnlfa, double* __restrict__ h2) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= np) return;
  h2[i] = dp[i] + ifnlfa * eta2[i];
}

// --------------------------------------------------------------- momentum
// Element loop of Mom_Eqs_New_NC (src/momentum.F:298-699): advection,
// barotropic pressure gradient, IBPV lateral stress; contributions x NCEle.
// With ALPHA the element's wet/dry ALPHAL is formed here first (the resident
// run folds k_wd_alpha into this pass; nothing reads ALPHAL in between).
template <bool ALPHA>
__global__ __launch_bounds__(256) void k_mom_elem(DevMesh m, DevState s, MomPrm q, double* __restrict__ cx,
                           double* __restrict__ cy, int* flags) {
  int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const double al = ALPHA ? wd_alpha(m, s, e, flags, 1.0e9) : s.alphal[e];
  const double THIRD = 1.0 / 3.0, DT = m.dt, GO2 = m.g / 2.0;
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  const int ncele = s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e];
  // the flags are final for this step: NCEle as the next GWCE matrix sees it
  s.ncele[e] = (unsigned char)ncele;
  // dry element: every contribution is NCEle*DT x (...) = 0 (see k_rhs_elem)
  double tx[3] = {0.0, 0.0, 0.0}, ty[3] = {0.0, 0.0, 0.0};
  if (ncele != 0) {
    const double U1N1 = s.uu1[n1], U1N2 = s.uu1[n2], U1N3 = s.uu1[n3];
    const double V1N1 = s.vv1[n1], V1N2 = s.vv1[n2], V1N3 = s.vv1[n3];
    const double H1N1 = s.h1[n1], H1N2 = s.h1[n2], H1N3 = s.h1[n3];
    const double AreaIE2 = m.areas[e];
    const double AreaIE = 0.5 * AreaIE2;
    const double FDX1 = m.fdx[3 * e], FDX2 = m.fdx[3 * e + 1], FDX3 = m.fdx[3 * e + 2];
    const double FDY1 = m.fdy[3 * e], FDY2 = m.fdy[3 * e + 1], FDY3 = m.fdy[3 * e + 2];
    const double U1Avg = THIRD * (U1N1 + U1N2 + U1N3);
    const double V1Avg = THIRD * (V1N1 + V1N2 + V1N3);
    const double DU1DXA = (U1N1 * FDX1 + U1N2 * FDX2 + U1N3 * FDX3) * 0.5;
    const double DU1DYA = (U1N1 * FDY1 + U1N2 * FDY2 + U1N3 * FDY3) * 0.5;
    const double DV1DXA = (V1N1 * FDX1 + V1N2 * FDX2 + V1N3 * FDX3) * 0.5;
    const double DV1DYA = (V1N1 * FDY1 + V1N2 * FDY2 + V1N3 * FDY3) * 0.5;
    const double EVMEle = THIRD * (m.evm[n1] + m.evm[n2] + m.evm[n3]);
    double B1N1 = al * s.eta1[n1], B2N1 = al * s.eta2[n1];
    double B1N2 = al * s.eta1[n2], B2N2 = al * s.eta2[n2];
    double B1N3 = al * s.eta1[n3], B2N3 = al * s.eta2[n3];
    if (q.nws) {
      B1N1 = B1N1 + s.pr1[n1]; B2N1 = B2N1 + s.pr2[n1];
      B1N2 = B1N2 + s.pr1[n2]; B2N2 = B2N2 + s.pr2[n2];
      B1N3 = B1N3 + s.pr1[n3]; B2N3 = B2N3 + s.pr2[n3];
    }
    if (s.tip1) {  // src/momentum.F:443-450
      B1N1 = B1N1 - s.tip1[n1]; B2N1 = B2N1 - s.tip2[n1];
      B1N2 = B1N2 - s.tip1[n2]; B2N2 = B2N2 - s.tip2[n2];
      B1N3 = B1N3 - s.tip1[n3]; B2N3 = B2N3 - s.tip2[n3];
    }
    const double DBTPDX1A = B1N1 * FDX1 + B1N2 * FDX2 + B1N3 * FDX3;
    const double DBTPDX2A = B2N1 * FDX1 + B2N2 * FDX2 + B2N3 * FDX3;
    const double DBTPDY1A = B1N1 * FDY1 + B1N2 * FDY2 + B1N3 * FDY3;
    const double DBTPDY2A = B2N1 * FDY1 + B2N2 * FDY2 + B2N3 * FDY3;
    const double DBTPDXA = 0.5 * (1.0 * DBTPDX1A + 1.0 * DBTPDX2A);
    const double DBTPDYA = 0.5 * (1.0 * DBTPDY1A + 1.0 * DBTPDY2A);
    const double UDU = U1Avg * DU1DXA, VDU = V1Avg * DU1DYA;
    const double UDV = U1Avg * DV1DXA, VDV = V1Avg * DV1DYA;
    const double DU1DX = DU1DXA / AreaIE, DU1DY = DU1DYA / AreaIE;
    const double DV1DX = DV1DXA / AreaIE, DV1DY = DV1DYA / AreaIE;
    const double NCDT = (double)ncele * DT;
    const double advx = -q.ifnlct * (UDU + VDU) - GO2 * DBTPDXA - 0.0;
    const double advy = -q.ifnlct * (UDV + VDV) - GO2 * DBTPDYA - 0.0;
    const double Hs[3] = {H1N1, H1N2, H1N3};
    const double FX[3] = {FDX1, FDX2, FDX3}, FY[3] = {FDY1, FDY2, FDY3};
  #pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double eh = EVMEle * Hs[k];
      const double lxx = eh * DU1DX, lxy = eh * DU1DY, lyx = eh * DV1DX, lyy = eh * DV1DY;
      // TANPHI = 0: the AreaIE*TANPHI*(...) terms are signed zeros
      tx[k] = NCDT * (advx - (1.5 * (lxx * FX[k] + lxy * FY[k]) - AreaIE * 0.0 * (lxy + lyx)) / Hs[k]);
      ty[k] = NCDT * (advy - (1.5 * (lyx * FX[k] + lyy * FY[k]) + AreaIE * 0.0 * (lxx - lyy)) / Hs[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int sl = m.slot[3 * e + k];
    cx[sl] = tx[k];
    cy[sl] = ty[k];
  }
}

template __global__ void k_mom_elem<false>(DevMesh, DevState, MomPrm, double*, double*, int*);
template __global__ void k_mom_elem<true>(DevMesh, DevState, MomPrm, double*, double*, int*);

// The momentum element pass in FP32 (the mixed-precision path): the terms of
// k_mom_elem with FP32 arithmetic and geometry, gradients from FP64
// differences (grad32), FP32 contributions.
template <bool ALPHA>
__global__ __launch_bounds__(256) void k_mom_elem32(DevMesh m, DevState s, MomPrm q,
                                                    float* __restrict__ cx,
                                                    float* __restrict__ cy, int* flags) {
  int e = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const double al64 = ALPHA ? wd_alpha(m, s, e, flags, 1.0e9) : s.alphal[e];
  const int n1 = m.nm[3 * e], n2 = m.nm[3 * e + 1], n3 = m.nm[3 * e + 2];
  const int ncele = s.nodecode[n1] * s.nodecode[n2] * s.nodecode[n3] * s.noff[e];
  s.ncele[e] = (unsigned char)ncele;
  float tx[3] = {0.0f, 0.0f, 0.0f}, ty[3] = {0.0f, 0.0f, 0.0f};
  if (ncele != 0) {
    const float THIRD = 1.0f / 3.0f, DT = (float)m.dt, GO2 = (float)(m.g / 2.0);
    const double ua = s.uu1[n1], ub = s.uu1[n2], uc = s.uu1[n3];
    const double va = s.vv1[n1], vb = s.vv1[n2], vc = s.vv1[n3];
    const float AreaIE2 = m.areas32[e];
    const float AreaIE = 0.5f * AreaIE2;
    const float FDX1 = m.fdx32[3 * e], FDX2 = m.fdx32[3 * e + 1], FDX3 = m.fdx32[3 * e + 2];
    const float FDY1 = m.fdy32[3 * e], FDY2 = m.fdy32[3 * e + 1], FDY3 = m.fdy32[3 * e + 2];
    const float U1Avg = THIRD * ((float)ua + (float)ub + (float)uc);
    const float V1Avg = THIRD * ((float)va + (float)vb + (float)vc);
    const float DU1DXA = grad32(ua, ub, uc, FDX2, FDX3) * 0.5f;
    const float DU1DYA = grad32(ua, ub, uc, FDY2, FDY3) * 0.5f;
    const float DV1DXA = grad32(va, vb, vc, FDX2, FDX3) * 0.5f;
    const float DV1DYA = grad32(va, vb, vc, FDY2, FDY3) * 0.5f;
    const float EVMEle = THIRD * ((float)m.evm[n1] + (float)m.evm[n2] + (float)m.evm[n3]);
    double b1a = al64 * s.eta1[n1], b2a = al64 * s.eta2[n1];
    double b1b = al64 * s.eta1[n2], b2b = al64 * s.eta2[n2];
    double b1c = al64 * s.eta1[n3], b2c = al64 * s.eta2[n3];
    if (q.nws) {
      b1a = b1a + s.pr1[n1]; b2a = b2a + s.pr2[n1];
      b1b = b1b + s.pr1[n2]; b2b = b2b + s.pr2[n2];
      b1c = b1c + s.pr1[n3]; b2c = b2c + s.pr2[n3];
    }
    if (s.tip1) {
      b1a = b1a - s.tip1[n1]; b2a = b2a - s.tip2[n1];
      b1b = b1b - s.tip1[n2]; b2b = b2b - s.tip2[n2];
      b1c = b1c - s.tip1[n3]; b2c = b2c - s.tip2[n3];
    }
    const float DBTPDXA = 0.5f * (grad32(b1a, b1b, b1c, FDX2, FDX3) +
                                  grad32(b2a, b2b, b2c, FDX2, FDX3));
    const float DBTPDYA = 0.5f * (grad32(b1a, b1b, b1c, FDY2, FDY3) +
                                  grad32(b2a, b2b, b2c, FDY2, FDY3));
    const float fnlct = (float)q.ifnlct;
    const float advx = -fnlct * (U1Avg * DU1DXA + V1Avg * DU1DYA) - GO2 * DBTPDXA;
    const float advy = -fnlct * (U1Avg * DV1DXA + V1Avg * DV1DYA) - GO2 * DBTPDYA;
    const float DU1DX = DU1DXA / AreaIE, DU1DY = DU1DYA / AreaIE;
    const float DV1DX = DV1DXA / AreaIE, DV1DY = DV1DYA / AreaIE;
    const float NCDT = (float)ncele * DT;
    const float Hs[3] = {(float)s.h1[n1], (float)s.h1[n2], (float)s.h1[n3]};
    const float FX[3] = {FDX1, FDX2, FDX3}, FY[3] = {FDY1, FDY2, FDY3};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float eh = EVMEle * Hs[k];
      const float lxx = eh * DU1DX, lxy = eh * DU1DY, lyx = eh * DV1DX, lyy = eh * DV1DY;
      tx[k] = NCDT * (advx - (1.5f * (lxx * FX[k] + lxy * FY[k])) / Hs[k]);
      ty[k] = NCDT * (advy - (1.5f * (lyx * FX[k] + lyy * FY[k])) / Hs[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int sl = m.slot[3 * e + k];
    cx[sl] = tx[k];
    cy[sl] = ty[k];
  }
}
template __global__ void k_mom_elem32<false>(DevMesh, DevState, MomPrm, float*, float*, int*);
template __global__ void k_mom_elem32<true>(DevMesh, DevState, MomPrm, float*, float*, int*);

// Nodal part of Mom_Eqs_New_NC (src/momentum.F:713-1182): gather, /TotalArea,
// wind, friction/Coriolis 2x2, land-boundary constraints, solve, fluxes;
// then the peak-surge maximum (src/write_output.F:6357-6366).  FP32
// contributions (the mixed path) are summed with compensation.
template <typename CT>
__global__ __launch_bounds__(256) void k_mom_node(DevMesh m, DevState s, MomPrm q, const CT* __restrict__ cx,
                           const CT* __restrict__ cy) {
  // the block's 256 nodes own one contiguous range of incidence slots: it is
  // read once, coalesced, into LDS, and each node then sums its own slots
  // from there in order (the same sums as reading them directly)
  __shared__ CT sx[NODE_STAGE], sy[NODE_STAGE];
  const int r0 = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x;
  const int i = r0 + threadIdx.x;
  const int p0 = m.inc_ptr[r0];
  const int n_st = m.inc_ptr[min(r0 + (int)blockDim.x, m.np)] - p0;
  const bool staged = n_st <= NODE_STAGE;
  if (staged) {
    for (int k = threadIdx.x; k < n_st; k += blockDim.x) {
      sx[k] = cx[p0 + k];
      sy[k] = cy[p0 + k];
    }
  }
  __syncthreads();
  if (i >= m.np) return;
  const double DTO2 = m.dt / 2.0;
  const int nci = s.nodecode[i];
  double mx = 0.0, my = 0.0;
  if (sizeof(CT) == 4) {
    Kahan32 kx, ky;
    const CT* px = staged ? sx : cx + p0;
    const CT* py = staged ? sy : cy + p0;
    for (int p = m.inc_ptr[i] - p0; p < m.inc_ptr[i + 1] - p0; ++p) {
      kx.add((float)px[p]);
      ky.add((float)py[p]);
    }
    mx = kx.value();
    my = ky.value();
  } else if (staged) {
    for (int p = m.inc_ptr[i] - p0; p < m.inc_ptr[i + 1] - p0; ++p) {
      mx = mx + sx[p];
      my = my + sy[p];
    }
  } else {
    for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) {
      mx = mx + cx[p];
      my = my + cy[p];
    }
  }
  const double ta = s.totalarea[i];
  if (ta != 0.0) {
    mx = mx / ta;
    my = my / ta;
  }
  double WSX = 0.0, WSY = 0.0;
  if (q.nws) {
    WSX = DTO2 * 1.0 * (s.wsx1[i] / s.h1[i] + s.wsx2[i] / s.h2[i]);
    WSY = DTO2 * 1.0 * (s.wsy1[i] / s.h1[i] + s.wsy2[i] / s.h2[i]);
  }
  const double tk = s.tk[i];
  const double VXX = DTO2 * (tk + 0.0), VYY = DTO2 * (tk + 0.0);
  const double corifa = m.corif[i] + q.ifnlct * 0.0 * s.uu1[i];
  const double VXY = DTO2 * (0.0 - corifa), VYX = DTO2 * (0.0 + corifa);
  const double u1 = s.uu1[i], v1 = s.vv1[i];
  const double dnci = (double)nci;
  mx = dnci * (mx + WSX + (1.0 - VXX) * u1 - VXY * v1);
  my = dnci * (my + WSY + (1.0 - VYY) * v1 - VYX * u1);
  double A0 = 1.0 + VXX * dnci, A1 = 1.0 + VYY * dnci, A2 = VXY * dnci, A3 = VYX * dnci;
  // land boundary constraints in ME2GW order (src/momentum.F:810-870).  The
  // isotropy test of :816 is ABS(NORM2(TKM) - SQRT(2)*TK) < EPSILON with
  // TKM = (TK, TK, 0): flang's scaled NORM2 gives |TK|*SQRT(2) exactly, so
  // the first branch is taken unless TK is negative (a node whose total
  // depth went negative) and not tiny.
  const double n2tkm = fabs(tk) * __builtin_sqrt(2.0);
  const bool isotropic = fabs(n2tkm - __builtin_sqrt(2.0) * tk) < 2.220446049250313e-16;
  for (int p = m.vbc_ptr[i]; p < m.vbc_ptr[i + 1]; ++p) {
    const int I = m.vbc[p];
    const int code = m.lbcodei[I];
    const double CS = m.csii[I], SI = m.siii[I];
    const double veln = -(s.qn2 ? s.qn2[I] : 0.0) / s.h2[i];
    if (code >= 0 && code <= 9) {
      if (isotropic) {
        mx = (SI * mx - CS * my - veln * A2) * dnci;
        my = veln * A0 * dnci;
        A2 = -CS * A0;
        A3 = -A2;
        A0 = SI * A0;
        A1 = A0;
      } else {
        mx = (SI * mx - CS * my) * dnci;
        my = veln * dnci;
        A0 = A0 * SI - A3 * CS;
        A2 = A2 * SI - A1 * CS;
        A3 = CS;
        A1 = SI;
      }
    } else if (code >= 10 && code <= 19) {
      mx = 0.0 * dnci;
      my = veln * dnci;
      A0 = SI;
      A2 = -CS;
      A3 = CS;
      A1 = SI;
    }
  }
  const double ddu = A0 * A1 - A2 * A3;
  const double u2 = (mx * A1 - my * A2) / ddu;
  const double v2 = (my * A0 - mx * A3) / ddu;
  s.uu2[i] = u2;
  s.vv2[i] = v2;
  const double h2 = s.h2[i];
  s.qx2[i] = u2 * h2;
  s.qy2[i] = v2 * h2;
  const double e2 = s.eta2[i];
  // run-time checks of the reference: ErrorElev over the wet nodes
  // (src/timestep.F:1242-1251, 1381-1395) and a NaN velocity anywhere
  // (src/timestep.F:1413-1471; an infinite one counts too).  The first step
  // that raises either is kept.
  const bool elev_bad = nci == 1 && fabs(e2) > q.errorelev;
  const bool vel_bad = !(fabs(u2) <= 1.7976931348623157e308 && fabs(v2) <= 1.7976931348623157e308);
  if (elev_bad || vel_bad) {
    atomicOr(&q.flags[F_ERROR], (elev_bad ? ERR_ELEV : 0) | (vel_bad ? ERR_NONFINITE : 0));
    atomicCAS(&q.flags[F_ERRSTEP], 0, q.it > 0 ? q.it : 1);
  }
  // peak surge and its step (src/write_output.F:6357-6364; the host turns
  // the step into ETAMAX_Time = IT*DTDP + STATIM*86400)
  if (nci == 1 && e2 > s.etamax[i]) {
    s.etamax[i] = e2;
    s.etamax_it[i] = q.it;
  }
  // peak speed and its step (src/write_output.F:6367-6376)
  if (nci == 1) {
    const double ut = __builtin_sqrt(u2 * u2 + v2 * v2);
    if (ut > s.umax[i]) {
      s.umax[i] = ut;
      s.umax_it[i] = q.it;
    }
  }
}

template __global__ void k_mom_node<double>(DevMesh, DevState, MomPrm, const double*, const double*);
template __global__ void k_mom_node<float>(DevMesh, DevState, MomPrm, const float*, const float*);

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, unsigned v, int n) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_i32(int* p, int v, int n) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// TotalArea / MJU at cold start (src/mesh.F:3360-3400), node gather.
__global__ __launch_bounds__(256) void k_total_area(DevMesh m, DevState s) {
  int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= m.np) return;
  int mju = 0;
  double ta = 0.0;
  for (int p = m.inc_ptr[i]; p < m.inc_ptr[i + 1]; ++p) {
    const int e = m.inc[p] / 3;
    const int ncele = s.nodecode[m.nm[3 * e]] * s.nodecode[m.nm[3 * e + 1]] *
                      s.nodecode[m.nm[3 * e + 2]] * s.noff[e];
    mju += ncele;
    ta = ta + (double)ncele * m.areas[e] / 2.0;
  }
  if (mju == 0) mju = 1;
  s.mju[i] = mju;
  s.totalarea[i] = ta;
}

}  // namespace adcgpu
