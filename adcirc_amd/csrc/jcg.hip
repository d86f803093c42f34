// ITPACK jcg (src/itpackv.F:63-456) for one persistent cooperative launch.
//
// Launch: blocks of JCG_CPB x 512 threads, each 512-thread group owning one
// 2048-row chunk, plus one eigenvalue block.
// Row thread t of chunk c owns rows c*2048 + t + 512*r (r = 0..3) and keeps
// their iterates (d, d_old, u, u_old) and the product N d in registers for
// the whole solve; only d is published to memory, for the neighbours'
// gathers.  The eigenvalue block runs chgcon/zbrent (the adaptive estimate of
// the largest eigenvalue, a serial root search) while the other blocks form
// the matrix-vector product, and publishes CME with the chunk partials.
// After each grid barrier every block reduces the chunk partials with the
// same fixed tree and evaluates the scalar recurrences (pstop, parcon)
// redundantly, so all blocks take identical decisions and leave together.
// Two grid barriers per iteration: after the product (partials and CME
// complete) and after the update (d complete).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "reduce.h"

namespace adcgpu {

// Grid barrier for cooperative launches whose blocks are all resident.
// Arrival is counted in two levels, so that no counter sees more than a few
// tens of atomic adds: blocks add to the counter of their group of
// JCG_BAR_GROUP, the last arrival of a group adds to the top counter, and the
// last arrival there publishes the barrier's generation number, which every
// block polls with relaxed loads.
// Memory ordering: the workgroup barrier orders every wave's stores before
// thread 0's agent-scope release fence (write-back of the XCD's L2), and
// thread 0's agent-scope acquire fence (invalidate) after the wait before
// every wave's later loads; the group- and top-level last arrivals pass the
// release on with acquire-release fences (fences synchronise through the
// release sequence of the counters' atomic adds).  The L2 write-back and
// invalidate are serialised per XCD, so they are issued once per block, not
// once per wave (measured on MI355X with 512 blocks: 15 us per barrier
// against 41 us), and never per poll.
// The wait is bounded in wall-clock time and a block that sees the error
// flag leaves at once, so a launch that went wrong drains quickly.  Returns
// false on time-out or error (a logic error; the caller leaves the kernel).
// `bid` and `nb` are the block's index and the number of blocks of the
// system it works on (the whole grid, or one system of a batched launch).
// `scoped`: the block reads other blocks' data after the barrier only with
// agent-scope atomic loads, which are coherent at agent scope by themselves
// (they miss the CU cache and the XCD's non-coherent L2 lines; the barrier's
// own polls rely on the same) and are issued after the poll that saw the new
// generation returned.  The L2 invalidate of the acquire fence (about a
// microsecond per block on MI355X) is then left out; only the compiler is
// kept from moving loads above the wait.
__device__ __forceinline__ bool grid_sync(unsigned* bar, unsigned& gen, int* err,
                                          long long ticks, unsigned bid, unsigned nb,
                                          bool scoped = false) {
  __shared__ int failed;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    failed = 0;
    ++gen;
    const unsigned ng = (nb + JCG_BAR_GROUP - 1) / JCG_BAR_GROUP;
    const unsigned g = bid / JCG_BAR_GROUP;
    const unsigned gcount = (g == ng - 1) ? nb - g * JCG_BAR_GROUP : JCG_BAR_GROUP;
    unsigned* genw = bar + JCG_BAR_STRIDE * (1 + ng);
    const unsigned o1 = __hip_atomic_fetch_add(bar + JCG_BAR_STRIDE * (1 + g), 1u,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (o1 == gen * gcount - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
      const unsigned o2 =
          __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o2 == gen * ng - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        __hip_atomic_store(genw, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    const long long t0 = wall_clock64();
    unsigned polls = 0;
    // a bound of zero allows no wait at all: every barrier fails at once (the
    // host forces the error path this way)
    if (ticks <= 0) {
      atomicOr(err, ERR_BARRIER);
      failed = 1;
    }
    while (!failed &&
           __hip_atomic_load(genw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen) {
      __builtin_amdgcn_s_sleep(1);
      if ((++polls & 63u) == 0) {
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_BARRIER) {
          failed = 1;
          break;
        }
        if (wall_clock64() - t0 > ticks) {
          atomicOr(err, ERR_BARRIER);
          failed = 1;
          break;
        }
      }
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_BARRIER)
      failed = 1;
    if (scoped)
      __asm__ volatile("" ::: "memory");
    else
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return failed == 0;
}

// Point-to-point wait after the update of the pseudo-residual: this tile
// publishes its iteration count, then waits until the tiles lo..hi (its rows'
// neighbours, the mesh bandwidth wide) have published theirs.  Only the
// neighbours' d is read before the next grid barrier, so this replaces a
// second grid-wide barrier per iteration.  Same release / acquire pattern
// (`scoped` as there) and the same bounded wait as grid_sync.
__device__ __forceinline__ bool halo_sync(unsigned* hflags, int tile, int lo, int hi,
                                          unsigned gen, int* err, long long ticks,
                                          bool scoped = false) {
  __shared__ int failed;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(hflags + JCG_HALO_STRIDE * tile, gen, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    failed = 0;
    const long long t0 = wall_clock64();
    unsigned polls = 0;
    if (ticks <= 0) {
      atomicOr(err, ERR_BARRIER);
      failed = 1;
    }
    for (int q = lo; q <= hi && !failed; ++q) {
      if (q == tile) continue;
      while (__hip_atomic_load(hflags + JCG_HALO_STRIDE * q, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) < gen) {
        __builtin_amdgcn_s_sleep(1);
        if ((++polls & 63u) == 0) {
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_BARRIER) {
            failed = 1;
            break;
          }
          if (wall_clock64() - t0 > ticks) {
            atomicOr(err, ERR_BARRIER);
            failed = 1;
            break;
          }
        }
      }
    }
    if (scoped)
      __asm__ volatile("" ::: "memory");
    else
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return failed == 0;
}

// Element access through a 32-bit byte offset from a uniform base: the
// addresses then take the scalar-base + 32-bit-offset form of the vector
// memory instructions instead of one 64-bit VGPR pair per row.  Arrays here
// stay below 4 GiB (checked on the host).
template <class T>
__device__ __forceinline__ T& at(T* base, unsigned i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ const T& at(const T* base, unsigned i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(T));
}

// Agent-scope load of a value another block wrote (coherent across XCDs).
__device__ __forceinline__ double ld_agent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The neighbouring tiles' rows lo .. t0-1 and t1 .. hi-1 of v into the staged
// rows (dsh holds rows slo ..); a thread's loads are all issued before their
// LDS stores.
__device__ __forceinline__ void stage_halo(double* dsh, const double* v, int slo, int lo, int t0,
                                           int t1, int hi) {
  constexpr int K = 4;
  const int nl = t0 - lo, nh = nl + (hi - t1);
  for (int k0 = (int)threadIdx.x; k0 < nh; k0 += K * JCG_BLOCK) {
    double x[K];
    int row[K];
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const int k = k0 + m * JCG_BLOCK;
      row[m] = k < nl ? lo + k : t1 + (k - nl);
      x[m] = k < nh ? ld_agent(&at(v, (unsigned)row[m])) : 0.0;
    }
#pragma unroll
    for (int m = 0; m < K; ++m)
      if (k0 + m * JCG_BLOCK < nh) dsh[row[m] - slo] = x[m];
  }
}

struct ItLocal {
  int in, adapt, halt, q1, maxfn, nsig, itmax, stop;
  double zeta, cme, gamma, rho, delnnm, delsnm, udnm, stptst, gamold, rhoold, cmold;
  double c1, c2, c3, c4, tz;
};

// determ (src/itpackv.F:1029-1068): the backward three-term recurrence with
// the two running values kept in registers (same operations, same order).
__device__ double jcg_determ(int n, const double* tri1, const double* tri2, double x) {
  double a2 = tri1[n - 1] - x;                            // wk2(n)
  double a1 = (tri1[n - 2] - x) * a2 + tri2[n - 1];       // wk2(n-1)
  for (int l = n - 3; l >= 0; --l) {
    const double w = (tri1[l] - x) * a1 + tri2[l + 1] * a2;  // wk2(l)
    a2 = a1;
    a1 = w;
  }
  return a1;
}

// zbrent (src/itpackv.F:2521-2781) with eps = 0; returns ier, result in *pb.
__device__ int jcg_zbrent(int n, const double* tri1, const double* tri2, double t, int maxfn,
                          double a, double* pb) {
  double b = 1.0, c = 0, d = 0, e = 0, p, qq, r, ss, fa, fb, fc = 0, tol, rm, rone, temp;
  int ic = 2;
  const double eps = 0.0;
  fa = jcg_determ(n, tri1, tri2, a);
  fb = jcg_determ(n, tri1, tri2, b);
  if ((fa > 0.0 && fb > 0.0) || (fa < 0.0 && fb < 0.0)) {
    *pb = b;
    return 502;
  }
  bool reset = true;
  while (true) {
    if (reset) {
      c = a; fc = fa; d = b - c; e = d;
    }
    if (!(fabs(fc) >= fabs(fb))) {
      a = b; b = c; c = a; fa = fb; fb = fc; fc = fa;
    }
    tol = t * fmax(fabs(b), 0.1);
    rm = (c - b) * 0.5;
    if (fabs(fb) <= eps || fabs(c - b) <= tol) { *pb = b; return 0; }
    if (ic >= maxfn) { *pb = b; return 501; }
    bool bisect = (fabs(e) < tol) || (fabs(fa) <= fabs(fb));
    if (!bisect) {
      ss = fb / fa;
      if (a == c) {
        p = (c - b) * ss;
        qq = 1.0 - ss;
      } else {
        qq = fa / fc;
        r = fb / fc;
        rone = r - 1.0;
        p = ss * ((c - b) * qq * (qq - r) - (b - a) * rone);
        qq = (qq - 1.0) * rone * (ss - 1.0);
      }
      if (p > 0.0) qq = -qq;
      if (p < 0.0) p = -p;
      ss = e;
      e = d;
      if (p + p >= 3.0 * rm * qq || p + p >= fabs(ss * qq)) {
        bisect = true;
      } else {
        d = p / qq;
      }
    }
    if (bisect) { e = rm; d = e; }
    a = b;
    fa = fb;
    temp = d;
    if (fabs(temp) <= 0.5 * tol) temp = copysign(0.5 * tol, rm);
    b = b + temp;
    fb = jcg_determ(n, tri1, tri2, b);
    ic = ic + 1;
    if (fb == 0.0 || fc == 0.0) { reset = false; continue; }
    if ((fb > 0.0 && fc < 0.0) || (fb < 0.0 && fc > 0.0)) { reset = false; continue; }
    reset = true;
  }
}

// chgcon(ibmth=1) (src/itpackv.F:893-1027); tri1/tri2 are the two columns.
__device__ void jcg_chgcon(ItLocal& L, double* tri1, double* tri2) {
  const int ip = L.in;
  double start = L.cme, end1;
  if (ip == 0) {
    end1 = 0.0;
    L.cmold = 0.0;
  } else if (ip == 1) {
    end1 = 1.0 - 1.0 / L.gamma;
    tri1[0] = end1;
    tri2[0] = 0.0;
  } else {
    if (fabs(start - L.cmold) <= L.zeta * start) {
      L.adapt = 0;
      return;
    }
    L.cmold = start;
    tri1[ip - 1] = 1.0 - 1.0 / L.gamma;
    tri2[ip - 1] = (1.0 - L.rho) / (L.rho * L.rhoold * L.gamma * L.gamold);
    double b;
    int ier = jcg_zbrent(ip, tri1, tri2, L.tz, L.maxfn, start, &b);
    if (ier != 0) {
      L.adapt = 0;
      return;
    }
    end1 = b;
  }
  L.cme = end1;
}

// pstop_nrms(iflag=1) (src/itpackv.F:809-890)
__device__ void jcg_pstop(ItLocal& L, double unrm, double dnrm) {
  const double con = L.cme;
  double uold, tr, tl;
  L.halt = 0;
  if (L.in < 1) {
    L.q1 = 0;
    L.udnm = 1.0;
    L.stptst = 1000.0;
  }
  if (!(L.q1 || (L.in > 5 && (L.in % 5) != 0))) {
    uold = L.udnm;
    L.udnm = unrm;
    if (L.udnm == 0.0) L.udnm = 1.0;
    if (L.in > 5 && fabs(L.udnm - uold) <= L.udnm * L.zeta) L.q1 = 1;
  }
  tr = __builtin_sqrt(L.udnm);
  tl = 1.0;
  if (con != 1.0) {
    tl = __builtin_sqrt(dnrm);
    tr = tr * (1.0 - con);
  }
  L.stptst = tl / tr;
  if (tl >= tr * L.zeta) return;
  L.halt = 1;
}

// parcon(ibmth=1) (src/itpackv.F:1774-1851)
__device__ void jcg_parcon(ItLocal& L, double dtnrm) {
  const int ip = L.in;
  const double rhoold = L.rho;
  L.gamold = L.gamma;
  L.gamma = 1.0 / (1.0 - dtnrm / L.delnnm);
  L.rho = 1.0;
  if (ip != 0) L.rho = 1.0 / (1.0 - L.gamma * L.delnnm / (L.gamold * L.delsnm * rhoold));
  L.delsnm = L.delnnm;
  L.rhoold = rhoold;
  L.c1 = L.rho * L.gamma;
  L.c2 = L.rho;
  L.c3 = 1.0 - L.rho;
  L.c4 = L.rho * (1.0 - L.gamma);
}


// The solve's scalar state at the start of a launch: ITPACK's SAVEd values
// from the last solve, the rest cleared (thread 0).
__device__ __forceinline__ void jcg_init_state(ItLocal& L, const JcgArgs& a) {
  L.in = 0;
  L.adapt = 1;
  L.halt = 0;
  L.q1 = a.saved[0] != 0.0;
  L.gamold = a.saved[1];
  L.rhoold = a.saved[2];
  L.cmold = a.saved[3];
  L.itmax = a.itmax;
  L.zeta = a.zeta;
  L.cme = 0.0;
  L.gamma = 0.0;
  L.rho = 0.0;
  L.delnnm = 0.0;
  L.delsnm = 0.0;
  L.udnm = 0.0;
  L.stptst = 0.0;
  L.maxfn = a.maxfn;
  L.tz = a.tz;
  L.stop = 0;
}

// After iteration in's grid barrier: the chunk partials reduced by every
// block with the same fixed tree (the loads of two chunks issued before
// their sums), with the eigenvalue block's estimate for the iteration; then
// pstop and parcon (thread 0), so that every block takes the same decision.
__device__ __forceinline__ void jcg_scalars(ItLocal& L, const JcgArgs& a, double* lds, int in,
                                            int nch, bool need_u) {
  const int t = threadIdx.x;
  double cme_in = 0.0, cmold_in = 0.0;
  if (t == 0) {
    cme_in = ld_agent(a.saved + 6 + 2 * (in & 1));
    cmold_in = ld_agent(a.saved + 7 + 2 * (in & 1));
  }
  double r0 = 0.0, r1 = 0.0, r2 = 0.0;
  const double* part = a.part + (size_t)(in & 1) * nch * 4;
  if (t < RED_XTHREADS) {
    int cc = t;
    for (; cc + RED_XTHREADS < nch; cc += 2 * RED_XTHREADS) {
      const double* pa = part + (size_t)cc * 4;
      const double* pb = part + (size_t)(cc + RED_XTHREADS) * 4;
      const double a0 = ld_agent(pa), a1 = ld_agent(pa + 1), a2 = ld_agent(pa + 2);
      const double b0 = ld_agent(pb), b1 = ld_agent(pb + 1), b2 = ld_agent(pb + 2);
      r0 = (r0 + a0) + b0;
      r1 = (r1 + a1) + b1;
      r2 = (r2 + a2) + b2;
    }
    if (cc < nch) {
      const double* pa = part + (size_t)cc * 4;
      r0 = r0 + ld_agent(pa);
      r1 = r1 + ld_agent(pa + 1);
      r2 = r2 + ld_agent(pa + 2);
    }
  }
  block_tree3(r0, r1, r2, lds);
  const double delnnm = r0, dtnrm = r1, unorm_s = r2;
  if (t == 0) {
    L.in = in;
    L.cme = cme_in;
    L.cmold = cmold_in;
    L.delnnm = delnnm;
    jcg_pstop(L, need_u ? unorm_s : 1.0, delnnm);
    if (!L.halt) jcg_parcon(L, dtnrm);
  }
  __syncthreads();
}

// The eigenvalue block's part of iteration in: the adaptive estimate,
// published by iteration parity (a block may still read the last
// iteration's pair).
__device__ __forceinline__ void jcg_eig_step(ItLocal& L, const JcgArgs& a, double* tri1,
                                             double* tri2, int in) {
  if (threadIdx.x == 0) {
    L.in = in;
    if (L.adapt) jcg_chgcon(L, tri1, tri2);
    a.saved[6 + 2 * (in & 1)] = L.cme;
    a.saved[7 + 2 * (in & 1)] = L.cmold;
  }
}

// The solve's outcome and SAVEd state (block 0, thread 0).
__device__ __forceinline__ void jcg_finish(const ItLocal& L, const JcgArgs& a, bool halted) {
  const int in_final = halted ? L.in : a.itmax;
  a.flags[F_NUMITR] = in_final;
  a.flags[F_IER] = halted ? 0 : 13;
  a.saved[0] = L.q1 ? 1.0 : 0.0;
  a.saved[1] = L.gamold;
  a.saved[2] = L.rhoold;
  a.saved[3] = L.cmold;
  a.saved[4] = L.cme;
  a.saved[5] = L.stptst;
  a.flags[F_ITSUM] += in_final;
  if (!halted) a.flags[F_NITMAX] += 1;
}

// The per-row partials of one 2048-row chunk (threads g*512 .. g*512+511 of
// the block, chunk c) into the chunk's slot of the partials by parity:
// ((w0+w1)+(w2+w3))+((w4+w5)+(w6+w7)) over the chunk's 8 waves.  Every
// thread of the block calls it.
__device__ __forceinline__ void jcg_chunk_partials(const JcgArgs& a, double (*wsum)[24], int c,
                                                   int nch, int in, double add, double adt,
                                                   double auu) {
  const int t = threadIdx.x;
  const double s0 = wave_sum(add), s1 = wave_sum(adt), s2 = wave_sum(auu);
  const int tt = t % RED_THREADS;
  double* ws = wsum[t / RED_THREADS];
  if ((tt & 63) == 0) {
    ws[0 + (tt >> 6)] = s0;
    ws[8 + (tt >> 6)] = s1;
    ws[16 + (tt >> 6)] = s2;
  }
  __syncthreads();
  if (tt < 3 && c < nch) {
    const double* v = ws + 8 * tt;
    a.part[(size_t)(in & 1) * nch * 4 + (size_t)c * 4 + tt] =
        ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  }
}

// DUAL: two initial guesses, a.u (the reference's) and a.u2 (the learned
// one); the pseudo-residuals of both are formed in the scaling pass and the
// iteration starts from the one with the smaller norm (the reference's on a
// tie).  Everything after that choice is the same ITPACK iteration.
// STAGED: the products read the pseudo-residual from LDS (a.hrows >= 0).
// `bid` and `nblocks`: this block's index among the blocks that solve this
// system, and their number (jcg_grid_for(a.n)).
template <bool DUAL, bool STAGED>
__device__ __forceinline__ void jcg_solve(const JcgArgs& a, int bid, int nblocks) {
  __shared__ ItLocal L;
  __shared__ double lds[12];
  __shared__ double tri1[JCG_MAX_ITMAX], tri2[JCG_MAX_ITMAX];
  extern __shared__ double dsh[];  // staged pseudo-residual, rows slo .. t1+hrows-1
  const int np = a.n;
  const int t = threadIdx.x;
  const int nch = (np + RED_CHUNK - 1) / RED_CHUNK;
  const int nrb = (nch + JCG_CPB - 1) / JCG_CPB;  // blocks with rows
  const bool eig_block = bid == nrb;  // the last block: no rows
  // each group of RED_THREADS threads owns one chunk; rows past NP (including
  // a whole missing chunk in the last block) are skipped like padding
  const int tile = eig_block ? nrb : xcd_tile(bid, nrb);
  const int c = tile * JCG_CPB + t / RED_THREADS;
  // halo flags follow the grid barrier's counters
  unsigned* hflags =
      a.barrier + ((nblocks + JCG_BAR_GROUP - 1) / JCG_BAR_GROUP + 2) * JCG_BAR_STRIDE;
  const int hlo = a.halo >= 0 ? (tile - a.halo > 0 ? tile - a.halo : 0) : 0;
  const int hhi = a.halo >= 0 ? (tile + a.halo < nrb - 1 ? tile + a.halo : nrb - 1) : 0;
  const int base = c * RED_CHUNK + t % RED_THREADS;
  unsigned gen = 0;
  int* err = a.flags + F_ERROR;
  // the staged rows: the tile's own rows t0 .. t1-1 and hrows each side
  const bool staged = STAGED && !eig_block;
  const int t0 = tile * JCG_CPB * RED_CHUNK;
  const int t1 = t0 + JCG_CPB * RED_CHUNK < np ? t0 + JCG_CPB * RED_CHUNK : np;
  const int slo = staged ? (t0 - a.hrows > 0 ? t0 - a.hrows : 0) : 0;
  // the rows the tile's rows reference, within slo .. t1 + hrows - 1
  const int hlo_row = staged ? a.hrange[2 * tile] : 0;
  const int hhi_row = staged ? a.hrange[2 * tile + 1] : 0;

  if (t == 0) jcg_init_state(L, a);
  __syncthreads();

  // per-row iterates, register resident
  double dc[RED_PER_THREAD], dp[RED_PER_THREAD], uc[RED_PER_THREAD], up[RED_PER_THREAD];
  double nd[RED_PER_THREAD];
  // neighbour counts of the rows, one byte each (MNEI <= 255)
  unsigned nnp[(RED_PER_THREAD + 3) / 4];
#pragma unroll
  for (int k = 0; k < (RED_PER_THREAD + 3) / 4; ++k) nnp[k] = 0x01010101u;
#define NN(r) ((int)((nnp[(r) >> 2] >> (8 * ((r) & 3))) & 255u))

  // ---- scal, part 1: diagonal scaling (src/itpackv.F:2374-2393), unless
  // the caller's last pass already did it (then w, u and rhs are scaled and
  // every row's w is in memory before this launch)
  if (!eig_block) {
#pragma unroll
    for (int r = 0; r < RED_PER_THREAD; ++r) {
      const int i = base + RED_THREADS * r;
      uc[r] = 0.0;
      if (i < np) {
        if (a.prescaled) {
          uc[r] = at(a.u, i);
        } else {
          const double dg = at(a.coef, i);
          if (!(dg > 0.0)) atomicOr(err, ERR_DIAG);
          const double sq = __builtin_sqrt(dg);
          const double u0 = at(a.u, i) * sq;
          at(a.u, i) = u0;
          uc[r] = u0;
          const double w = 1.0 / sq;
          at(a.w, i) = w;
          at(a.rhs, i) = at(a.rhs, i) * w;
        }
        const unsigned sh = 8 * (r & 3);
        nnp[r >> 2] = (nnp[r >> 2] & ~(255u << sh)) | ((unsigned)at(a.nneigh, i) << sh);
      }
    }
  }
  if (!a.prescaled)
    if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;

  // ---- scal, part 2: scale off-diagonals; initial pseudo-residual
  // d0 = rhs - N u - u (src/itpackv.F:2394-2400, 346-353)
  double u2c[DUAL ? RED_PER_THREAD : 1], d2c[DUAL ? RED_PER_THREAD : 1];
  if (!eig_block) {
    double wi[RED_PER_THREAD];
#pragma unroll
    for (int r = 0; r < RED_PER_THREAD; ++r) {
      const int i = base + RED_THREADS * r;
      wi[r] = 0.0;
      dc[r] = 0.0;
      if (DUAL) {
        u2c[r] = 0.0;
        d2c[r] = 0.0;
      }
      if (i < np) {
        wi[r] = at(a.w, i);
        dc[r] = at(a.rhs, i);
        if (DUAL) {
          u2c[r] = at(a.u2, i);
          d2c[r] = dc[r];
        }
      }
    }
    // the next slot's column index and coefficient are loaded while this
    // slot's w and u are gathered
#pragma unroll
    for (int h = 0; h < RED_PER_THREAD; h += 4) {
      int jm = 1;
      unsigned cn[4];
      double kn[4];
#pragma unroll
      for (int r = h; r < h + 4; ++r) {
        jm = NN(r) > jm ? NN(r) : jm;
        const unsigned q = (unsigned)(np + base + RED_THREADS * r);
        cn[r - h] = 1 < NN(r) ? (unsigned)at(a.col, q) : 0u;
        kn[r - h] = 1 < NN(r) ? at(a.coef, q) : 0.0;
      }
      for (int j = 1; j < jm; ++j) {
        unsigned cc[4];
        double kc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          cc[k] = cn[k];
          kc[k] = kn[k];
        }
        if (j + 1 < jm) {
          const unsigned q0 = (unsigned)((j + 1) * np + base);
#pragma unroll
          for (int r = h; r < h + 4; ++r) {
            if (j + 1 < NN(r)) {
              cn[r - h] = (unsigned)at(a.col, q0 + RED_THREADS * r);
              kn[r - h] = at(a.coef, q0 + RED_THREADS * r);
            }
          }
        }
        const unsigned q0 = (unsigned)(j * np + base);
#pragma unroll
        for (int r = h; r < h + 4; ++r) {
          if (j < NN(r)) {
            const double cs = kc[r - h] * wi[r] * at(a.w, cc[r - h]);
            at(a.cs, q0 + RED_THREADS * r) = cs;
            dc[r] = dc[r] - cs * at(a.u, cc[r - h]);
            if (DUAL) d2c[r] = d2c[r] - cs * at(a.u2, cc[r - h]);
          }
        }
      }
    }
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int r = 0; r < RED_PER_THREAD; ++r) {
      const int i = base + RED_THREADS * r;
      dc[r] = dc[r] - uc[r];
      dp[r] = 0.0;
      up[r] = 0.0;
      if (i < np) at(a.d, i) = dc[r];
      if (DUAL) {
        d2c[r] = d2c[r] - u2c[r];
        if (i < np) {
          at(a.d2, i) = d2c[r];
          s0 = s0 + dc[r] * dc[r];
          s1 = s1 + d2c[r] * d2c[r];
        }
      }
    }
    if (DUAL) {
      // chunk values of |d0|^2 for both starts, in the inner-product tree,
      // in the second parity slot of the partials (the first iteration
      // writes the first)
      s0 = wave_sum(s0);
      s1 = wave_sum(s1);
      __shared__ double wd[JCG_CPB][16];
      const int tt = t % RED_THREADS;
      double* ws = wd[t / RED_THREADS];
      if ((tt & 63) == 0) {
        ws[tt >> 6] = s0;
        ws[8 + (tt >> 6)] = s1;
      }
      __syncthreads();
      if (tt < 2 && c < nch) {
        const double* v = ws + 8 * tt;
        a.part[(size_t)nch * 4 + (size_t)c * 4 + tt] =
            ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
      }
    }
  }
  if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;
  // the start with the smaller initial pseudo-residual
  double* dv = a.d;
  if (DUAL) {
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    const double* part = a.part + (size_t)nch * 4;
    if (t < RED_XTHREADS)
      for (int cc = t; cc < nch; cc += RED_XTHREADS) {
        r0 = r0 + part[(size_t)cc * 4];
        r1 = r1 + part[(size_t)cc * 4 + 1];
      }
    block_tree3(r0, r1, r2, lds);
    __shared__ int use2;
    if (t == 0) {
      use2 = r1 < r0 ? 1 : 0;
      if (bid == 0) {
        a.saved[10] = r0;
        a.saved[11] = r1;
        if (use2) a.flags[F_MLUSED] += 1;
      }
    }
    __syncthreads();
    if (use2) {
      dv = a.d2;
      if (!eig_block)
#pragma unroll
        for (int r = 0; r < RED_PER_THREAD; ++r) {
          uc[r] = u2c[r];
          dc[r] = d2c[r];
        }
    }
  }

  // the initial pseudo-residual into LDS: the block's rows from registers, the
  // rows each side from memory (complete since the last grid barrier)
  if (staged) {
#pragma unroll
    for (int r = 0; r < RED_PER_THREAD; ++r) {
      const int i = base + RED_THREADS * r;
      if (i < np) dsh[i - slo] = dc[r];
    }
    stage_halo(dsh, dv, slo, hlo_row, t0, t1, hhi_row);
  }
  __syncthreads();

  // ---- iterations
  for (int in = 0; in <= a.itmax; ++in) {
    const bool need_u = !(L.q1 || (in > 5 && (in % 5) != 0));
    if (eig_block) {
      jcg_eig_step(L, a, tri1, tri2, in);
    } else {
      // N d, rows in groups of four (bounded register pressure); per row
      // the neighbours are taken in slot order.  The next slot's column
      // index and coefficient are loaded while this slot's d is gathered, so
      // each slot costs one memory round trip instead of two.
      // Staged, the column indices come as LDS positions, 16 bits per slot,
      // two slots per word (a.lcol): one index load per two slots.
#pragma unroll
      for (int h = 0; h < RED_PER_THREAD; h += 4) {
        int jm = 1;
        unsigned cn[4];  // unstaged: the next slot's column; staged: its pair word
        double sn[4];
#pragma unroll
        for (int r = h; r < h + 4; ++r) {
          nd[r] = 0.0;
          jm = NN(r) > jm ? NN(r) : jm;
          const unsigned q = (unsigned)(np + base + RED_THREADS * r);
          cn[r - h] = 1 < NN(r) ? (STAGED ? at(a.lcol, (unsigned)(base + RED_THREADS * r))
                                          : (unsigned)at(a.col, q))
                                : 0u;
          sn[r - h] = 1 < NN(r) ? at(a.cs, q) : 0.0;
        }
        for (int j = 1; j < jm; ++j) {
          unsigned cc[4];
          double sc[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            // slot j is the low half of its pair word when j is odd
            cc[k] = STAGED ? ((j & 1) ? (cn[k] & 0xffffu) : (cn[k] >> 16)) : cn[k];
            sc[k] = sn[k];
          }
          if (j + 1 < jm) {
            const unsigned q0 = (unsigned)((j + 1) * np + base);
            const unsigned p0 = (unsigned)((j >> 1) * np + base);  // pair of slot j+1
#pragma unroll
            for (int r = h; r < h + 4; ++r) {
              if (j + 1 < NN(r)) {
                if (!STAGED)
                  cn[r - h] = (unsigned)at(a.col, q0 + RED_THREADS * r);
                else if (!(j & 1))
                  cn[r - h] = at(a.lcol, p0 + RED_THREADS * r);
                sn[r - h] = at(a.cs, q0 + RED_THREADS * r);
              }
            }
          }
#pragma unroll
          for (int r = h; r < h + 4; ++r)
            if (j < NN(r))
              nd[r] = nd[r] - sc[r - h] * (STAGED ? dsh[cc[r - h]] : at(dv, cc[r - h]));
        }
      }
      double add = 0.0, adt = 0.0, auu = 0.0;
#pragma unroll
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        if (base + RED_THREADS * r < np) {
          add = add + dc[r] * dc[r];
          adt = adt + dc[r] * nd[r];
          if (need_u) auu = auu + uc[r] * uc[r];
        }
      }
      __shared__ double wsum[JCG_CPB][24];
      jcg_chunk_partials(a, wsum, c, nch, in, add, adt, auu);
    }
    if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks, STAGED)) return;

    jcg_scalars(L, a, lds, in, nch, need_u);
    if (L.halt) break;
    const double c1 = L.c1, c2 = L.c2, c3 = L.c3, c4 = L.c4;
    if (!eig_block) {
#pragma unroll
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        const int i = base + RED_THREADS * r;
        const double un = c1 * dc[r] + c2 * uc[r] + c3 * up[r];
        const double dn = c1 * nd[r] + c4 * dc[r] + c3 * dp[r];
        up[r] = uc[r];
        uc[r] = un;
        dp[r] = dc[r];
        dc[r] = dn;
        // written through to memory: the next grid barrier's release then
        // has no pseudo-residual lines to write back from this XCD's L2
        if (i < np) {
          __hip_atomic_store(&at(dv, i), dn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (staged) dsh[i - slo] = dn;
        }
      }
    }
    if (a.halo < 0) {
      if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;
    } else if (!eig_block) {
      // the partials and the eigenvalue pair are double-buffered by
      // iteration parity, so only the neighbours' d has to be waited for
      if (!halo_sync(hflags, tile, hlo, hhi, (unsigned)(in + 1), err, a.barrier_ticks, STAGED))
        return;
    }
    // the neighbouring tiles' new d into LDS (each barrier above ends with
    // the block's own rows written and the neighbours' visible)
    if (staged) {
      stage_halo(dsh, dv, slo, hlo_row, t0, t1, hhi_row);
      __syncthreads();
    }
  }
  // the result is the iterate of the last iteration index: u at the halt,
  // or the one before the final update when ITMAX ran out
  const bool halted = L.halt != 0;

  // ---- unscal (src/itpackv.F:2424-2479), ETAS and ETA2 (src/gwce.F:2006-2008)
  if (!eig_block) {
    double ri[RED_PER_THREAD];
    int jm = 1;
#pragma unroll
    for (int r = 0; r < RED_PER_THREAD; ++r) {
      const int i = base + RED_THREADS * r;
      ri[r] = 1.0;
      if (i < np) {
        const double wi = at(a.w, i);
        const double u = (halted ? uc[r] : up[r]) * wi;
        ri[r] = 1.0 / wi;
        at(a.coef, i) = ri[r] * ri[r];
        at(a.rhs, i) = at(a.rhs, i) * ri[r];
        at(a.u, i) = u;
        if (a.eta2)
          at(a.eta2, i) = (double)at(a.nodecode, i) * (u + 0.0) + at(a.eta1, i);
      }
      jm = NN(r) > jm ? NN(r) : jm;
    }
    // off-diagonals, four rows per slot so that their gathers overlap;
    // deferred to the next matrix pass when the caller asks
    for (int j = 1; j < (a.defer_unscal ? 1 : jm); ++j) {
      const unsigned q0 = (unsigned)(j * np + base);
#pragma unroll
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        if (j < NN(r)) {
          const unsigned q = q0 + RED_THREADS * r;
          at(a.coef, q) = at(a.cs, q) * ri[r] * (1.0 / at(a.w, (unsigned)at(a.col, q)));
        }
      }
    }
  }
#undef NN
  if (bid == 0 && t == 0) jcg_finish(L, a, halted);
}

template <bool DUAL, bool STAGED>
__global__ __launch_bounds__(JCG_BLOCK) void k_jcg(JcgArgs a) {
  jcg_solve<DUAL, STAGED>(a, (int)blockIdx.x, (int)gridDim.x);
}

// Several systems in one cooperative launch (an ensemble's members): blocks
// first[m] .. first[m+1]-1 solve system m with args[m].  A system's blocks
// synchronise only among themselves, through its own barrier words, and
// leave when it has converged; its arithmetic is that of its own launch.
template <bool DUAL, bool STAGED>
__global__ __launch_bounds__(JCG_BLOCK) void k_jcg_batch(const JcgArgs* __restrict__ args,
                                                         const int* __restrict__ first, int nsys) {
  const int b = blockIdx.x;
  int m = 0;
  while (m + 1 < nsys && first[m + 1] <= b) ++m;
  jcg_solve<DUAL, STAGED>(args[m], b - first[m], first[m + 1] - first[m]);
}

// ITPACK jcg for systems with more row tiles than blocks (more than the
// resident blocks allow, or a caller's limit): each row block works through
// the tiles bid, bid + nrb, ... with the iterates in memory (u in a.u, the
// previous u in a.uo, d in a.d, the previous d in a.dpv, N d in a.ndv), the
// products gathered from memory and two grid barriers per iteration (after
// the products, after the update).  Every row's operations and the
// reduction tree are those of jcg_solve, so the results are the same.
__device__ void jcg_solve_mem(const JcgArgs& a, int bid, int nblocks) {
  __shared__ ItLocal L;
  __shared__ double lds[12];
  __shared__ double tri1[JCG_MAX_ITMAX], tri2[JCG_MAX_ITMAX];
  __shared__ double wsum[JCG_CPB][24];
  const int np = a.n;
  const int t = threadIdx.x;
  const int nch = (np + RED_CHUNK - 1) / RED_CHUNK;
  const int ntile = (nch + JCG_CPB - 1) / JCG_CPB;
  const int nrb = nblocks - 1;  // row blocks; the last block estimates the eigenvalue
  const bool eig_block = bid == nrb;
  const int g = t / RED_THREADS, tt = t % RED_THREADS;
  unsigned gen = 0;
  int* err = a.flags + F_ERROR;
  double* u = a.u;
  double* d = a.d;
  if (t == 0) jcg_init_state(L, a);
  __syncthreads();

  // ---- scal, part 1 (src/itpackv.F:2374-2393), unless already done
  if (!eig_block && !a.prescaled)
    for (int tl = bid; tl < ntile; tl += nrb)
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        const int i = (tl * JCG_CPB + g) * RED_CHUNK + tt + RED_THREADS * r;
        if (i >= np) continue;
        const double dg = at(a.coef, i);
        if (!(dg > 0.0)) atomicOr(err, ERR_DIAG);
        const double sq = __builtin_sqrt(dg);
        at(u, i) = at(u, i) * sq;
        const double w = 1.0 / sq;
        at(a.w, i) = w;
        at(a.rhs, i) = at(a.rhs, i) * w;
      }
  if (!a.prescaled)
    if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;

  // ---- scal, part 2 and d0 = rhs - N u - u (src/itpackv.F:2394-2400, 346-353)
  if (!eig_block)
    for (int tl = bid; tl < ntile; tl += nrb)
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        const int i = (tl * JCG_CPB + g) * RED_CHUNK + tt + RED_THREADS * r;
        if (i >= np) continue;
        const double wi = at(a.w, i);
        double dci = at(a.rhs, i);
        const int nn = at(a.nneigh, i);
        for (int j = 1; j < nn; ++j) {
          const unsigned q = (unsigned)(j * np + i);
          const unsigned cc = (unsigned)at(a.col, q);
          const double cs = at(a.coef, q) * wi * at(a.w, cc);
          at(a.cs, q) = cs;
          dci = dci - cs * at(u, cc);
        }
        at(d, i) = dci - at(u, i);
        at(a.dpv, i) = 0.0;
        at(a.uo, i) = 0.0;
      }
  if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;

  // ---- iterations
  for (int in = 0; in <= a.itmax; ++in) {
    const bool need_u = !(L.q1 || (in > 5 && (in % 5) != 0));
    if (eig_block) {
      jcg_eig_step(L, a, tri1, tri2, in);
    } else {
      for (int tl = bid; tl < ntile; tl += nrb) {
        const int c = tl * JCG_CPB + g;
        double add = 0.0, adt = 0.0, auu = 0.0;
        for (int r = 0; r < RED_PER_THREAD; ++r) {
          const int i = c * RED_CHUNK + tt + RED_THREADS * r;
          if (i >= np) continue;
          const int nn = at(a.nneigh, i);
          double ndi = 0.0;
          for (int j = 1; j < nn; ++j) {
            const unsigned q = (unsigned)(j * np + i);
            ndi = ndi - at(a.cs, q) * at(d, (unsigned)at(a.col, q));
          }
          at(a.ndv, i) = ndi;
          const double dci = at(d, i);
          add = add + dci * dci;
          adt = adt + dci * ndi;
          if (need_u) {
            const double uci = at(u, i);
            auu = auu + uci * uci;
          }
        }
        jcg_chunk_partials(a, wsum, c, nch, in, add, adt, auu);
        __syncthreads();  // wsum is reused by the next tile
      }
    }
    if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;
    jcg_scalars(L, a, lds, in, nch, need_u);
    if (L.halt) break;
    const double c1 = L.c1, c2 = L.c2, c3 = L.c3, c4 = L.c4;
    if (!eig_block)
      for (int tl = bid; tl < ntile; tl += nrb)
        for (int r = 0; r < RED_PER_THREAD; ++r) {
          const int i = (tl * JCG_CPB + g) * RED_CHUNK + tt + RED_THREADS * r;
          if (i >= np) continue;
          const double dci = at(d, i), uci = at(u, i);
          const double un = c1 * dci + c2 * uci + c3 * at(a.uo, i);
          const double dn = c1 * at(a.ndv, i) + c4 * dci + c3 * at(a.dpv, i);
          at(a.uo, i) = uci;
          at(u, i) = un;
          at(a.dpv, i) = dci;
          at(d, i) = dn;
        }
    if (!grid_sync(a.barrier, gen, err, a.barrier_ticks, bid, nblocks)) return;
  }
  // the iterate of the last iteration index: u at the halt, or the one
  // before the final update when ITMAX ran out
  const bool halted = L.halt != 0;

  // ---- unscal (src/itpackv.F:2424-2479), ETAS and ETA2 (src/gwce.F:2006-2008)
  if (!eig_block)
    for (int tl = bid; tl < ntile; tl += nrb)
      for (int r = 0; r < RED_PER_THREAD; ++r) {
        const int i = (tl * JCG_CPB + g) * RED_CHUNK + tt + RED_THREADS * r;
        if (i >= np) continue;
        const double wi = at(a.w, i);
        const double ur = (halted ? at(u, i) : at(a.uo, i)) * wi;
        const double ri = 1.0 / wi;
        at(a.coef, i) = ri * ri;
        at(a.rhs, i) = at(a.rhs, i) * ri;
        at(u, i) = ur;
        if (a.eta2)
          at(a.eta2, i) = (double)at(a.nodecode, i) * (ur + 0.0) + at(a.eta1, i);
        if (!a.defer_unscal) {
          const int nn = at(a.nneigh, i);
          for (int j = 1; j < nn; ++j) {
            const unsigned q = (unsigned)(j * np + i);
            at(a.coef, q) = at(a.cs, q) * ri * (1.0 / at(a.w, (unsigned)at(a.col, q)));
          }
        }
      }
  if (bid == 0 && t == 0) jcg_finish(L, a, halted);
}

__global__ __launch_bounds__(JCG_BLOCK) void k_jcg_mem(JcgArgs a) {
  jcg_solve_mem(a, (int)blockIdx.x, (int)gridDim.x);
}

template __global__ void k_jcg<false, false>(JcgArgs);
template __global__ void k_jcg<true, false>(JcgArgs);
template __global__ void k_jcg<false, true>(JcgArgs);
template __global__ void k_jcg<true, true>(JcgArgs);
template __global__ void k_jcg_batch<false, false>(const JcgArgs*, const int*, int);
template __global__ void k_jcg_batch<true, false>(const JcgArgs*, const int*, int);
template __global__ void k_jcg_batch<false, true>(const JcgArgs*, const int*, int);
template __global__ void k_jcg_batch<true, true>(const JcgArgs*, const int*, int);

}  // namespace adcgpu
