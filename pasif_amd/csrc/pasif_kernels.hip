// gfx950 kernels for the batched RK4 forward integrator and its adjoint.
//
// One workgroup runs one tile (a packed set of independent instances) through
// a small state machine:
//
//   Forward    q_0 .. q_T, checkpoint every `ck_every` steps (adjoint runs)
//   Recompute  segment j: q from checkpoint j, steps jk .. min(jk+k, T),
//              every state stored to the tile's segment buffer
//   Backward   segment j: adjoint steps n = e .. jk+1 reading x_n, x_{n-1}
//   Done
//
// A launch advances each tile by at most `budget` steps and saves its
// registers (state, running maxima, phase) to HBM, so the same kernel is the
// persistent solver (budget = everything), the plain per-step baseline
// (budget = 1) and the serving tick (budget = a slice).
//
// Arithmetic restates the reference's step (src/__GpuDriver.cu:693-713,
// 746-783; src/kernels.cu:21-63, 71-126, 162-229):
//   k_s  = sum_terms coef * S_j [* X_k [* X_l [* X_m]]] + P_o u_s
//   y    = q + c_s k_s                  c = h/2, h/2, h
//   acc  = ((k1 + 2 k2) + 2 k3) + k4,   q <- q + h6 acc
// with fused multiply-adds where the reference's compiler contracts them.
#include <hip/hip_runtime.h>

#include "pasif_tiles.hpp"

namespace pasif {

template <typename R> __device__ __forceinline__ R fmaR(R a, R b, R c);
template <> __device__ __forceinline__ double fmaR<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <> __device__ __forceinline__ float fmaR<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <typename R> __device__ __forceinline__ R absR(R a);
template <> __device__ __forceinline__ double absR<double>(double a) { return __builtin_fabs(a); }
template <> __device__ __forceinline__ float absR<float>(float a) { return __builtin_fabsf(a); }

__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// LDS element at byte offset `off` (+ compile-time element offset E)
template <typename R, int E>
__device__ __forceinline__ R lds(const unsigned char* base, uint32_t off) {
  return *reinterpret_cast<const R*>(base + off + E * sizeof(R));
}

// One lane's segment program: coefficients and packed LDS byte offsets of
// its terms, held in registers.  Trip counts are per tile (uniform).
template <typename R, int C1, int C2, int C3, int C4>
struct LaneProg {
  static constexpr int CT = C1 + C2 + C3 + C4;
  static constexpr int CW = C1 + C2 + 2 * C3 + 2 * C4;
  R v1[C1];
  uint32_t w1[C1];
  R v2[C2];
  uint32_t w2[C2];
  R v3[C3];
  uint32_t w3a[C3], w3b[C3];
  R v4[C4];
  uint32_t w4a[C4], w4b[C4];
  uint32_t n1, n2, n3, n4;

  __device__ __forceinline__ void load(const Program<R>& p, const uint32_t* cnt, size_t tile,
                                       uint32_t L, uint32_t lane) {
    n1 = uni(cnt[0]);
    n2 = uni(cnt[1]);
    n3 = uni(cnt[2]);
    n4 = uni(cnt[3]);
    // The program is reloaded at every phase switch; keep the compiler from
    // hoisting its per-slot addresses out of the phase loop (they would stay
    // live in VGPRs for the whole kernel).
    uint32_t ln = lane;
    asm volatile("" : "+v"(ln));
    const R* V = p.val + tile * CT * L + ln;
    const uint32_t* I = p.idx + tile * CW * L + ln;
#pragma unroll
    for (int i = 0; i < C1; ++i)
      if (i < (int)n1) {
        v1[i] = V[i * L];
        w1[i] = I[i * L];
      }
#pragma unroll
    for (int i = 0; i < C2; ++i)
      if (i < (int)n2) {
        v2[i] = V[(C1 + i) * L];
        w2[i] = I[(C1 + i) * L];
      }
#pragma unroll
    for (int i = 0; i < C3; ++i)
      if (i < (int)n3) {
        v3[i] = V[(C1 + C2 + i) * L];
        w3a[i] = I[(C1 + C2 + 2 * i) * L];
        w3b[i] = I[(C1 + C2 + 2 * i + 1) * L];
      }
#pragma unroll
    for (int i = 0; i < C4; ++i)
      if (i < (int)n4) {
        v4[i] = V[(C1 + C2 + C3 + i) * L];
        w4a[i] = I[(C1 + C2 + 2 * C3 + 2 * i) * L];
        w4b[i] = I[(C1 + C2 + 2 * C3 + 2 * i + 1) * L];
      }
  }

  // Partial row sum of this segment.  S factors come from the stage buffer
  // SB of `S`, set-point factors from element XB of `X`.
  template <int SB, int XB>
  __device__ __forceinline__ R eval(const unsigned char* S, const unsigned char* X) const {
    R a = R(0);
#pragma unroll
    for (int i = 0; i < C1; ++i)
      if (i < (int)n1) a = fmaR(v1[i], lds<R, SB>(S, w1[i] & 0xFFFFu), a);
#pragma unroll
    for (int i = 0; i < C2; ++i)
      if (i < (int)n2) {
        const R t = v2[i] * lds<R, SB>(S, w2[i] & 0xFFFFu);
        a = fmaR(t, lds<R, XB>(X, w2[i] >> 16), a);
      }
#pragma unroll
    for (int i = 0; i < C3; ++i)
      if (i < (int)n3) {
        R t = v3[i] * lds<R, SB>(S, w3a[i] & 0xFFFFu);
        t = t * lds<R, XB>(X, w3a[i] >> 16);
        a = fmaR(t, lds<R, XB>(X, w3b[i] & 0xFFFFu), a);
      }
#pragma unroll
    for (int i = 0; i < C4; ++i)
      if (i < (int)n4) {
        R t = v4[i] * lds<R, SB>(S, w4a[i] & 0xFFFFu);
        t = t * lds<R, XB>(X, w4a[i] >> 16);
        t = t * lds<R, XB>(X, w4b[i] & 0xFFFFu);
        a = fmaR(t, lds<R, XB>(X, w4b[i] >> 16), a);
      }
    return a;
  }
};

// Excitation samples of one forced row for one step, fetched a step ahead.
template <typename R>
struct Samples {
  R e0, e1, e2;
};

template <typename R>
__device__ __forceinline__ Samples<R> fetch_samples(const TileArgs<R>& a, int32_t exc, uint32_t n) {
  const uint32_t Lx = a.exc_len;
  const R* E = a.excit + static_cast<size_t>(exc) * Lx;
  Samples<R> f;
  if (a.interp_w == 0u) {
    // zero-order hold (src/kernels.cu:83-86): e[n] for stages 1-3, e[n+1]
    // for stage 4, each only when 0 < s < L
    f.e0 = (n > 0u && n < Lx) ? E[n] : R(0);
    f.e1 = (n + 1u < Lx) ? E[n + 1u] : R(0);
    f.e2 = R(0);
  } else {
    // samples i0, i0+1, i0+2 cover hs = 2n .. 2n+2 for any window >= 2;
    // e[L] and beyond read as 0
    const uint32_t i0 = (2u * n) / a.interp_w;
    f.e0 = i0 < Lx ? E[i0] : R(0);
    f.e1 = i0 + 1u < Lx ? E[i0 + 1u] : R(0);
    f.e2 = i0 + 2u < Lx ? E[i0 + 2u] : R(0);
  }
  return f;
}

// forcing of half-step hs from samples starting at i0
// (src/kernels.cu:114-124: Mi[1][j] e[i] + Mi[0][j] e[i+1], 0 < i < L)
template <typename R>
__device__ __forceinline__ R interp_at(const Samples<R>& f, const R* wts, uint32_t W, uint32_t Lx,
                                       uint32_t i0, uint32_t hs) {
  const uint32_t i = hs / W;
  const uint32_t j = hs - i * W;
  const bool same = i == i0;
  const R ei = same ? f.e0 : f.e1;
  const R ei1 = same ? f.e1 : f.e2;
  return (i > 0u && i < Lx) ? fmaR(wts[W + j], ei, wts[j] * ei1) : R(0);
}

// One RK4 step of every row of the tile.  The owner's state is `st`; the
// stage-1 input y = st must already be published in buffer 0 of `yb`.
// Adjoint steps (ADJ) read set points x_n / xbar / x_{n-1} from elements
// 0 / 1 / 2 of `xb`; forward steps read them from the stage buffer itself.
template <bool ADJ, typename R, class Prog>
__device__ __forceinline__ void rk4_step(const Prog& P, unsigned char* yb, const unsigned char* xb,
                                         R* part, uint32_t lane, bool split, bool owner,
                                         uint32_t nseg, uint32_t yoff, R& st, bool forced, R u0,
                                         R u1, R u3, R force, bool active, R h, R h2, R h6) {
  R acc = R(0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    constexpr int dummy = 0;
    (void)dummy;
    R p;
    if (ADJ) {
      if (s == 0) p = P.template eval<0, 0>(yb, xb);
      else if (s == 1) p = P.template eval<1, 1>(yb, xb);
      else if (s == 2) p = P.template eval<0, 1>(yb, xb);
      else p = P.template eval<1, 2>(yb, xb);
    } else {
      if (s & 1) p = P.template eval<1, 1>(yb, yb);
      else p = P.template eval<0, 0>(yb, yb);
    }
    if (split) {
      part[lane] = p;
      __syncthreads();
      if (owner)
        for (uint32_t j = 1; j < nseg; ++j) p += part[lane + j];
    }
    if (owner) {
      const R us = s == 0 ? u0 : (s == 3 ? u3 : u1);
      const R k = forced ? fmaR(force, us, p) : p;
      R y;
      if (s == 0) {
        acc = k;
        y = fmaR(h2, k, st);
      } else if (s == 1) {
        acc = fmaR(R(2), k, acc);
        y = fmaR(h2, k, st);
      } else if (s == 2) {
        acc = fmaR(R(2), k, acc);
        y = fmaR(h, k, st);
      } else {
        acc = acc + k;
        const R qn = fmaR(h6, acc, st);
        if (active) st = qn;
        y = st;
      }
      *reinterpret_cast<R*>(yb + yoff + ((s + 1) & 1) * sizeof(R)) = y;
    }
    __syncthreads();
  }
}

// Dynamic LDS layout (bytes): yb [L+1][2] R (stage double buffer, row L = 0),
// part [L] R, xb [L+1][4] R (adjoint set points, row L = 0), wts [2][W] R.
template <typename R>
__host__ __device__ inline size_t lds_bytes(uint32_t L, bool adjoint, uint32_t W) {
  return (2 * (L + 1) + L + (adjoint ? 4 * (L + 1) : 0) + 2 * W) * sizeof(R);
}

template <typename R, int C1, int C2, int C3, int C4, int MAXL, bool ADJ>
__global__ void __launch_bounds__(MAXL) tile_kernel(TileArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using Prog = LaneProg<R, C1, C2, C3, C4>;
  const uint32_t L = a.lanes;
  constexpr bool adjoint = ADJ;
  unsigned char* yb = smem_raw;
  R* part = reinterpret_cast<R*>(smem_raw + 2 * (L + 1) * sizeof(R));
  unsigned char* xbb = smem_raw + (3 * L + 2) * sizeof(R);
  R* xb = reinterpret_cast<R*>(xbb);
  R* wts = reinterpret_cast<R*>(xbb + (adjoint ? 4 * (L + 1) * sizeof(R) : 0));
  const uint32_t lane = threadIdx.x;
  const size_t tile = blockIdx.x;
  const TileHeader* H = a.hdr + tile;
  const uint32_t T = uni(H->steps);
  const uint32_t K = adjoint ? a.ck_every : 1u;
  const size_t rb = tile * L;
  const size_t segb = tile * (K + 1);
  const uint32_t W = a.interp_w;

  uint32_t* pst = a.progress + 4 * tile;
  uint32_t n = uni(pst[0]);
  uint32_t phase = uni(pst[1]);
  uint32_t seg = uni(pst[2]);
  uint32_t budget = a.budget;

  if (lane == 0) {
    R* yz = reinterpret_cast<R*>(yb) + 2 * L;
    yz[0] = R(0);
    yz[1] = R(0);
    if (adjoint)
      for (int e = 0; e < 4; ++e) xb[4 * L + e] = R(0);
  }
  for (uint32_t i = lane; i < 2 * W; i += L) wts[i] = a.interp[i];

  // forward rows: ownership, per-row data, state
  const uint32_t finfo = a.fwd.info[rb + lane];
  const bool fown = lane_owner(finfo);
  const uint32_t frow = lane_row(finfo);
  const uint32_t fnseg = lane_nseg(finfo);
  const uint32_t fyoff = frow * 2 * sizeof(R);
  uint32_t fsteps = 0, fout = 0;
  int32_t fexc = -1;
  R fforce = R(0);
  R q = R(0), mx = R(0);
  if (fown) {
    fsteps = a.fwd.steps[rb + lane];
    fout = a.fwd.out[rb + lane];
    fexc = a.fwd.exc[rb + lane];
    fforce = a.fwd.force[rb + lane];
    if (n == 0u && phase == kPhaseForward) {
      q = a.fwd.init[a.fwd.src[rb + lane]];
      if (fsteps == 0u) a.out_q[fout] = q;
    } else {
      q = a.qstate[rb + frow];
      if (a.window) mx = a.mxstate[rb + frow];
    }
  }
  const bool forced = fown && fexc >= 0;
  const bool fsplit = uni(H->split_f) != 0u;
  const uint32_t rows_f = uni(H->rows_f);

  // adjoint rows
  bool aown = false, asplit = false;
  uint32_t anseg = 0, aout = 0, asteps = 0, asrc = 0, ayoff = 0, arow = 0;
  R lam = R(0);
  if (adjoint) {
    const uint32_t ainfo = a.adj.info[rb + lane];
    aown = lane_owner(ainfo);
    arow = lane_row(ainfo);
    anseg = lane_nseg(ainfo);
    ayoff = arow * 2 * sizeof(R);
    asplit = uni(H->split_a) != 0u;
    if (aown) {
      aout = a.adj.out[rb + lane];
      asteps = a.adj.steps[rb + lane];
      asrc = a.adj.src[rb + lane];
      if (phase != kPhaseForward) lam = a.lstate[rb + arow];
    }
  }

  Prog P;
  while (budget > 0u && phase != kPhaseDone) {
    if (phase != kPhaseBackward) {
      // ---------------- forward / recompute ----------------
      const bool recompute = phase == kPhaseRecompute;
      P.load(a.fwd, H->cnt_f, tile, L, lane);
      const uint32_t s0 = seg * K;
      if (recompute && n == s0 && fown) {
        q = a.ckpt[(tile * a.nck + seg) * L + frow];
        a.segbuf[segb * L + frow] = q;
      }
      if (fown) *reinterpret_cast<R*>(yb + fyoff) = q;
      __syncthreads();
      const uint32_t e = recompute ? min(T, s0 + K) : T;
      const uint32_t nend = (e - n < budget) ? e : n + budget;
      budget -= nend - n;
      Samples<R> fcur{R(0), R(0), R(0)}, fnext{R(0), R(0), R(0)};
      if (forced) fcur = fetch_samples(a, fexc, n);
      for (; n < nend; ++n) {
        R u0 = R(0), u1 = R(0), u3 = R(0);
        if (forced) {
          fnext = fetch_samples(a, fexc, n + 1u);
          if (W == 0u) {
            u0 = fcur.e0;
            u1 = fcur.e0;
            u3 = fcur.e1;
          } else {
            const uint32_t i0 = (2u * n) / W;
            u0 = interp_at(fcur, wts, W, a.exc_len, i0, 2u * n);
            u1 = interp_at(fcur, wts, W, a.exc_len, i0, 2u * n + 1u);
            u3 = interp_at(fcur, wts, W, a.exc_len, i0, 2u * n + 2u);
          }
        }
        if (adjoint && !recompute && fown && (n % K) == 0u)
          a.ckpt[(tile * a.nck + n / K) * L + frow] = q;
        const bool active = n < fsteps;
        rk4_step<false>(P, yb, xbb, part, lane, fsplit, fown, fnseg, fyoff, q, forced, u0, u1,
                        u3, fforce, active, a.h, a.h2, a.h6);
        if (fown) {
          if (recompute) {
            a.segbuf[(segb + (n + 1u - s0)) * L + frow] = q;
          } else {
            if (a.save_every && (n % a.save_every) == 0u) {
              const uint32_t slot = n / a.save_every;
              if (slot < a.nslots) a.traj[static_cast<size_t>(slot) * a.out_len + fout] = q;
            }
            if (a.window && active && n + 1u + a.window > fsteps) mx = max(mx, absR(q));
            if (n + 1u == fsteps) {
              a.out_q[fout] = q;
              if (a.window) a.out_max[fout] = mx;
            }
          }
        }
        fcur = fnext;
      }
      if (n == e) {
        if (recompute) {
          phase = kPhaseBackward;
        } else if (adjoint && T > 0u) {
          seg = (T - 1u) / K;
          n = seg * K;
          phase = kPhaseRecompute;
          if (aown) lam = a.adj.init[asrc];
        } else {
          phase = kPhaseDone;
          if (adjoint && aown) a.out_grad[aout] = a.adj.init[asrc];
        }
      }
      __syncthreads();
    } else if constexpr (ADJ) {
      // ---------------- adjoint (reverse) ----------------
      P.load(a.adj, H->cnt_a, tile, L, lane);
      const uint32_t s0 = seg * K;
      if (aown) *reinterpret_cast<R*>(yb + ayoff) = lam;
      R xcur = R(0), xprev = R(0);
      if (lane < rows_f) {
        xcur = a.segbuf[(segb + (n - s0)) * L + lane];
        if (n > s0) xprev = a.segbuf[(segb + (n - 1u - s0)) * L + lane];
      }
      __syncthreads();
      const uint32_t nend = (n - s0 < budget) ? s0 : n - budget;
      budget -= n - nend;
      for (; n > nend; --n) {
        if (lane < rows_f) {
          const R xnext = (n - 1u > s0) ? a.segbuf[(segb + (n - 2u - s0)) * L + lane] : R(0);
          // request / half / previous set point (src/kernels.cu:162-195)
          xb[4 * lane] = xcur;
          xb[4 * lane + 1] = R(0.5) * (xcur + xprev);
          xb[4 * lane + 2] = xprev;
          xcur = xprev;
          xprev = xnext;
        }
        __syncthreads();
        rk4_step<true>(P, yb, xbb, part, lane, asplit, aown, anseg, ayoff, lam, false, R(0), R(0),
                       R(0), R(0), n <= asteps, a.h, a.h2, a.h6);
      }
      if (n == s0) {
        if (seg == 0u) {
          phase = kPhaseDone;
          if (aown) a.out_grad[aout] = lam;
        } else {
          seg -= 1u;
          n = seg * K;
          phase = kPhaseRecompute;
        }
      }
      __syncthreads();
    }
  }

  if (fown) {
    a.qstate[rb + frow] = q;
    if (a.window) a.mxstate[rb + frow] = mx;
  }
  if (adjoint && aown) a.lstate[rb + arow] = lam;
  if (lane == 0) {
    pst[0] = n;
    pst[1] = phase;
    pst[2] = seg;
  }
}

template <typename R, int C1, int C2, int C3, int C4, int MAXL>
static hipError_t launch_cfg(const TileArgs<R>& a, uint32_t ntiles, hipStream_t s) {
  const size_t lds = lds_bytes<R>(a.lanes, a.adjoint != 0u, a.interp_w);
  auto kern = a.adjoint ? tile_kernel<R, C1, C2, C3, C4, MAXL, true>
                        : tile_kernel<R, C1, C2, C3, C4, MAXL, false>;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(ntiles), dim3(a.lanes), lds, s, a);
  return hipGetLastError();
}

template <typename R>
static hipError_t launch_tiles_impl(const TileArgs<R>& a, uint32_t ntiles, int slot_cfg,
                                    hipStream_t s) {
  if (a.lanes == 0 || a.lanes > kMaxLanes || (a.lanes % 64) != 0) return hipErrorInvalidValue;
  if (lds_bytes<R>(a.lanes, a.adjoint != 0u, a.interp_w) > kMaxLdsBytes)
    return hipErrorInvalidValue;
  if (a.lanes <= 256) {
    if (slot_cfg == 2)
      return launch_cfg<R, kSlotWide.c1, kSlotWide.c2, kSlotWide.c3, kSlotWide.c4, 256>(a, ntiles,
                                                                                      s);
    return launch_cfg<R, kSlotNarrow.c1, kSlotNarrow.c2, kSlotNarrow.c3, kSlotNarrow.c4, 256>(
        a, ntiles, s);
  }
  if (slot_cfg == 2) return hipErrorInvalidValue;
  return launch_cfg<R, kSlotNarrow.c1, kSlotNarrow.c2, kSlotNarrow.c3, kSlotNarrow.c4, 1024>(
      a, ntiles, s);
}

hipError_t launch_tiles(const TileArgs<double>& a, uint32_t ntiles, int slot_cfg, hipStream_t s) {
  return launch_tiles_impl(a, ntiles, slot_cfg, s);
}
hipError_t launch_tiles(const TileArgs<float>& a, uint32_t ntiles, int slot_cfg, hipStream_t s) {
  return launch_tiles_impl(a, ntiles, slot_cfg, s);
}

}  // namespace pasif
