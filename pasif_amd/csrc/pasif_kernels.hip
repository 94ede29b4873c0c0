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

#include <mutex>
#include <vector>

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

// Workgroup barrier ordering LDS only.  Global memory stays unordered, so
// prefetched excitation / trajectory loads in flight are not waited for; the
// full __syncthreads() is kept where global data written by one lane is read
// by another (checkpoint and segment buffers across phase changes).
__device__ __forceinline__ void sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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
    // A kind that the tile uses is loaded whole: its unused slots hold value 0
    // and point at the zero row, so evaluating them adds exact zeros.
    if (n1) {
#pragma unroll
      for (int i = 0; i < C1; ++i) {
        v1[i] = V[i * L];
        w1[i] = I[i * L];
      }
    }
    if (n2) {
#pragma unroll
      for (int i = 0; i < C2; ++i) {
        v2[i] = V[(C1 + i) * L];
        w2[i] = I[(C1 + i) * L];
      }
    }
    if (n3) {
#pragma unroll
      for (int i = 0; i < C3; ++i) {
        v3[i] = V[(C1 + C2 + i) * L];
        w3a[i] = I[(C1 + C2 + 2 * i) * L];
        w3b[i] = I[(C1 + C2 + 2 * i + 1) * L];
      }
    }
    if (n4) {
#pragma unroll
      for (int i = 0; i < C4; ++i) {
        v4[i] = V[(C1 + C2 + C3 + i) * L];
        w4a[i] = I[(C1 + C2 + 2 * C3 + 2 * i) * L];
        w4b[i] = I[(C1 + C2 + 2 * C3 + 2 * i + 1) * L];
      }
    }
  }

  // Partial row sum of this segment.  S factors come from the stage buffer
  // SB of `S`, set-point factors from element XB of `X`.  Slots of a kind are
  // evaluated without per-slot branches so their LDS reads issue together.
  template <int SB, int XB>
  __device__ __forceinline__ R eval(const unsigned char* S, const unsigned char* X) const {
    // two partial sums per kind (even / odd slots) shorten the dependent
    // FMA chains; they are combined as ((k1 + k2) + (k3 + k4))
    R a0 = R(0), a1 = R(0), b0 = R(0), b1 = R(0), c0 = R(0), c1 = R(0), d0 = R(0), d1 = R(0);
    if (n1) {
#pragma unroll
      for (int i = 0; i < C1; ++i) {
        const R y = lds<R, SB>(S, w1[i] & 0xFFFFu);
        if (i & 1) a1 = fmaR(v1[i], y, a1);
        else a0 = fmaR(v1[i], y, a0);
      }
    }
    if (n2) {
#pragma unroll
      for (int i = 0; i < C2; ++i) {
        const R t = v2[i] * lds<R, SB>(S, w2[i] & 0xFFFFu);
        const R x = lds<R, XB>(X, w2[i] >> 16);
        if (i & 1) b1 = fmaR(t, x, b1);
        else b0 = fmaR(t, x, b0);
      }
    }
    if (n3) {
#pragma unroll
      for (int i = 0; i < C3; ++i) {
        R t = v3[i] * lds<R, SB>(S, w3a[i] & 0xFFFFu);
        t = t * lds<R, XB>(X, w3a[i] >> 16);
        const R x = lds<R, XB>(X, w3b[i] & 0xFFFFu);
        if (i & 1) c1 = fmaR(t, x, c1);
        else c0 = fmaR(t, x, c0);
      }
    }
    if (n4) {
#pragma unroll
      for (int i = 0; i < C4; ++i) {
        R t = v4[i] * lds<R, SB>(S, w4a[i] & 0xFFFFu);
        t = t * lds<R, XB>(X, w4a[i] >> 16);
        t = t * lds<R, XB>(X, w4b[i] & 0xFFFFu);
        const R x = lds<R, XB>(X, w4b[i] >> 16);
        if (i & 1) d1 = fmaR(t, x, d1);
        else d0 = fmaR(t, x, d0);
      }
    }
    return ((a0 + a1) + (b0 + b1)) + ((c0 + c1) + (d0 + d1));
  }
};

// Forcing samples.  A forced row stages the samples of its excitation that
// the next kForceChunk steps read into its own LDS slot, so the step loop
// issues no global loads.  Zero-order hold (src/kernels.cu:83-86) reads
// e[n] (stages 1-3) and e[n+1] (stage 4), each only when 0 < s < L;
// interpolation (src/kernels.cu:114-124) on the half-step grid
// hs = 2n .. 2n+2 reads e[i .. i+1] with i = hs / W (any window W >= 2).
// Samples at index L and beyond read as 0.  A chunk starting at sample b
// holds e[b .. b + kForceStage).
template <typename R>
__device__ __forceinline__ void stage_samples(R* slot, const R* E, uint32_t Lx, uint32_t b) {
  // loads from clamped indices are unconditional, so each batch of 16 is in
  // flight together; out-of-range samples are zeroed afterwards
  constexpr uint32_t B = 16;
#pragma unroll
  for (uint32_t t0 = 0; t0 < kForceStage; t0 += B) {
    R v[B];
#pragma unroll
    for (uint32_t u = 0; u < B; ++u) {
      const uint32_t i = b + t0 + u;
      v[u] = E[i < Lx ? i : Lx - 1u];
    }
#pragma unroll
    for (uint32_t u = 0; u < B; ++u)
      if (t0 + u < kForceStage) slot[t0 + u] = (b + t0 + u < Lx) ? v[u] : R(0);
  }
}

// Mi[1][j] e[i] + Mi[0][j] e[i+1] for the half step hs = i0*W + j0 + t
// (t = 0, 1, 2); zero unless 0 < i < L.  `slot` holds samples from index b.
template <typename R>
__device__ __forceinline__ R interp_at(const R* slot, uint32_t b, const R* wts, uint32_t W,
                                       uint32_t Lx, uint32_t i0, uint32_t j0, uint32_t t) {
  uint32_t j = j0 + t;
  const bool next = j >= W;
  j = next ? j - W : j;
  const uint32_t i = i0 + (next ? 1u : 0u);
  const R ei = slot[i - b];
  const R ei1 = slot[i + 1u - b];
  return (i > 0u && i < Lx) ? fmaR(wts[W + j], ei, wts[j] * ei1) : R(0);
}

// One RK4 step of every row of the tile.  The owner's state is `st`; the
// stage-1 input y = st must already be published in buffer 0 of `yb`.
// Adjoint steps (ADJ) read set points x_n / xbar / x_{n-1} from elements
// 0 / 1 / 2 of `xb`; forward steps read them from the stage buffer itself.
// `rec` (forward steps, owner lanes; null: none) receives the step's stage
// inputs y1 = st, y2, y3, y4 at rec[0], rec[L], rec[2L], rec[3L].
template <bool ADJ, typename R, class Prog>
__device__ __forceinline__ void rk4_step(const Prog& P, unsigned char* yb, const unsigned char* xb,
                                         R* part, uint32_t lane, bool split, bool owner,
                                         uint32_t nseg, uint32_t yoff, R& st, bool forced, R u0,
                                         R u1, R u3, R force, bool active, R h, R h2, R h6,
                                         R* rec = nullptr, uint32_t L = 0) {
  R acc = R(0);
  if (!ADJ && rec) rec[0] = st;
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
      sync_lds();
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
      if (!ADJ && rec && s < 3) rec[(s + 1) * L] = y;
      *reinterpret_cast<R*>(yb + yoff + ((s + 1) & 1) * sizeof(R)) = y;
    }
    sync_lds();
  }
}

// One step of the exact discrete adjoint (reverse mode of rk4_step's
// forward map) for every adjoint row of the tile.  The owner's state is
// `lam`; h6 * lam must already be published in buffer 0 of `yb`, and the
// step's forward stage inputs y1..y4 in elements 0..3 of `xb`:
//   a4 = h6 lam;             mu4 = g(a4, y4)
//   a3 = 2 h6 lam + h mu4;   mu3 = g(a3, y3)
//   a2 = 2 h6 lam + h2 mu3;  mu2 = g(a2, y2)
//   a1 = h6 lam + h2 mu2;    mu1 = g(a1, y1)
//   lam += ((mu4 + mu3) + mu2) + mu1
// and h6 * lam is published again for the next step.
template <typename R, class Prog>
__device__ __forceinline__ void rk4_exact_adjoint_step(const Prog& P, unsigned char* yb,
                                                       const unsigned char* xb, R* part,
                                                       uint32_t lane, bool split, bool owner,
                                                       uint32_t nseg, uint32_t yoff, R& lam,
                                                       bool active, R h, R h2, R h6) {
  R acc = R(0);
  const R h3 = R(2) * h6;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    R p;
    if (s == 0) p = P.template eval<0, 3>(yb, xb);
    else if (s == 1) p = P.template eval<1, 2>(yb, xb);
    else if (s == 2) p = P.template eval<0, 1>(yb, xb);
    else p = P.template eval<1, 0>(yb, xb);
    if (split) {
      part[lane] = p;
      sync_lds();
      if (owner)
        for (uint32_t j = 1; j < nseg; ++j) p += part[lane + j];
    }
    if (owner) {
      R y;
      if (s == 0) {
        acc = p;
        y = fmaR(h, p, h3 * lam);
      } else if (s == 1) {
        acc = acc + p;
        y = fmaR(h2, p, h3 * lam);
      } else if (s == 2) {
        acc = acc + p;
        y = fmaR(h2, p, h6 * lam);
      } else {
        acc = acc + p;
        const R ln = lam + acc;
        if (active) lam = ln;
        y = h6 * lam;
      }
      *reinterpret_cast<R*>(yb + yoff + ((s + 1) & 1) * sizeof(R)) = y;
    }
    sync_lds();
  }
}

// Dynamic LDS layout (elements of R): yb [L+1][2] (stage double buffer,
// row L = 0), part [L], xb [L+1][4] (adjoint set points, row L = 0),
// wts [2][W] (interpolation rows 0 and 1), eb [max_forced][kForceStage]
// (staged excitation samples).
template <typename R>
__host__ __device__ inline size_t lds_bytes(uint32_t L, bool adjoint, uint32_t W, uint32_t nf) {
  return (2 * (L + 1) + L + (adjoint ? 4 * (L + 1) : 0) + 2 * W + nf * kForceStage) * sizeof(R);
}

// One tile of one job through its phases (see the file comment).  ADJ
// compiles the adjoint phases in; a job without an adjoint program runs the
// forward phase only.
template <typename R, int C1, int C2, int C3, int C4, bool ADJ>
__device__ __forceinline__ void run_tile(const TileArgs<R>& a, const size_t tile,
                                         unsigned char* smem_raw) {
  using Prog = LaneProg<R, C1, C2, C3, C4>;
  const uint32_t L = a.lanes;
  const bool adjoint = ADJ && a.adjoint != 0u;
  unsigned char* yb = smem_raw;
  R* part = reinterpret_cast<R*>(smem_raw + 2 * (L + 1) * sizeof(R));
  unsigned char* xbb = smem_raw + (3 * L + 2) * sizeof(R);
  R* xb = reinterpret_cast<R*>(xbb);
  R* wts = reinterpret_cast<R*>(xbb + (adjoint ? 4 * (L + 1) * sizeof(R) : 0));
  R* eb = wts + 2 * a.interp_w;
  const uint32_t lane = threadIdx.x;
  const TileHeader* H = a.hdr + tile;
  const uint32_t T = uni(H->steps);
  const uint32_t K = adjoint ? a.ck_every : 1u;
  const size_t rb = tile * L;
  const bool exact = adjoint && a.exact != 0u;
  // segment buffer of the tile: post-step states [K+1][L], or (exact) the
  // four stage inputs of each step [K+1][4][L]
  const size_t segb = tile * (K + 1) * (exact ? 4u : 1u);
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
  uint32_t fslot = 0;
  R fforce = R(0);
  R q = R(0), mx = R(0);
  if (fown) {
    fsteps = a.fwd.steps[rb + lane];
    fout = a.fwd.out[rb + lane];
    fexc = a.fwd.exc[rb + lane];
    fslot = a.fwd.fslot[rb + lane];
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
        q = a.ckpt[size_t(a.ck_tile[tile] + seg) * L + frow];
        if (!exact) a.segbuf[segb * L + frow] = q;
      }
      if (fown) *reinterpret_cast<R*>(yb + fyoff) = q;
      sync_lds();
      const uint32_t e = recompute ? min(T, s0 + K) : T;
      const uint32_t nend = (e - n < budget) ? e : n + budget;
      budget -= nend - n;
      // half-step position hs = 2n = hi*W + hj, advanced incrementally
      uint32_t hi = W ? (2u * n) / W : 0u, hj = W ? 2u * n - hi * W : 0u;
      const R* E = a.excit + static_cast<size_t>(forced ? fexc : 0) * a.exc_len;
      R* eslot = eb + fslot * kForceStage;
      uint32_t ebase = 0, left = 0;  // first staged sample, steps left in the chunk
      for (; n < nend; ++n) {
        R u0 = R(0), u1 = R(0), u3 = R(0);
        if (left == 0u) {
          ebase = W ? hi : n;
          if (forced) stage_samples(eslot, E, a.exc_len, ebase);
          left = kForceChunk;
        }
        --left;
        if (forced) {
          if (W == 0u) {
            const R e0 = eslot[n - ebase], e1 = eslot[n + 1u - ebase];
            u0 = (n > 0u && n < a.exc_len) ? e0 : R(0);
            u1 = u0;
            u3 = e1;
          } else {
            u0 = interp_at(eslot, ebase, wts, W, a.exc_len, hi, hj, 0u);
            u1 = interp_at(eslot, ebase, wts, W, a.exc_len, hi, hj, 1u);
            u3 = interp_at(eslot, ebase, wts, W, a.exc_len, hi, hj, 2u);
          }
        }
        hj += 2u;
        if (W && hj >= W) {
          hj -= W;
          hi += 1u;
        }
        if (adjoint && !recompute && fown && (n % K) == 0u)
          a.ckpt[size_t(a.ck_tile[tile] + n / K) * L + frow] = q;
        const bool active = n < fsteps;
        rk4_step<false>(P, yb, xbb, part, lane, fsplit, fown, fnseg, fyoff, q, forced, u0, u1,
                        u3, fforce, active, a.h, a.h2, a.h6,
                        (exact && recompute && fown)
                            ? a.segbuf + (segb + 4u * (n - s0)) * L + frow : nullptr, L);
        if (fown) {
          if (recompute) {
            if (!exact) a.segbuf[(segb + (n + 1u - s0)) * L + frow] = q;
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
      }
      if (n == e && !recompute && a.save_every && fown) {
        // trajectory slots after the tile's last step hold the final state
        for (uint32_t slot = (T + a.save_every - 1u) / a.save_every; slot < a.nslots; ++slot)
          a.traj[static_cast<size_t>(slot) * a.out_len + fout] = q;
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
    } else if (ADJ) {
      // ---------------- adjoint (reverse) ----------------
      P.load(a.adj, H->cnt_a, tile, L, lane);
      const uint32_t s0 = seg * K;
      if (aown) *reinterpret_cast<R*>(yb + ayoff) = exact ? a.h6 * lam : lam;
      R xcur = R(0), xprev = R(0);
      if (lane < rows_f && !exact) {
        xcur = a.segbuf[(segb + (n - s0)) * L + lane];
        if (n > s0) xprev = a.segbuf[(segb + (n - 1u - s0)) * L + lane];
      }
      sync_lds();
      const uint32_t nend = (n - s0 < budget) ? s0 : n - budget;
      budget -= n - nend;
      if (exact) {
        for (; n > nend; --n) {
          if (lane < rows_f) {
            // stage inputs of the forward step n-1 -> n
            const R* y = a.segbuf + (segb + 4u * (n - 1u - s0)) * L + lane;
            const R y1 = y[0], y2 = y[L], y3 = y[2 * L], y4 = y[3 * L];
            xb[4 * lane] = y1;
            xb[4 * lane + 1] = y2;
            xb[4 * lane + 2] = y3;
            xb[4 * lane + 3] = y4;
          }
          sync_lds();
          rk4_exact_adjoint_step(P, yb, xbb, part, lane, asplit, aown, anseg, ayoff, lam,
                                 n <= asteps, a.h, a.h2, a.h6);
        }
      }
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
        sync_lds();
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

template <typename R, int C1, int C2, int C3, int C4, int MAXL, bool ADJ>
__global__ void __launch_bounds__(MAXL) tile_kernel(TileArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  run_tile<R, C1, C2, C3, C4, ADJ>(a, blockIdx.x, smem_raw);
}

// Tiles of many jobs in one launch: workgroup b runs tile map[2b+1] of job
// map[2b].  Every job shares the workgroup size, slot configuration and
// precision; its arguments are read from the job table.
template <typename R, int C1, int C2, int C3, int C4, int MAXL>
__global__ void __launch_bounds__(MAXL) pool_kernel(const TileArgs<R>* __restrict__ jobs,
                                                     const uint32_t* __restrict__ map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const uint32_t job = uni(map[2 * blockIdx.x]);
  const uint32_t tile = uni(map[2 * blockIdx.x + 1]);
  const TileArgs<R> a = jobs[job];
  run_tile<R, C1, C2, C3, C4, true>(a, tile, smem_raw);
}

template <typename R, int C1, int C2, int C3, int C4, int MAXL>
static hipError_t launch_cfg(const TileArgs<R>& a, uint32_t ntiles, hipStream_t s) {
  const size_t lds = lds_bytes<R>(a.lanes, a.adjoint != 0u, a.interp_w, a.max_forced);
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
  if (lds_bytes<R>(a.lanes, a.adjoint != 0u, a.interp_w, a.max_forced) > kMaxLdsBytes)
    return hipErrorInvalidValue;
  if (a.lanes <= 256) {
    if (slot_cfg == 2)
      return launch_cfg<R, kSlotWide.c1, kSlotWide.c2, kSlotWide.c3, kSlotWide.c4, 256>(a, ntiles,
                                                                                      s);
    if (slot_cfg == 3)
      return launch_cfg<R, kSlotSmall.c1, kSlotSmall.c2, kSlotSmall.c3, kSlotSmall.c4, 256>(
          a, ntiles, s);
    return launch_cfg<R, kSlotNarrow.c1, kSlotNarrow.c2, kSlotNarrow.c3, kSlotNarrow.c4, 256>(
        a, ntiles, s);
  }
  if (slot_cfg != 1) return hipErrorInvalidValue;
  return launch_cfg<R, kSlotNarrow.c1, kSlotNarrow.c2, kSlotNarrow.c3, kSlotNarrow.c4, 1024>(
      a, ntiles, s);
}

template <typename R, int C1, int C2, int C3, int C4, int MAXL>
static hipError_t launch_pool_cfg(const TileArgs<R>* jobs, const uint32_t* map, uint32_t nblocks,
                                  uint32_t lanes, size_t lds, hipStream_t s) {
  auto kern = pool_kernel<R, C1, C2, C3, C4, MAXL>;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(lanes), lds, s, jobs, map);
  return hipGetLastError();
}

template <typename R>
static hipError_t launch_pool_impl(const TileArgs<R>* jobs, const uint32_t* map, uint32_t nblocks,
                                   uint32_t lanes, int slot_cfg, size_t lds, hipStream_t s) {
  if (lanes == 0 || lanes > kMaxLanes || (lanes % 64) != 0 || lds > kMaxLdsBytes)
    return hipErrorInvalidValue;
  if (lanes <= 256) {
    if (slot_cfg == 2)
      return launch_pool_cfg<R, kSlotWide.c1, kSlotWide.c2, kSlotWide.c3, kSlotWide.c4, 256>(
          jobs, map, nblocks, lanes, lds, s);
    if (slot_cfg == 3)
      return launch_pool_cfg<R, kSlotSmall.c1, kSlotSmall.c2, kSlotSmall.c3, kSlotSmall.c4, 256>(
          jobs, map, nblocks, lanes, lds, s);
    return launch_pool_cfg<R, kSlotNarrow.c1, kSlotNarrow.c2, kSlotNarrow.c3, kSlotNarrow.c4,
                           256>(jobs, map, nblocks, lanes, lds, s);
  }
  if (slot_cfg != 1) return hipErrorInvalidValue;
  return launch_pool_cfg<R, kSlotNarrow.c1, kSlotNarrow.c2, kSlotNarrow.c3, kSlotNarrow.c4, 1024>(
      jobs, map, nblocks, lanes, lds, s);
}

hipError_t launch_pool(const TileArgs<double>* jobs, const uint32_t* map, uint32_t nblocks,
                       uint32_t lanes, int slot_cfg, size_t lds, hipStream_t s) {
  return launch_pool_impl(jobs, map, nblocks, lanes, slot_cfg, lds, s);
}
hipError_t launch_pool(const TileArgs<float>* jobs, const uint32_t* map, uint32_t nblocks,
                       uint32_t lanes, int slot_cfg, size_t lds, hipStream_t s) {
  return launch_pool_impl(jobs, map, nblocks, lanes, slot_cfg, lds, s);
}
size_t tile_lds_bytes(size_t elem, uint32_t lanes, bool adjoint, uint32_t W, uint32_t nforced) {
  return elem == 8 ? lds_bytes<double>(lanes, adjoint, W, nforced)
                   : lds_bytes<float>(lanes, adjoint, W, nforced);
}

hipError_t launch_tiles(const TileArgs<double>& a, uint32_t ntiles, int slot_cfg, hipStream_t s) {
  return launch_tiles_impl(a, ntiles, slot_cfg, s);
}
hipError_t launch_tiles(const TileArgs<float>& a, uint32_t ntiles, int slot_cfg, hipStream_t s) {
  return launch_tiles_impl(a, ntiles, slot_cfg, s);
}

// ---------------------------------------------------------------- big path
// Systems too large for a workgroup's lanes: one workgroup per instance, rows
// strided over its threads, every vector of the instance in global memory
// and a barrier between stages (written by some threads, read by all).
// The arithmetic is the tile kernel's: kinds summed in order, each from zero
// in entry order, then the force; the same RK4 update, adjoint recursions
// and checkpoint schedule.

// u of a forced row at (step n, stage), read from the excitation: zero-order
// hold e[n] (stages 1-3) / e[n+1] (stage 4) when 0 < s < L (src/kernels.cu:
// 83-86), or the interpolation on the half-step grid (src/kernels.cu:114-124)
// with e[L] read as 0.
template <typename R>
__device__ __forceinline__ R big_force(const BigArgs<R>& a, int32_t exc, uint32_t n, int stage) {
  if (exc < 0) return R(0);
  const R* E = a.excit + static_cast<size_t>(exc) * a.exc_len;
  const uint32_t Lx = a.exc_len;
  const uint32_t n1 = n + (stage == 3 ? 1u : 0u);
  if (a.interp_w == 0u) return (n1 > 0u && n1 < Lx) ? E[n1] : R(0);
  const uint32_t W = a.interp_w;
  const uint32_t hs = 2u * n1 + ((stage == 1 || stage == 2) ? 1u : 0u);
  const uint32_t i = hs / W, j = hs - i * W;
  if (!(i > 0u && i < Lx)) return R(0);
  const R ei = E[i];
  const R ei1 = (i + 1u < Lx) ? E[i + 1u] : R(0);
  return fmaR(a.interp[W + j], ei, a.interp[j] * ei1);
}

// Sum of terms [b, e) of kind k, in entry order from zero, four at a time:
// the four terms' index words and values are requested together, then their
// factors, so a segment of 8 terms costs four dependent memory round trips
// instead of two per term.  One copy (not inlined into every stage).
template <typename R>
__device__ __noinline__ R big_terms(const R* val, const uint32_t* idx, uint32_t b, uint32_t e,
                                    uint32_t k, const R* S, const R* X) {
  R t = R(0);
  for (uint32_t j0 = b; j0 < e; j0 += 4u) {
    const uint32_t c = min(4u, e - j0);
    uint4 ix[4];
    R v[4], f0[4], f1[4], f2[4], f3[4];
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j)
      if (j < c) {
        ix[j] = *reinterpret_cast<const uint4*>(idx + 4u * (j0 + j));
        v[j] = val[j0 + j];
      }
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j)
      if (j < c) {
        f0[j] = S[ix[j].x];
        f1[j] = k >= 1u ? X[ix[j].y] : R(0);
        f2[j] = k >= 2u ? X[ix[j].z] : R(0);
        f3[j] = k >= 3u ? X[ix[j].w] : R(0);
      }
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j)
      if (j < c) {
        if (k == 0u) {
          t = fmaR(v[j], f0[j], t);
        } else {
          R w = v[j] * f0[j];
          if (k >= 2u) w = w * f1[j];
          if (k >= 3u) w = w * f2[j];
          t = fmaR(w, k == 1u ? f1[j] : (k == 2u ? f2[j] : f3[j]), t);
        }
      }
  }
  return t;
}

// Right-hand side of row r (every kind at most kBigSegTerms terms): S
// factors from `S`, set-point factors from `X`; kinds added in order.
template <typename R>
__device__ __forceinline__ R big_row(const uint32_t* rp, const R* val, const uint32_t* idx,
                                     uint32_t r, const R* S, const R* X) {
  R part[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    part[k] = big_terms(val, idx, rp[4u * r + k], rp[4u * r + k + 1u], uint32_t(k), S, X);
  return ((part[0] + part[1]) + part[2]) + part[3];
}

// Sum of the terms of segment g (one kind of one row, at most kBigSegTerms
// terms, in entry order from zero).
template <typename R>
__device__ __forceinline__ R big_seg(const uint32_t* sg, const R* val, const uint32_t* idx,
                                     uint32_t g, const R* S, const R* X) {
  return big_terms(val, idx, sg[3u * g], sg[3u * g + 1u], sg[3u * g + 2u], S, X);
}

// Right-hand side of row r from its segments' partial sums: each kind the
// sum of its segments in order, kinds added in order.
template <typename R>
__device__ __forceinline__ R big_row_parts(const uint32_t* sp, const R* part, uint32_t r) {
  R pk[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    R t = R(0);
    const uint32_t g1 = sp[4u * r + k + 1u];
    uint32_t g = sp[4u * r + k];
    // eight partial sums requested together, added in order
    for (; g + 8u <= g1; g += 8u) {
      R x[8];
#pragma unroll
      for (uint32_t j = 0; j < 8u; ++j) x[j] = part[g + j];
#pragma unroll
      for (uint32_t j = 0; j < 8u; ++j) t = t + x[j];
    }
    for (; g < g1; ++g) t = t + part[g];
    pk[k] = t;
  }
  return ((pk[0] + pk[1]) + pk[2]) + pk[3];
}

template <typename R, bool ADJ>
__global__ void __launch_bounds__(kBigThreads) big_kernel(BigArgs<R> a) {
  const BigInst I = a.inst[blockIdx.x];
  const uint32_t tid = threadIdx.x, NT = blockDim.x;
  const uint32_t nf = I.n_f, na = I.n_a, T = I.steps;
  const bool adjoint = ADJ && a.adjoint != 0u;
  const bool exact = adjoint && a.exact != 0u;
  const uint32_t K = adjoint ? a.ck_every : 1u;
  const uint32_t* rpf = a.fwd.rptr + I.rptr_f;
  const R* vf = a.fwd.val + I.term_f;
  const uint32_t* xf = a.fwd.idx + 4u * I.term_f;
  const R* ff = a.fwd.force + I.row_f;
  const uint32_t* sgf = a.fwd.seg + 3u * I.sseg_f;
  const uint32_t* spf = a.fwd.sptr + I.sptr_f;
  extern __shared__ unsigned char big_smem[];
  R* const lds_part = reinterpret_cast<R*>(big_smem);
  R* pf = a.part_lds ? lds_part : a.part + I.part_f;
  const bool splitf = I.split_f != 0u;
  R* q = a.vec_f + I.scr_f;
  R* y0 = q + nf;
  R* y1 = y0 + nf;
  R* acc = y1 + nf;
  uint32_t* pst = a.progress + 4u * blockIdx.x;
  uint32_t n = pst[0], phase = pst[1], seg = pst[2];
  uint32_t budget = a.budget;
  const R h = a.h, h2 = a.h2, h6 = a.h6;

  if (n == 0u && phase == kPhaseForward) {
    for (uint32_t r = tid; r < nf; r += NT) {
      q[r] = a.fwd.init[I.row_f + r];
      if (T == 0u) a.out_q[I.out_f + r] = q[r];
    }
  }
  __syncthreads();
  while (budget > 0u && phase != kPhaseDone) {
    if (phase != kPhaseBackward) {
      // ---------------- forward / recompute ----------------
      const bool recompute = phase == kPhaseRecompute;
      const uint32_t s0 = seg * K;
      if (recompute && n == s0) {
        for (uint32_t r = tid; r < nf; r += NT) {
          q[r] = a.ckpt[I.ck + static_cast<size_t>(seg) * nf + r];
          if (!exact) a.segbuf[I.seg + r] = q[r];
        }
        __syncthreads();
      }
      const uint32_t e = recompute ? min(T, s0 + K) : T;
      const uint32_t nend = (e - n < budget) ? e : n + budget;
      budget -= nend - n;
      for (; n < nend; ++n) {
        if (adjoint && !recompute && (n % K) == 0u)
          for (uint32_t r = tid; r < nf; r += NT)
            a.ckpt[I.ck + static_cast<size_t>(n / K) * nf + r] = q[r];
        R* rec = (exact && recompute)
                     ? a.segbuf + I.seg + static_cast<size_t>(4u * (n - s0)) * nf : nullptr;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const R* Y = s == 0 ? q : (s == 2 ? y1 : y0);
          R* Yn = s == 1 ? y1 : y0;
          R u = R(0);
          if (I.exc >= 0) u = big_force(a, I.exc, n, s);
          if (splitf) {
            // rows with many terms: their segments' sums on every thread first
            for (uint32_t g = tid; g < I.nseg_f; g += NT) pf[g] = big_seg(sgf, vf, xf, g, Y, Y);
            __syncthreads();
          }
          for (uint32_t r = tid; r < nf; r += NT) {
            const R p = splitf ? big_row_parts(spf, pf, r) : big_row(rpf, vf, xf, r, Y, Y);
            const R k = fmaR(ff[r], u, p);
            const R qr = q[r];
            if (s == 0) {
              acc[r] = k;
              const R y = fmaR(h2, k, qr);
              Yn[r] = y;
              if (rec) {
                rec[r] = qr;
                rec[nf + r] = y;
              }
            } else if (s == 1) {
              acc[r] = fmaR(R(2), k, acc[r]);
              const R y = fmaR(h2, k, qr);
              Yn[r] = y;
              if (rec) rec[2u * nf + r] = y;
            } else if (s == 2) {
              acc[r] = fmaR(R(2), k, acc[r]);
              const R y = fmaR(h, k, qr);
              Yn[r] = y;
              if (rec) rec[3u * nf + r] = y;
            } else {
              acc[r] = acc[r] + k;
            }
          }
          __syncthreads();
        }
        for (uint32_t r = tid; r < nf; r += NT) {
          const R qn = fmaR(h6, acc[r], q[r]);
          q[r] = qn;
          if (recompute) {
            if (!exact) a.segbuf[I.seg + static_cast<size_t>(n + 1u - s0) * nf + r] = qn;
          } else if (n + 1u == T) {
            a.out_q[I.out_f + r] = qn;
          }
        }
        __syncthreads();
      }
      if (n == e) {
        if (recompute) {
          phase = kPhaseBackward;
        } else if (adjoint && T > 0u) {
          seg = (T - 1u) / K;
          n = seg * K;
          phase = kPhaseRecompute;
          for (uint32_t r = tid; r < na; r += NT) a.vec_a[I.scr_a + r] = a.adj.init[I.row_a + r];
        } else {
          phase = kPhaseDone;
          if (adjoint)
            for (uint32_t r = tid; r < na; r += NT)
              a.out_grad[I.out_a + r] = a.adj.init[I.row_a + r];
        }
      }
      __syncthreads();
    } else if (ADJ) {
      // ---------------- adjoint (reverse) ----------------
      const uint32_t* rpa = a.adj.rptr + I.rptr_a;
      const R* va = a.adj.val + I.term_a;
      const uint32_t* xa = a.adj.idx + 4u * I.term_a;
      const uint32_t* sga = a.adj.seg + 3u * I.sseg_a;
      const uint32_t* spa = a.adj.sptr + I.sptr_a;
      R* pa = a.part_lds ? lds_part : a.part + I.part_a;
      const bool splita = I.split_a != 0u;
      R* lam = a.vec_a + I.scr_a;
      R* z0 = lam + na;
      R* z1 = z0 + na;
      R* za = z1 + na;
      R* X = a.setp + I.scr_f;
      const uint32_t s0 = seg * K;
      const uint32_t nend = (n - s0 < budget) ? s0 : n - budget;
      budget -= n - nend;
      for (; n > nend; --n) {
        for (uint32_t r = tid; r < nf; r += NT) {
          if (exact) {
            const R* y = a.segbuf + I.seg + static_cast<size_t>(4u * (n - 1u - s0)) * nf;
#pragma unroll
            for (int v = 0; v < 4; ++v) X[v * nf + r] = y[v * nf + r];
          } else {
            const R xn = a.segbuf[I.seg + static_cast<size_t>(n - s0) * nf + r];
            const R xp = a.segbuf[I.seg + static_cast<size_t>(n - 1u - s0) * nf + r];
            X[r] = xn;
            X[nf + r] = R(0.5) * (xn + xp);
            X[2u * nf + r] = xp;
          }
        }
        if (exact)
          for (uint32_t r = tid; r < na; r += NT) z0[r] = h6 * lam[r];
        __syncthreads();
        const bool active = n <= T;
        if (!exact) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const R* Y = s == 0 ? lam : (s == 2 ? z1 : z0);
            R* Yn = s == 1 ? z1 : z0;
            const R* Xs = X + (s == 0 ? 0u : (s == 3 ? 2u : 1u)) * nf;
            if (splita) {
              for (uint32_t g = tid; g < I.nseg_a; g += NT) pa[g] = big_seg(sga, va, xa, g, Y, Xs);
              __syncthreads();
            }
            for (uint32_t r = tid; r < na; r += NT) {
              const R p = splita ? big_row_parts(spa, pa, r) : big_row(rpa, va, xa, r, Y, Xs);
              const R l = lam[r];
              if (s == 0) {
                za[r] = p;
                Yn[r] = fmaR(h2, p, l);
              } else if (s == 1) {
                za[r] = fmaR(R(2), p, za[r]);
                Yn[r] = fmaR(h2, p, l);
              } else if (s == 2) {
                za[r] = fmaR(R(2), p, za[r]);
                Yn[r] = fmaR(h, p, l);
              } else {
                za[r] = za[r] + p;
              }
            }
            __syncthreads();
          }
          for (uint32_t r = tid; r < na; r += NT) {
            const R ln = fmaR(h6, za[r], lam[r]);
            if (active) lam[r] = ln;
          }
        } else {
          const R h3 = R(2) * h6;
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const R* Y = (s & 1) ? z1 : z0;
            R* Yn = (s & 1) ? z0 : z1;
            const R* Xs = X + static_cast<uint32_t>(3 - s) * nf;
            if (splita) {
              for (uint32_t g = tid; g < I.nseg_a; g += NT) pa[g] = big_seg(sga, va, xa, g, Y, Xs);
              __syncthreads();
            }
            for (uint32_t r = tid; r < na; r += NT) {
              const R p = splita ? big_row_parts(spa, pa, r) : big_row(rpa, va, xa, r, Y, Xs);
              const R l = lam[r];
              if (s == 0) {
                za[r] = p;
                Yn[r] = fmaR(h, p, h3 * l);
              } else if (s == 1) {
                za[r] = za[r] + p;
                Yn[r] = fmaR(h2, p, h3 * l);
              } else if (s == 2) {
                za[r] = za[r] + p;
                Yn[r] = fmaR(h2, p, h6 * l);
              } else {
                za[r] = za[r] + p;
              }
            }
            __syncthreads();
          }
          for (uint32_t r = tid; r < na; r += NT) {
            const R ln = lam[r] + za[r];
            if (active) lam[r] = ln;
          }
        }
        __syncthreads();
      }
      if (n == s0) {
        if (seg == 0u) {
          phase = kPhaseDone;
          for (uint32_t r = tid; r < na; r += NT) a.out_grad[I.out_a + r] = lam[r];
        } else {
          seg -= 1u;
          n = seg * K;
          phase = kPhaseRecompute;
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0u) {
    pst[0] = n;
    pst[1] = phase;
    pst[2] = seg;
  }
}

template <typename R>
static hipError_t launch_big_impl(const BigArgs<R>& a, uint32_t ninst, size_t lds,
                                  hipStream_t s) {
  if (ninst == 0u) return hipSuccess;
  if (lds > kBigPartLds) return hipErrorInvalidValue;
  if (a.adjoint)
    hipLaunchKernelGGL((big_kernel<R, true>), dim3(ninst), dim3(kBigThreads), lds, s, a);
  else
    hipLaunchKernelGGL((big_kernel<R, false>), dim3(ninst), dim3(kBigThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_big(const BigArgs<double>& a, uint32_t ninst, size_t lds, hipStream_t s) {
  return launch_big_impl(a, ninst, lds, s);
}
hipError_t launch_big(const BigArgs<float>& a, uint32_t ninst, size_t lds, hipStream_t s) {
  return launch_big_impl(a, ninst, lds, s);
}

}  // namespace pasif

// Run scratch of pasif_solve_noalloc: one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (on the
// host the name is a stand-in, not a device address): scratch_region looks
// up each device's copy once, with that device current, and every use goes
// through that address.
static __device__ __align__(256) unsigned char scratch[pasif::kScratchBytes];

namespace pasif {
hipError_t scratch_region(int dev, void** ptr, size_t* bytes) {
  static std::mutex mu;
  static std::vector<void*> addr;
  std::lock_guard<std::mutex> lk(mu);
  if (dev < 0) return hipErrorInvalidDevice;
  if (size_t(dev) >= addr.size()) addr.resize(size_t(dev) + 1, nullptr);
  if (!addr[dev]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    addr[dev] = p;
  }
  *ptr = addr[dev];
  *bytes = kScratchBytes;
  return hipSuccess;
}
}  // namespace pasif
