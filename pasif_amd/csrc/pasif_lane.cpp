// Lane kernels: source generation per structure class and hipRTC
// compilation with a per-process and an on-disk cache.
//
// The generated kernel restates the same RK4 step and adjoint schedule as
// the tile kernel (src/__GpuDriver.cu:693-713, 746-783; src/kernels.cu:21-63,
// 71-126, 162-229): k_s = sum of the row's terms + P u_s, y = q + c_s k_s,
// acc = ((k1 + 2 k2) + 2 k3) + k4, q <- q + h6 acc; forward checkpoints every
// K steps, segment recompute, reverse steps with set points x_n, xbar,
// x_{n-1}.  Only the data layout differs: one lane per instance, state in
// registers, the row loops unrolled with compile-time indices.
#include "pasif_lane.hpp"

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

namespace pasif {

uint32_t LaneSide::coefs() const {
  size_t t = forced.size();
  for (const auto& r : rows) t += r.size();
  return uint32_t(t);
}

uint32_t LaneShape::live_values() const {
  const uint32_t fwd = 4 * f.n + f.coefs();
  const uint32_t adj = adjoint ? 4 * a.n + 4 * f.n + a.coefs() : 0;
  return fwd > adj ? fwd : adj;
}

namespace {

// Device-side runtime shared by every class (R, LaneArgs and the class
// policies are defined before it).
const char* kLaneRuntime = R"LANE(
#define GP(T) __attribute__((address_space(1))) T*
__device__ __forceinline__ u32 uni(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ u32 umin(u32 a, u32 b) { return a < b ? a : b; }
enum : u32 { kForward = 0, kRecompute = 1, kBackward = 2, kDone = 3 };
template <bool B> struct Flag { static constexpr bool value = B; };

// e[b .. b + STAGE) of one excitation (indices clamped into [0, L))
__device__ __forceinline__ void fetch(R (&v)[STAGE], GP(const R) E, u32 Lx, u32 b) {
#pragma unroll
  for (u32 u = 0; u < STAGE; ++u) {
    const u32 i = b + u;
    v[u] = E[i < Lx ? i : Lx - 1u];
  }
}
// into the lane's LDS column; 0 beyond L and for unforced lanes
__device__ __forceinline__ void put(R* eb, u32 lane, const R (&v)[STAGE], u32 Lx, u32 b,
                                    bool forced) {
#pragma unroll
  for (u32 u = 0; u < STAGE; ++u)
    eb[u * WAVE + lane] = (forced && b + u < Lx) ? v[u] : R(0);
}

// Mi[1][j] e[i] + Mi[0][j] e[i+1] on the half step i0*W + j0 + t, zero
// unless 0 < i < L (src/kernels.cu:114-124)
__device__ __forceinline__ R interp_at(const R* eb, u32 lane, u32 b, GP(const R) wts, u32 W,
                                       u32 Lx, u32 i0, u32 j0, u32 t) {
  u32 j = j0 + t;
  const bool next = j >= W;
  j = next ? j - W : j;
  const u32 i = i0 + (next ? 1u : 0u);
  const R ei = eb[(i - b) * WAVE + lane];
  const R ei1 = eb[(i + 1u - b) * WAVE + lane];
  return (i > 0u && i < Lx) ? fmaR(wts[W + j], ei, wts[j] * ei1) : R(0);
}

// one RK4 step of N rows; rhs(k, y, stage).  ALL: every lane is active.
template <int N, bool ALL, class F>
__device__ __forceinline__ void rk4(R* q, F rhs, R h, R h2, R h6, bool active) {
  R y[N], k[N], acc[N];
#pragma unroll
  for (int r = 0; r < N; ++r) y[r] = q[r];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    rhs(k, y, s);
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if (s == 0) {
        acc[r] = k[r];
        y[r] = fmaR(h2, k[r], q[r]);
      } else if (s == 1) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        y[r] = fmaR(h2, k[r], q[r]);
      } else if (s == 2) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        y[r] = fmaR(h, k[r], q[r]);
      } else {
        acc[r] = acc[r] + k[r];
        const R qn = fmaR(h6, acc[r], q[r]);
        q[r] = (ALL || active) ? qn : q[r];
      }
    }
  }
}

// One wave of one class through its phases; same schedule and resumable
// state as the tile kernel.  LDS: eb [STAGE][WAVE] staged samples of the
// current chunk of steps (the next chunk is prefetched into registers),
// then seg [K+1][NF][WAVE] (adjoint) the recomputed segment.
template <class S>
__device__ __forceinline__ void run(const LaneArgs& a, u32 wave, R* smem) {
  constexpr int NF = S::NF;
  constexpr int NA = S::NA > 0 ? S::NA : 1;
  constexpr int TF = S::TF > 0 ? S::TF : 1;
  constexpr int TA = S::TA > 0 ? S::TA : 1;
  const u32 lane = threadIdx.x;
  const u64 st = a.stride;
  const u32 i = wave * WAVE + lane;
  const bool valid = i < (u32)a.count;
  R* eb = smem;
  R* seg = smem + STAGE * WAVE;
  const bool adjoint = S::ADJ && a.adjoint != 0ull;
  const u32 K = adjoint ? (u32)a.ck_every : 1u;
  const u32 W = (u32)a.interp_w;
  const u32 Lx = (u32)a.exc_len;
  const u32 SE = (u32)a.save_every;
  const R h = R(a.h), h2 = R(a.h2), h6 = R(a.h6);
  GP(const u32) steps = (GP(const u32))a.steps;
  // instances are sorted by step count, descending: the wave's first lane
  // has the most steps, its last valid lane the fewest
  const u32 T = uni(steps[wave * WAVE]);
  const u32 Tmin = uni(steps[umin(wave * WAVE + WAVE, (u32)a.count) - 1u]);
  const u32 fsteps = steps[i];
  const i32 exc = ((GP(const i32))a.exc)[i];
  const bool forced = exc >= 0;
  const u32 fout = ((GP(const u32))a.out_f)[i];
  const u32 aout = ((GP(const u32))a.out_a)[i];
  GP(const R) cf = (GP(const R))a.coef_f + i;
  GP(const R) ca = (GP(const R))a.coef_a + i;
  GP(const R) qi = (GP(const R))a.init_f + i;
  GP(const R) li = (GP(const R))a.init_a + i;
  GP(R) qs = (GP(R))a.qstate + i;
  GP(R) ms = (GP(R))a.mxstate + i;
  GP(R) ls = (GP(R))a.lstate + i;
  GP(R) ck = (GP(R))a.ckpt + i;
  GP(R) sb = (GP(R))a.segbuf + i;
  GP(R) outq = (GP(R))a.out_q;
  GP(R) outm = (GP(R))a.out_max;
  GP(R) outg = (GP(R))a.out_grad;
  GP(R) traj = (GP(R))a.traj;
  GP(const R) wts = (GP(const R))a.interp;
  GP(const R) E = (GP(const R))a.excit + (u64)(forced ? exc : 0) * Lx;
  GP(u32) pst = (GP(u32))a.progress + 4 * wave;
  u32 n = uni(pst[0]);
  u32 phase = uni(pst[1]);
  u32 sg = uni(pst[2]);
  u32 budget = (u32)a.budget;

  R q[NF], mx[NF], lam[NA];
#pragma unroll
  for (int r = 0; r < NF; ++r) mx[r] = R(0);
#pragma unroll
  for (int r = 0; r < NA; ++r) lam[r] = R(0);
  if (n == 0u && phase == kForward) {
#pragma unroll
    for (int r = 0; r < NF; ++r) q[r] = qi[r * st];
    if (fsteps == 0u && valid)
#pragma unroll
      for (int r = 0; r < NF; ++r) outq[fout + r] = q[r];
  } else {
#pragma unroll
    for (int r = 0; r < NF; ++r) {
      q[r] = qs[r * st];
      if (a.window) mx[r] = ms[r * st];
    }
  }
  if (adjoint && phase != kForward)
#pragma unroll
    for (int r = 0; r < NA; ++r) lam[r] = ls[r * st];
  bool seg_here = false;  // LDS holds the current segment up to step n

  while (budget > 0u && phase != kDone) {
    if (phase != kBackward) {
      const bool recompute = phase == kRecompute;
      R c[TF];
#pragma unroll
      for (int t = 0; t < S::TF; ++t) c[t] = cf[t * st];
      const u32 s0 = sg * K;
      if (recompute) {
        if (n == s0) {
#pragma unroll
          for (int r = 0; r < NF; ++r) {
            q[r] = ck[((u64)sg * NF + r) * st];
            seg[r * WAVE + lane] = q[r];
            sb[r * st] = q[r];
          }
        } else if (!seg_here) {
          // resumed inside a recompute: the LDS copy of the steps recomputed
          // by earlier launches comes back from HBM
          for (u32 m = 0; m <= n - s0; ++m)
#pragma unroll
            for (int r = 0; r < NF; ++r)
              seg[(m * NF + r) * WAVE + lane] = sb[((u64)m * NF + r) * st];
        }
        seg_here = true;
      }
      const u32 e = recompute ? umin(T, s0 + K) : T;
      const u32 nend = (e - n < budget) ? e : n + budget;
      budget -= nend - n;
      const u32 nfirst = n;
      // per-step bookkeeping (checkpoints, segment, trajectory, window,
      // per-lane output) only when a run needs it
      const bool gen = recompute || adjoint || SE != 0u || a.window != 0ull;
      u32 next_ck = ((n + K - 1u) / K) * K;
      u32 next_save = SE ? ((n + SE - 1u) / SE) * SE : 0u;
      // forcing cursor: half-step position 2n = hi*W + hj, staged chunk
      u32 hi = W ? (2u * n) / W : 0u, hj = W ? 2u * n - hi * W : 0u;
      u32 ebase = 0, left = 0, pbase = 0;
      bool have = false;
      R pre[STAGE];
      auto steps_to = [&](auto all_t, auto gen_t, auto zoh_t, u32 stop) {
        constexpr bool ALL = decltype(all_t)::value;
        constexpr bool GEN = decltype(gen_t)::value;
        constexpr bool ZOH = decltype(zoh_t)::value;
        for (; n < stop; ++n) {
          R u0, u1, u3;
          if (ZOH) {
            const R e0 = eb[(n - ebase) * WAVE + lane], e1 = eb[(n + 1u - ebase) * WAVE + lane];
            u0 = (n > 0u && n < Lx) ? e0 : R(0);
            u1 = u0;
            u3 = e1;
          } else {
            u0 = interp_at(eb, lane, ebase, wts, W, Lx, hi, hj, 0u);
            u1 = interp_at(eb, lane, ebase, wts, W, Lx, hi, hj, 1u);
            u3 = interp_at(eb, lane, ebase, wts, W, Lx, hi, hj, 2u);
            hj += 2u;
            if (hj >= W) {
              hj -= W;
              hi += 1u;
            }
          }
          if (GEN && adjoint && !recompute && n == next_ck) {
#pragma unroll
            for (int r = 0; r < NF; ++r) ck[((u64)(n / K) * NF + r) * st] = q[r];
            next_ck += K;
          }
          rk4<NF, ALL>(q, [&](R* k, const R* y, int s) {
            S::rhs_f(k, y, c, s == 0 ? u0 : (s == 3 ? u3 : u1));
          }, h, h2, h6, n < fsteps);
          if (GEN) {
            if (recompute) {
              const u32 m = n + 1u - s0;
#pragma unroll
              for (int r = 0; r < NF; ++r) {
                seg[(m * NF + r) * WAVE + lane] = q[r];
                sb[((u64)m * NF + r) * st] = q[r];
              }
            } else {
              if (SE && n == next_save) {
                const u32 slot = n / SE;
                if (slot < (u32)a.nslots && valid)
#pragma unroll
                  for (int r = 0; r < NF; ++r) traj[(u64)slot * a.out_len + fout + r] = q[r];
                next_save += SE;
              }
              if (a.window && (ALL || n < fsteps) && n + 1u + (u32)a.window > fsteps)
#pragma unroll
                for (int r = 0; r < NF; ++r) mx[r] = fmaxR(mx[r], absR(q[r]));
              if (n + 1u == fsteps && valid) {
#pragma unroll
                for (int r = 0; r < NF; ++r) {
                  outq[fout + r] = q[r];
                  if (a.window) outm[fout + r] = mx[r];
                }
              }
            }
          }
        }
      };
      while (n < nend) {
        if (left == 0u) {
          // a new chunk: its samples come from the prefetch (or are loaded
          // now), and the next chunk's are requested
          const u32 base = W ? hi : n;
          if (have && pbase == base) {
            put(eb, lane, pre, Lx, base, forced);
          } else {
            R v[STAGE];
            fetch(v, E, Lx, base);
            put(eb, lane, v, Lx, base, forced);
          }
          ebase = base;
          left = CHUNK;
          pbase = W ? (2u * (n + CHUNK)) / W : n + CHUNK;
          fetch(pre, E, Lx, pbase);
          have = true;
        }
        u32 stop = umin(nend, n + left);
        if (n < Tmin && Tmin < stop) stop = Tmin;
        left -= stop - n;
        const bool all = n < Tmin;
        if (gen) {
          if (W == 0u) {
            if (all) steps_to(Flag<true>(), Flag<true>(), Flag<true>(), stop);
            else steps_to(Flag<false>(), Flag<true>(), Flag<true>(), stop);
          } else {
            if (all) steps_to(Flag<true>(), Flag<true>(), Flag<false>(), stop);
            else steps_to(Flag<false>(), Flag<true>(), Flag<false>(), stop);
          }
        } else {
          if (W == 0u) {
            if (all) steps_to(Flag<true>(), Flag<false>(), Flag<true>(), stop);
            else steps_to(Flag<false>(), Flag<false>(), Flag<true>(), stop);
          } else {
            if (all) steps_to(Flag<true>(), Flag<false>(), Flag<false>(), stop);
            else steps_to(Flag<false>(), Flag<false>(), Flag<false>(), stop);
          }
        }
      }
      // without per-step bookkeeping: a lane whose last step fell in this
      // range kept its final state since (inactive steps leave it unchanged)
      if (!gen && valid && fsteps > nfirst && fsteps <= nend)
#pragma unroll
        for (int r = 0; r < NF; ++r) outq[fout + r] = q[r];
      if (n == e) {
        if (recompute) {
          phase = kBackward;
        } else if (adjoint && T > 0u) {
          sg = (T - 1u) / K;
          n = sg * K;
          phase = kRecompute;
#pragma unroll
          for (int r = 0; r < NA; ++r) lam[r] = li[r * st];
        } else {
          phase = kDone;
          if (adjoint && valid)
#pragma unroll
            for (int r = 0; r < S::NA; ++r) outg[aout + r] = li[r * st];
        }
      }
    } else if (S::ADJ) {
      R c[TA];
#pragma unroll
      for (int t = 0; t < S::TA; ++t) c[t] = ca[t * st];
      const u32 s0 = sg * K;
      if (!seg_here) {
        // resumed in the reverse phase: reload the segment this launch needs
        for (u32 m = 0; m <= n - s0; ++m)
#pragma unroll
          for (int r = 0; r < NF; ++r) seg[(m * NF + r) * WAVE + lane] = sb[((u64)m * NF + r) * st];
        seg_here = true;
      }
      const u32 nend = (n - s0 < budget) ? s0 : n - budget;
      budget -= n - nend;
      auto back_to = [&](auto all_t, u32 stop) {
        constexpr bool ALL = decltype(all_t)::value;
        for (; n > stop; --n) {
          R xc[NF], xb[NF], xp[NF];
#pragma unroll
          for (int r = 0; r < NF; ++r) {
            xc[r] = seg[((n - s0) * NF + r) * WAVE + lane];
            xp[r] = seg[((n - 1u - s0) * NF + r) * WAVE + lane];
            xb[r] = R(0.5) * (xc[r] + xp[r]);
          }
          rk4<NA, ALL>(lam, [&](R* k, const R* y, int s) {
            S::rhs_a(k, y, s == 0 ? xc : (s == 3 ? xp : xb), c);
          }, h, h2, h6, n <= fsteps);
        }
      };
      // steps n > Tmin leave some lanes inactive; from Tmin down all are active
      if (n > Tmin) back_to(Flag<false>(), Tmin > nend ? Tmin : nend);
      if (n > nend) back_to(Flag<true>(), nend);
      if (n == s0) {
        if (sg == 0u) {
          phase = kDone;
          if (valid)
#pragma unroll
            for (int r = 0; r < S::NA; ++r) outg[aout + r] = lam[r];
        } else {
          sg -= 1u;
          n = sg * K;
          phase = kRecompute;
          seg_here = false;
        }
      }
    } else {
      phase = kDone;  // a class without an adjoint program never reverses
    }
  }
#pragma unroll
  for (int r = 0; r < NF; ++r) {
    qs[r * st] = q[r];
    if (a.window) ms[r * st] = mx[r];
  }
  if (adjoint)
#pragma unroll
    for (int r = 0; r < S::NA; ++r) ls[r * st] = lam[r];
  if (lane == 0u) {
    pst[0] = n;
    pst[1] = phase;
    pst[2] = sg;
  }
}
)LANE";

std::string type_text(const char* t) {
  const std::string s(t);
  if (s == "uint64_t") return "u64";
  return s;
}

// Right-hand side of one side of a class: k[r] for every row.  Factor 0 of
// each term reads `S`, the others `X`.
void emit_rhs(std::ostringstream& o, const LaneSide& sd, bool forced_input, const char* S,
              const char* X) {
  uint32_t t = 0;
  std::vector<uint32_t> fslot(sd.n, ~0u);
  {
    uint32_t nterms = 0;
    for (const auto& r : sd.rows) nterms += uint32_t(r.size());
    for (size_t i = 0; i < sd.forced.size(); ++i) fslot[sd.forced[i]] = nterms + uint32_t(i);
  }
  for (uint32_t r = 0; r < sd.n; ++r) {
    const auto& terms = sd.rows[r];
    std::string sums[4];
    for (int kind = 0; kind < 4; ++kind) {
      std::string acc;
      for (const LaneTerm& tm : terms) {
        if (int(tm.kind) != kind) continue;
        std::ostringstream prod;
        const uint32_t slot = t++;
        if (kind == 0) {
          if (acc.empty())
            prod << "c[" << slot << "] * " << S << "[" << tm.f[0] << "]";
          else
            prod << "fmaR(c[" << slot << "], " << S << "[" << tm.f[0] << "], " << acc << ")";
        } else {
          std::ostringstream lead;
          lead << "c[" << slot << "] * " << S << "[" << tm.f[0] << "]";
          std::string l = lead.str();
          for (int f = 1; f < int(kind); ++f)
            l = "(" + l + ") * " + X + "[" + std::to_string(tm.f[f]) + "]";
          const std::string last = std::string(X) + "[" + std::to_string(tm.f[kind]) + "]";
          if (acc.empty())
            prod << "(" << l << ") * " << last;
          else
            prod << "fmaR(" << l << ", " << last << ", " << acc << ")";
        }
        const std::string name = "s" + std::to_string(kind) + "_" + std::to_string(r);
        if (acc.empty())
          o << "    R " << name << " = " << prod.str() << ";\n";
        else
          o << "    " << name << " = " << prod.str() << ";\n";
        acc = name;
      }
      sums[kind] = acc;
    }
    std::string k;
    for (int kind = 0; kind < 4; ++kind) {
      if (sums[kind].empty()) continue;
      k = k.empty() ? sums[kind] : "(" + k + " + " + sums[kind] + ")";
    }
    if (k.empty()) k = "R(0)";
    if (forced_input && fslot[r] != ~0u)
      k = "fmaR(c[" + std::to_string(fslot[r]) + "], u, " + k + ")";
    o << "    k[" << r << "] = " << k << ";\n";
  }
}

uint64_t fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_dir() {
  if (const char* e = std::getenv("PASIF_JIT_CACHE")) return e;
  if (const char* hm = std::getenv("HOME")) return std::string(hm) + "/.cache/pasif_amd/jit";
  return std::string();
}

void mkdirs(const std::string& path) {
  std::string cur;
  std::stringstream ss(path);
  std::string part;
  if (!path.empty() && path[0] == '/') cur = "/";
  while (std::getline(ss, part, '/')) {
    if (part.empty()) continue;
    cur += part + "/";
    (void)mkdir(cur.c_str(), 0755);
  }
}

const char* kJitOptions[] = {"--gpu-architecture=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
constexpr int kJitOptionCount = sizeof(kJitOptions) / sizeof(kJitOptions[0]);

std::mutex g_jit_mu;
std::map<std::pair<int, uint64_t>, LaneKernel> g_jit;

}  // namespace

std::string lane_source(const std::vector<LaneShape>& shapes, int bits, uint32_t seg_len) {
  std::ostringstream o;
  o << "// generated lane kernel: " << shapes.size() << " structure class(es), FP" << bits
    << ", segment " << seg_len << "\n";
  o << "typedef unsigned int u32;\ntypedef int i32;\ntypedef unsigned long long u64;\n";
  o << "typedef " << (bits == 64 ? "double" : "float") << " R;\n";
  o << "struct LaneArgs {\n";
#define PASIF_LANE_TEXT(T, name) o << "  " << type_text(#T) << " " << #name << ";\n";
  PASIF_LANE_FIELDS(PASIF_LANE_TEXT)
#undef PASIF_LANE_TEXT
  o << "};\n";
  o << "#define WAVE " << kLaneWave << "u\n#define CHUNK " << kLaneChunk << "u\n#define STAGE "
    << kLaneStage << "u\n";
  if (bits == 64) {
    o << "__device__ __forceinline__ R fmaR(R a, R b, R c) { return __builtin_fma(a, b, c); }\n"
         "__device__ __forceinline__ R absR(R a) { return __builtin_fabs(a); }\n"
         "__device__ __forceinline__ R fmaxR(R a, R b) { return __builtin_fmax(a, b); }\n";
  } else {
    o << "__device__ __forceinline__ R fmaR(R a, R b, R c) { return __builtin_fmaf(a, b, c); }\n"
         "__device__ __forceinline__ R absR(R a) { return __builtin_fabsf(a); }\n"
         "__device__ __forceinline__ R fmaxR(R a, R b) { return __builtin_fmaxf(a, b); }\n";
  }
  o << kLaneRuntime;
  for (size_t c = 0; c < shapes.size(); ++c) {
    const LaneShape& sh = shapes[c];
    o << "struct C" << c << " {\n";
    o << "  enum { NF = " << sh.f.n << ", NA = " << (sh.adjoint ? sh.a.n : 0)
      << ", TF = " << sh.f.coefs() << ", TA = " << (sh.adjoint ? sh.a.coefs() : 0)
      << ", ADJ = " << (sh.adjoint ? 1 : 0) << " };\n";
    o << "  static __device__ __forceinline__ void rhs_f(R* k, const R* y, const R* c, R u) {\n";
    emit_rhs(o, sh.f, true, "y", "y");
    o << "  }\n";
    o << "  static __device__ __forceinline__ void rhs_a(R* k, const R* y, const R* x, const R* c) "
         "{\n";
    if (sh.adjoint) emit_rhs(o, sh.a, false, "y", "x");
    o << "  }\n};\n";
  }
  o << "extern \"C\" __global__ void __launch_bounds__(" << kLaneWave
    << ") pasif_lane_kernel(const LaneArgs* __restrict__ cls, u32 ncls) {\n"
       "  extern __shared__ R smem[];\n"
       "  const u32 b = blockIdx.x;\n"
       "  u32 c = 0;\n"
       "  while (c + 1u < ncls && b >= (u32)cls[c + 1u].block_begin) ++c;\n"
       "  c = uni(c);\n"
       "  switch (c) {\n";
  for (size_t c = 0; c < shapes.size(); ++c)
    o << "    case " << c << ": run<C" << c << ">(cls[" << c << "], b - (u32)cls[" << c
      << "].block_begin, smem); break;\n";
  o << "    default: break;\n  }\n}\n";
  (void)seg_len;
  return o.str();
}

size_t lane_lds_bytes(const std::vector<LaneShape>& shapes, int bits, uint32_t seg_len) {
  uint32_t nf = 0;
  bool adj = false;
  for (const LaneShape& s : shapes) {
    if (s.adjoint) {
      adj = true;
      nf = s.f.n > nf ? s.f.n : nf;
    }
  }
  const size_t el = bits == 64 ? 8 : 4;
  size_t v = size_t(kLaneStage) * kLaneWave;
  if (adj) v += size_t(seg_len + 1) * nf * kLaneWave;
  return v * el;
}

namespace {
// hipRTC compilation of `src` to a gfx950 code object
bool rtc_compile(const std::string& src, std::vector<char>* code, std::string* err) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "pasif_lane.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    *err = "hipRTC: cannot create the lane kernel program";
    return false;
  }
  const hiprtcResult rc = hiprtcCompileProgram(prog, kJitOptionCount, kJitOptions);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    *err = std::string("hipRTC compilation of the lane kernel failed: ") +
           hiprtcGetErrorString(rc) + "\n" + log.substr(0, 4000);
    return false;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code->resize(cs);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return true;
}
}  // namespace

bool lane_compile_only(const std::string& src, size_t* code_bytes, std::string* err) {
  std::vector<char> code;
  if (!rtc_compile(src, &code, err)) return false;
  if (code_bytes) *code_bytes = code.size();
  return true;
}

bool lane_compile(const std::string& src, int dev, LaneKernel* out, std::string* err) {
  std::string opts;
  for (int i = 0; i < kJitOptionCount; ++i) opts += std::string(kJitOptions[i]) + " ";
  const uint64_t key = fnv1a(src + opts);
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_jit.find({dev, key});
  if (it != g_jit.end()) {
    *out = it->second;
    return true;
  }
  std::vector<char> code;
  const std::string dir = cache_dir();
  char name[64];
  snprintf(name, sizeof name, "lane_%016llx.co", (unsigned long long)key);
  const std::string path = dir.empty() ? std::string() : dir + "/" + name;
  if (!path.empty()) {
    if (FILE* f = fopen(path.c_str(), "rb")) {
      fseek(f, 0, SEEK_END);
      const long sz = ftell(f);
      fseek(f, 0, SEEK_SET);
      if (sz > 0) {
        code.resize(size_t(sz));
        if (fread(code.data(), 1, code.size(), f) != code.size()) code.clear();
      }
      fclose(f);
    }
  }
  if (code.empty()) {
    if (!rtc_compile(src, &code, err)) return false;
    if (!path.empty()) {
      mkdirs(dir);
      const std::string tmp = path + ".tmp" + std::to_string(long(getpid()));
      if (FILE* f = fopen(tmp.c_str(), "wb")) {
        const bool ok = fwrite(code.data(), 1, code.size(), f) == code.size();
        fclose(f);
        if (ok) std::rename(tmp.c_str(), path.c_str());
        else std::remove(tmp.c_str());
      }
    }
  }
  LaneKernel k;
  hipError_t e = hipModuleLoadData(&k.mod, code.data());
  if (e != hipSuccess) {
    *err = std::string("loading the lane kernel module failed: ") + hipGetErrorName(e);
    return false;
  }
  e = hipModuleGetFunction(&k.fn, k.mod, "pasif_lane_kernel");
  if (e != hipSuccess) {
    *err = std::string("lane kernel not found in its module: ") + hipGetErrorName(e);
    return false;
  }
  g_jit[{dev, key}] = k;
  *out = k;
  return true;
}

}  // namespace pasif
