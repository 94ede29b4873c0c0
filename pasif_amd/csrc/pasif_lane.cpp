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

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <sys/stat.h>
#include <thread>
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

// LDS stage slot of every row that a term of `sd` reads as its own-state
// factor (and, for the forward side, as a set point); -1 for rows nothing
// reads (accumulators), which need no slot.
std::vector<int> read_slots(const LaneSide& sd, bool x_is_self) {
  std::vector<bool> used(sd.n, false);
  for (const auto& row : sd.rows)
    for (const LaneTerm& t : row) {
      used[t.f[0]] = true;
      if (x_is_self)
        for (uint32_t a = 1; a <= t.kind; ++a) used[t.f[a]] = true;
    }
  std::vector<int> slot(sd.n, -1);
  int k = 0;
  for (uint32_t r = 0; r < sd.n; ++r)
    if (used[r]) slot[r] = k++;
  return slot;
}

uint32_t lane_partition(LaneShape& sh, uint32_t waves) {
  std::vector<uint32_t> cf(waves, 0), ca(waves, 0), rws(waves, 0);
  auto part = [&](const LaneSide& sd, std::vector<uint32_t>& owner, std::vector<uint32_t>& coefs) {
    owner.assign(sd.n, 0);
    std::vector<uint32_t> cost(sd.n), order(sd.n);
    std::vector<bool> forced(sd.n, false);
    for (uint32_t r : sd.forced) forced[r] = true;
    for (uint32_t r = 0; r < sd.n; ++r) {
      uint32_t c = 3 + (forced[r] ? 2 : 0);
      for (const LaneTerm& t : sd.rows[r]) c += t.kind + 2;
      cost[r] = c;
      order[r] = r;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
    std::vector<uint64_t> load(waves, 0);
    for (uint32_t r : order) {
      uint32_t w = 0;
      for (uint32_t v = 1; v < waves; ++v)
        if (load[v] < load[w]) w = v;
      owner[r] = w;
      load[w] += cost[r];
      coefs[w] += uint32_t(sd.rows[r].size()) + (forced[r] ? 1 : 0);
      rws[w] += 1;
    }
  };
  part(sh.f, sh.fwave, cf);
  if (sh.adjoint) part(sh.a, sh.awave, ca);
  uint32_t worst = 0;
  for (uint32_t w = 0; w < waves; ++w) worst = std::max(worst, std::max(cf[w], ca[w]) + 3 * rws[w]);
  return worst;
}

uint32_t GSide::coefs() const {
  uint32_t t = 0;
  for (const GRow& r : rows) t += uint32_t(r.terms.size()) + (r.force.empty() ? 0u : 1u);
  return t;
}

namespace {

uint32_t log2u(uint32_t G) {
  uint32_t l = 0;
  while ((1u << l) < G) ++l;
  return l;
}

// Bands of G consecutive rows and shared rows of a side, by dynamic
// programming over per-stage cost estimates: a band costs the union of its
// rows' terms (terms compared by kind and factor offsets from their row), a
// shared row its terms spread over G lanes plus the butterfly.
void plan_bands(const LaneSide& sd, uint32_t G, uint32_t lo, std::vector<uint32_t>* band,
                std::vector<uint32_t>* shared) {
  const uint32_t n = sd.n, lg = log2u(G);
  auto key_of = [](const LaneTerm& t, uint32_t g) {
    std::vector<int64_t> k{int64_t(t.kind)};
    for (uint32_t a = 0; a <= t.kind; ++a) k.push_back(int64_t(t.f[a]) - int64_t(g));
    return k;
  };
  auto band_cost = [&](uint32_t g) {
    std::map<std::pair<std::vector<int64_t>, int>, int> uni;
    double c = 2.0;
    for (uint32_t p = 0; p < G; ++p) {
      std::map<std::vector<int64_t>, int> occ;
      for (const LaneTerm& t : sd.rows[g + p]) {
        std::vector<int64_t> k = key_of(t, g + p);
        const int o = occ[k]++;
        if (uni.emplace(std::make_pair(k, o), 1).second) c += t.kind + 1.0;
      }
    }
    return c;  // every lane of the band runs the union once
  };
  auto shared_cost = [&](uint32_t g) {
    double c = 0;
    for (const LaneTerm& t : sd.rows[g]) c += t.kind + 1.0;
    return (G > 1 ? std::ceil(c / G) + 3.0 * lg : c) + 2.0;
  };
  std::vector<double> best(n + 1, 0.0);
  std::vector<uint8_t> take(n + 1, 0);
  for (uint32_t g = n; g-- > lo;) {
    best[g] = shared_cost(g) + best[g + 1];
    if (g + G <= n) {
      const double b = band_cost(g) + best[g + G];
      if (b <= best[g]) {
        best[g] = b;
        take[g] = 1;
      }
    }
  }
  for (uint32_t g = lo; g < n;) {
    if (take[g]) {
      band->push_back(g);
      g += G;
    } else {
      shared->push_back(g);
      g += 1;
    }
  }
}

// Blocks of G m rows from lo while they fit, position p holding rows
// m p .. m p + m - 1 of each (m = 1: bands of G consecutive rows); the rest
// shared.
void blocked_bands(uint32_t n, uint32_t G, uint32_t lo, uint32_t m, std::vector<uint32_t>* band,
                   std::vector<uint32_t>* stride, std::vector<uint32_t>* shared) {
  uint32_t g = lo;
  for (; g + G * m <= n; g += G * m)
    for (uint32_t j = 0; j < m; ++j) {
      band->push_back(g + j);
      stride->push_back(m);
    }
  for (; g < n; ++g) shared->push_back(g);
}

// Lane slot and group position of every row of a side.
struct GMap {
  std::vector<uint32_t> slot, pos;
  std::vector<bool> shared;
};

GMap group_map(uint32_t n, uint32_t G, const GSide& gs) {
  GMap m;
  m.slot.assign(n, 0);
  m.pos.assign(n, 0);
  m.shared.assign(n, false);
  for (uint32_t r = 0; r < gs.RL; ++r)
    for (uint32_t p = 0; p < G; ++p) {
      m.slot[gs.band[r] + p * gs.stride[r]] = r;
      m.pos[gs.band[r] + p * gs.stride[r]] = p;
    }
  for (uint32_t s = 0; s < gs.NS; ++s) {
    m.slot[gs.shared[s]] = gs.RL + s;
    m.shared[gs.shared[s]] = true;
  }
  return m;
}

GSrc src_of(const GMap& m, uint32_t row, uint32_t p, uint32_t G) {
  GSrc s;
  s.slot = m.slot[row];
  s.shared = m.shared[row];
  s.d = s.shared ? 0u : (m.pos[row] + G - p) % G;
  return s;
}

// Union programs of a side laid out as gs.band / gs.shared; `own` maps the
// side's rows, `setp` the rows its set-point factors index.
void build_gside(const LaneSide& sd, uint32_t G, const GMap& own, const GMap& setp, bool forward,
                 GSide* gs) {
  std::vector<uint32_t> base(sd.n);
  std::vector<int32_t> force(sd.n, -1);
  uint32_t t = 0;
  for (uint32_t r = 0; r < sd.n; ++r) {
    base[r] = t;
    t += uint32_t(sd.rows[r].size());
  }
  for (size_t i = 0; i < sd.forced.size(); ++i) force[sd.forced[i]] = int32_t(t + i);
  gs->rows.assign(gs->nl(), GRow());
  using Key = std::vector<uint32_t>;
  auto add_term = [&](GRow& row, std::map<Key, size_t>& where, std::map<Key, uint32_t>& occ,
                      const LaneTerm& tm, uint32_t p, int32_t coef) {
    GTerm gt;
    gt.kind = tm.kind;
    Key k{tm.kind};
    for (uint32_t a = 0; a <= tm.kind; ++a) gt.f[a] = src_of(a == 0 ? own : setp, tm.f[a], p, G);
    // one monomial, one slot: the factors in a canonical order relative to
    // the lane (all of them on the forward side, where every factor is the
    // state; the set-point factors on the adjoint side)
    std::sort(gt.f + (forward ? 0 : 1), gt.f + tm.kind + 1);
    for (uint32_t a = 0; a <= tm.kind; ++a) {
      k.push_back(gt.f[a].shared ? 1u : 0u);
      k.push_back(gt.f[a].slot);
      k.push_back(gt.f[a].d);
    }
    const uint32_t o = occ[k]++;
    k.push_back(o);
    auto it = where.find(k);
    if (it == where.end()) {
      gt.coef.assign(G, -1);
      row.terms.push_back(gt);
      it = where.emplace(k, row.terms.size() - 1).first;
    }
    row.terms[it->second].coef[p] = coef;
  };
  for (uint32_t r = 0; r < gs->RL; ++r) {
    GRow& row = gs->rows[r];
    std::map<Key, size_t> where;
    for (uint32_t p = 0; p < G; ++p) {
      const uint32_t g = gs->band[r] + p * gs->stride[r];
      std::map<Key, uint32_t> occ;
      for (size_t e = 0; e < sd.rows[g].size(); ++e)
        add_term(row, where, occ, sd.rows[g][e], p, int32_t(base[g] + e));
      if (force[g] >= 0) {
        if (row.force.empty()) row.force.assign(G, -1);
        row.force[p] = force[g];
      }
    }
  }
  for (uint32_t s = 0; s < gs->NS; ++s) {
    const uint32_t g = gs->shared[s];
    GRow& row = gs->rows[gs->RL + s];
    std::map<Key, size_t> where;
    std::vector<std::map<Key, uint32_t>> occ(G);
    for (size_t e = 0; e < sd.rows[g].size(); ++e) {
      const LaneTerm& tm = sd.rows[g][e];
      // the term runs on the position owning its first banded factor
      uint32_t pt = 0;
      for (uint32_t a = 0; a <= tm.kind; ++a) {
        const GMap& m = a == 0 ? own : setp;
        if (!m.shared[tm.f[a]]) {
          pt = m.pos[tm.f[a]];
          break;
        }
      }
      add_term(row, where, occ[pt], tm, pt, int32_t(base[g] + e));
    }
    if (force[g] >= 0) {
      row.force.assign(G, -1);
      row.force[0] = force[g];
    }
  }
}

uint32_t remote_count(const GSide& gs, bool setpoints_own) {
  std::vector<GSrc> v;
  for (const GRow& row : gs.rows)
    for (const GTerm& t : row.terms)
      for (uint32_t a = 0; a <= (setpoints_own ? t.kind : 0u); ++a)
        if (!t.f[a].shared && t.f[a].d != 0) v.push_back(t.f[a]);
  std::sort(v.begin(), v.end());
  return uint32_t(std::unique(v.begin(), v.end()) - v.begin());
}

}  // namespace

namespace {

// Per-stage vector instructions of one lane of a group side: its union
// terms, kind sums, forces, butterflies of shared rows and rotations.
double gside_stage_cost(const GSide& gs, uint32_t G, bool setpoints_own, int bits) {
  const uint32_t lg = log2u(G);
  const double words = bits == 64 ? 2.0 : 1.0;
  const double rot = G == 8 ? 3.0 * words : (G >= 32 ? 2.0 * words : words);
  double stage = 2.0 * gs.nl();
  for (uint32_t r = 0; r < gs.nl(); ++r) {
    const GRow& row = gs.rows[r];
    bool kinds[4] = {false, false, false, false};
    for (const GTerm& t : row.terms) {
      stage += t.kind + 1.0;
      kinds[t.kind] = true;
    }
    const int nk = kinds[0] + kinds[1] + kinds[2] + kinds[3];
    stage += nk > 1 ? nk - 1 : 0;
    stage += row.force.empty() ? 0 : 1;
    if (r >= gs.RL) stage += lg * (rot + 1.0);
  }
  return stage + remote_count(gs, setpoints_own) * rot;
}

// The cheapest of the candidate band plans of a side (each built exactly).
// `mirror` (adjoint side): the forward plan, whose rows the first rows of
// the adjoint state mirror.
bool best_gside(const LaneSide& sd, uint32_t G, const GSide* mirror, uint32_t nmirror,
                const GMap* setp, int bits, GSide* out) {
  struct Plan {
    std::vector<uint32_t> band, stride, shared;
  };
  std::vector<Plan> plans;
  // rows from lo: dynamic programming over consecutive bands, or aligned
  // blocks of G m rows for m = 1, 2, 4, ...
  auto from = [&](const Plan& head, uint32_t lo) {
    Plan pl = head;
    plan_bands(sd, G, lo, &pl.band, &pl.shared);
    pl.stride.resize(pl.band.size(), 1u);
    plans.push_back(pl);
    for (uint32_t m = 1; lo + G * m <= sd.n; m *= 2) {
      Plan q = head;
      blocked_bands(sd.n, G, lo, m, &q.band, &q.stride, &q.shared);
      plans.push_back(q);
    }
  };
  from(Plan(), 0);
  if (mirror && nmirror <= sd.n) from(Plan{mirror->band, mirror->stride, mirror->shared}, nmirror);
  double best = -1;
  for (auto& pl : plans) {
    GSide gs;
    gs.band = pl.band;
    gs.stride = pl.stride;
    gs.shared = pl.shared;
    std::sort(gs.shared.begin(), gs.shared.end());
    gs.RL = uint32_t(gs.band.size());
    gs.NS = uint32_t(gs.shared.size());
    if (gs.RL == 0) continue;
    const GMap own = group_map(sd.n, G, gs);
    build_gside(sd, G, own, setp ? *setp : own, setp == nullptr, &gs);
    const double c = gside_stage_cost(gs, G, setp == nullptr, bits);
    if (best < 0 || c < best) {
      best = c;
      *out = std::move(gs);
    }
  }
  return best >= 0;
}

}  // namespace

namespace {
// Group layouts already planned in this process, by structure and group
// size (planning builds several candidate programs per size).
struct GroupPlan {
  bool ok = false;
  GSide gf, ga;
  std::vector<GSrc> gx;
};
std::mutex g_plan_mu;
std::map<std::string, GroupPlan> g_plans;

void key_side(const LaneSide& sd, std::string& k) {
  auto put = [&](uint32_t v) { k.append(reinterpret_cast<const char*>(&v), sizeof v); };
  put(sd.n);
  for (const auto& row : sd.rows) {
    put(uint32_t(row.size()));
    for (const LaneTerm& t : row) {
      put(t.kind);
      for (uint32_t a = 0; a <= t.kind; ++a) put(t.f[a]);
    }
  }
  put(uint32_t(sd.forced.size()));
  for (uint32_t r : sd.forced) put(r);
}
bool plan_group(LaneShape& sh, uint32_t G);
}  // namespace

bool lane_group(LaneShape& sh, uint32_t G) {
  if (G <= 1) return plan_group(sh, G);
  std::string key(reinterpret_cast<const char*>(&G), sizeof G);
  key.push_back(sh.adjoint ? 'a' : 'f');
  key_side(sh.f, key);
  if (sh.adjoint) key_side(sh.a, key);
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
      sh.gf = it->second.gf;
      sh.ga = it->second.ga;
      sh.gx = it->second.gx;
      sh.group = it->second.ok ? G : 1u;
      return it->second.ok;
    }
  }
  GroupPlan gp;
  gp.ok = plan_group(sh, G);
  gp.gf = sh.gf;
  gp.ga = sh.ga;
  gp.gx = sh.gx;
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_plans.emplace(std::move(key), std::move(gp));
  return sh.group == G;
}

namespace {
bool plan_group(LaneShape& sh, uint32_t G) {
  sh.gf = GSide();
  sh.ga = GSide();
  sh.gx.clear();
  sh.group = 1;
  if (G == 1) return true;
  if (G > kLaneWave || (kLaneWave % G) != 0 || (G & (G - 1)) != 0) return false;
  if (!best_gside(sh.f, G, nullptr, 0, nullptr, 64, &sh.gf)) return false;
  if (sh.adjoint) {
    const GMap mf = group_map(sh.f.n, G, sh.gf);
    if (!best_gside(sh.a, G, &sh.gf, sh.f.n, &mf, 64, &sh.ga)) return false;
    for (const GRow& row : sh.ga.rows)
      for (const GTerm& t : row.terms)
        for (uint32_t a = 1; a <= t.kind; ++a) sh.gx.push_back(t.f[a]);
    std::sort(sh.gx.begin(), sh.gx.end());
    sh.gx.erase(std::unique(sh.gx.begin(), sh.gx.end()), sh.gx.end());
  }
  // forward shared slots no forward term and no adjoint set point reads
  sh.gf.sink.assign(sh.gf.nl(), 0);
  for (uint32_t r = sh.gf.RL; r < sh.gf.nl(); ++r) sh.gf.sink[r] = 1;
  for (const GRow& row : sh.gf.rows)
    for (const GTerm& t : row.terms)
      for (uint32_t a = 0; a <= t.kind; ++a)
        if (t.f[a].shared) sh.gf.sink[t.f[a].slot] = 0;
  for (const GSrc& x : sh.gx)
    if (x.shared) sh.gf.sink[x.slot] = 0;
  sh.ga.sink.assign(sh.ga.nl(), 0);
  sh.group = G;
  return true;
}
}  // namespace

uint32_t lane_seg_rows(const LaneShape& sh, std::vector<int>* map) {
  const bool grp = sh.group > 1;
  const uint32_t nf = grp ? sh.gf.nl() : sh.f.n;
  std::vector<int> m(nf, -1);
  if (sh.adjoint) {
    if (grp) {
      for (const GSrc& x : sh.gx) m[x.slot] = 0;
    } else {
      for (const auto& row : sh.a.rows)
        for (const LaneTerm& t : row)
          for (uint32_t a = 1; a <= t.kind; ++a) m[t.f[a]] = 0;
    }
  }
  uint32_t rows = 0;
  for (int& v : m)
    if (v == 0) v = int(rows++);
  if (map) *map = m;
  return std::max(1u, rows);
}

uint32_t lane_live(const LaneShape& sh) {
  if (sh.group <= 1) return sh.live_values();
  const uint32_t fwd = 4 * sh.gf.nl() + sh.gf.coefs() + remote_count(sh.gf, true);
  const uint32_t adj = sh.adjoint ? 4 * sh.ga.nl() + sh.ga.coefs() + remote_count(sh.ga, false) +
                                        3 * uint32_t(sh.gx.size()) + sh.gf.nl()
                                  : 0;
  return std::max(fwd, adj);
}

double lane_step_cost(const LaneShape& sh, int bits) {
  const uint32_t G = sh.group;
  auto side = [&](const LaneSide& sd, const GSide& gs, bool setpoints_own) {
    double stage = 0, rows = 0;
    if (G <= 1) {
      rows = sd.n;
      for (const auto& row : sd.rows) {
        bool kinds[4] = {false, false, false, false};
        for (const LaneTerm& t : row) {
          stage += t.kind + 1.0;
          kinds[t.kind] = true;
        }
        const int nk = kinds[0] + kinds[1] + kinds[2] + kinds[3];
        stage += nk > 1 ? nk - 1 : 0;
      }
      stage += double(sd.forced.size()) + 2.0 * rows;
    } else {
      rows = gs.nl();
      stage = gside_stage_cost(gs, G, setpoints_own, bits);
    }
    return 4.0 * stage + 2.0 * rows;
  };
  const double forcing = 14.0, loop = 8.0;
  const double f = side(sh.f, sh.gf, true) + forcing + loop;
  if (!sh.adjoint) return f;
  const double nx = G <= 1 ? sh.f.n : double(sh.gx.size());
  const double a = side(sh.a, sh.ga, false) + 3.0 * nx + loop;
  return 2.0 * f + a;
}

void lane_group_coefs(const GSide& gs, const std::vector<double>& sys, uint32_t p,
                      std::vector<double>& out) {
  out.clear();
  for (const GRow& row : gs.rows) {
    for (const GTerm& t : row.terms) out.push_back(t.coef[p] >= 0 ? sys[size_t(t.coef[p])] : 0.0);
    if (!row.force.empty()) out.push_back(row.force[p] >= 0 ? sys[size_t(row.force[p])] : 0.0);
  }
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
// per-step bookkeeping of a forward step loop: none, checkpoints (forward
// pass of an adjoint run), the recomputed segment, or trajectory / window /
// per-step outputs (with checkpoints when the run also has an adjoint)
enum : u32 { mPlain = 0, mCk = 1, mRec = 2, mFull = 3 };
template <u32 M> struct Mode { static constexpr u32 value = M; };

// ------------------------------------------------------------ lane groups
// A value moved between lanes by a DPP control (32-bit halves for doubles).
// A lane whose source falls outside its row reads 0 (bound_ctrl); no old
// value is kept, so no register is initialised for it.
template <int CTRL, class T>
__device__ __forceinline__ T dpp(T x) {
  if constexpr (sizeof(T) == 8) {
    const u64 b = __builtin_bit_cast(u64, x);
    const u32 lo = (u32)__builtin_amdgcn_mov_dpp((int)(u32)b, CTRL, 0xF, 0xF, true);
    const u32 hi = (u32)__builtin_amdgcn_mov_dpp((int)(u32)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(T, ((u64)hi << 32) | (u64)lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xF, 0xF,
                                                          true));
  }
}
// through the LDS crossbar (no LDS storage): lane l reads lane src
template <class T>
__device__ __forceinline__ T bperm(T x, u32 src) {
  const int addr = (int)(src << 2);
  if constexpr (sizeof(T) == 8) {
    const u64 b = __builtin_bit_cast(u64, x);
    const u32 lo = (u32)__builtin_amdgcn_ds_bpermute(addr, (int)(u32)b);
    const u32 hi = (u32)__builtin_amdgcn_ds_bpermute(addr, (int)(u32)(b >> 32));
    return __builtin_bit_cast(T, ((u64)hi << 32) | (u64)lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, x)));
  }
}
// quad_perm control: lane l of each quad reads lane (l + D) mod G of its
// aligned group of G (G = 2 or 4)
constexpr int quad_ctrl(u32 G, u32 D) {
  int c = 0;
  for (u32 l = 0; l < 4u; ++l) c |= (int)(((l & ~(G - 1u)) | ((l + D) & (G - 1u))) & 3u) << (2u * l);
  return c;
}
// Value of position (p + D) mod G of the lane's aligned group of G lanes.
template <u32 G, u32 D>
__device__ __forceinline__ R rot(R x) {
  if constexpr (D % G == 0u) {
    return x;
  } else if constexpr (G == 2u || G == 4u) {
    return dpp<quad_ctrl(G, D), R>(x);
  } else if constexpr (G == 8u) {
    // row_shl:D where p + D stays in the group, row_shr:(8 - D) where it wraps
    const R a = dpp<0x100 + (int)D, R>(x);
    const R b = dpp<0x110 + 8 - (int)D, R>(x);
    return ((threadIdx.x & 7u) + D < 8u) ? a : b;
  } else if constexpr (G == 16u) {
    return dpp<0x120 + 16 - (int)D, R>(x);  // row_ror:(16 - D)
  } else {
    const u32 l = threadIdx.x & 63u;
    return bperm<R>(x, (l & ~(G - 1u)) | ((l + D) & (G - 1u)));
  }
}
// Sum over the group by an xor butterfly: lanes add their partner's value
// in the same order of operands up to commutation, so every lane of the
// group ends with bitwise the same sum.
template <u32 G>
__device__ __forceinline__ R gsum(R x) {
  if constexpr (G >= 2u) x = x + rot<2u, 1u>(x);
  if constexpr (G >= 4u) x = x + rot<4u, 2u>(x);
  if constexpr (G >= 8u) x = x + rot<8u, 4u>(x);
  if constexpr (G >= 16u) x = x + rot<16u, 8u>(x);
  if constexpr (G >= 32u) x = x + rot<32u, 16u>(x);
  if constexpr (G >= 64u) x = x + rot<64u, 32u>(x);
  return x;
}
// lane of position (p + D) mod G of the lane's group
template <u32 G, u32 D>
__device__ __forceinline__ u32 glane(u32 lane) {
  return (lane & ~(G - 1u)) | ((lane + D) & (G - 1u));
}


// e[b .. b + STAGE) of one excitation (indices clamped into [0, L))
__device__ __forceinline__ void fetch(R (&v)[STAGE], GP(const R) E, u32 Lx, u32 b) {
#pragma unroll
  for (u32 u = 0; u < STAGE; ++u) {
    const u32 i = b + u;
    v[u] = E[i < Lx ? i : Lx - 1u];
  }
}
// into the lane's LDS column (`e` points at the lane's first sample, rows
// SS apart); 0 beyond L and for unforced lanes
template <u32 SS>
__device__ __forceinline__ void put(R* e, const R (&v)[STAGE], u32 Lx, u32 b, bool forced) {
#pragma unroll
  for (u32 u = 0; u < STAGE; ++u) e[u * SS] = (forced && b + u < Lx) ? v[u] : R(0);
}

// Samples of a group's chunk: position p of a group of G lanes loads
// e[b + u] for u = p, p + G, ... < STAGE and stores them into the group's
// column of the stage [STAGE][WAVE / G] (`e` points at it); the group then
// reads its samples there.  Under zero-order hold (zoh) sample 0 is staged
// as 0: it only ever drives step 0's first stages, which take no forcing.
template <u32 G, u32 NPF>
__device__ __forceinline__ void fetch_g(R (&v)[NPF], GP(const R) E, u32 Lx, u32 b, u32 p) {
#pragma unroll
  for (u32 k = 0; k < NPF; ++k) {
    const u32 i = b + p + k * G;
    v[k] = E[i < Lx ? i : Lx - 1u];
  }
}
template <u32 G, u32 NPF>
__device__ __forceinline__ void put_g(R* e, const R (&v)[NPF], u32 Lx, u32 b, bool forced, u32 p,
                                      bool zoh) {
#pragma unroll
  for (u32 k = 0; k < NPF; ++k) {
    const u32 u = p + k * G;
    if (u < STAGE)
      e[u * (WAVE / G)] = (forced && b + u < Lx && !(zoh && b + u == 0u)) ? v[k] : R(0);
  }
  if constexpr (G > 1u) {
    // the other lanes of the group read these samples next
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  }
}

// Mi[1][j] e[i] + Mi[0][j] e[i+1] on the half step i0*W + j0 + t, zero
// unless 0 < i < L (src/kernels.cu:114-124)
template <u32 SS>
__device__ __forceinline__ R interp_at(const R* e, u32 b, GP(const R) wts, u32 W, u32 i0, u32 j0,
                                       u32 t) {
  u32 j = j0 + t;
  const bool next = j >= W;
  j = next ? j - W : j;
  const u32 i = i0 + (next ? 1u : 0u);
  const R ei = e[(i - b) * SS];
  const R ei1 = e[(i + 1u - b) * SS];
  // samples past the end are staged as 0; i = 0 is masked by a product with
  // a uniform 0 or 1 (a select here turns into a branch around the loads,
  // at which a lone wave waits for them at once)
  const R keep = i > 0u ? R(1) : R(0);
  return fmaR(wts[W + j], ei, wts[j] * ei1) * keep;
}

// one RK4 step of N rows; rhs(k, y, stage).  ALL: every lane is active.
// K::sink(r): row r is a shared row that no term reads; rhs leaves its
// group-partial slope in k[r], which is accumulated per lane over the four
// stages and summed across the group (K::G lanes) once per step.
struct NoSink {
  static constexpr unsigned G = 1;
  static __device__ __forceinline__ constexpr bool sink(int) { return false; }
};
template <int N, bool ALL, class K, class F>
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
        if (!K::sink(r)) y[r] = fmaR(h2, k[r], q[r]);
      } else if (s == 1) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        if (!K::sink(r)) y[r] = fmaR(h2, k[r], q[r]);
      } else if (s == 2) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        if (!K::sink(r)) y[r] = fmaR(h, k[r], q[r]);
      } else {
        acc[r] = acc[r] + k[r];
        if (K::sink(r)) acc[r] = gsum<K::G>(acc[r]);
        const R qn = fmaR(h6, acc[r], q[r]);
        q[r] = (ALL || active) ? qn : q[r];
      }
    }
  }
}

// rk4 that hands the inputs of stages 2..4 to rec(s, y) before evaluating
// them (the exact discrete adjoint's recompute keeps them)
template <int N, bool ALL, class K, class F, class Rec>
__device__ __forceinline__ void rk4rec(R* q, F rhs, Rec rec, R h, R h2, R h6, bool active) {
  R y[N], k[N], acc[N];
#pragma unroll
  for (int r = 0; r < N; ++r) y[r] = q[r];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s > 0) rec(s, y);
    rhs(k, y, s);
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if (s == 0) {
        acc[r] = k[r];
        if (!K::sink(r)) y[r] = fmaR(h2, k[r], q[r]);
      } else if (s == 1) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        if (!K::sink(r)) y[r] = fmaR(h2, k[r], q[r]);
      } else if (s == 2) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        if (!K::sink(r)) y[r] = fmaR(h, k[r], q[r]);
      } else {
        acc[r] = acc[r] + k[r];
        if (K::sink(r)) acc[r] = gsum<K::G>(acc[r]);
        const R qn = fmaR(h6, acc[r], q[r]);
        q[r] = (ALL || active) ? qn : q[r];
      }
    }
  }
}

// One wave of one class through its phases; same schedule and resumable
// state as the tile kernel.  S::G lanes integrate one instance (1: one
// instance per lane); per-lane arrays (coefficients, states, checkpoints)
// are indexed by the lane's column, per-instance ones (steps, excitation,
// output bases) by the instance.  LDS: eb [STAGE][WAVE / G] staged samples
// of the current chunk of steps, one column per instance (the next chunk is
// prefetched into registers), then seg [K+1][NSEG][WAVE] (adjoint) the
// recomputed segment's post-step states and, for the exact discrete
// adjoint, seg3 [K][3][NSEG][WAVE] each step's stage inputs y2, y3, y4 (HBM
// mirror sb likewise, by forward slot).
template <class S>
__device__ __forceinline__ void run(const LaneArgs& a, u32 wave, R* smem) {
  constexpr int NF = S::NF;
  constexpr int NA = S::NA > 0 ? S::NA : 1;
  constexpr int TF = S::TF > 0 ? S::TF : 1;
  constexpr int TA = S::TA > 0 ? S::TA : 1;
  constexpr int NX = S::NX > 0 ? S::NX : 1;
  constexpr u32 NSEG = S::NSEG;  // rows per step of the recomputed segment
  constexpr u32 G = S::G;
  constexpr u32 IPW = WAVE / G;  // instances per wave
  constexpr u32 NPF = (STAGE + G - 1u) / G;  // samples each lane of a group stages
  // groups with interpolated forcing: the chunk's forcing values (3 per
  // step) computed once by the group's lanes into a table
  constexpr u32 FT = G > 1u ? 3u * CHUNK : 0u;
  constexpr u32 NFT = (FT + G - 1u) / G;
  const u32 lane = threadIdx.x;
  const u32 p = lane & (G - 1u);
  const u64 st = a.stride;
  const u32 col = wave * WAVE + lane;
  const u32 i = wave * IPW + lane / G;
  const bool valid = i < (u32)a.count;
  R* eb = smem + lane / G;                // the instance's staged samples
  R* ft = smem + STAGE * IPW + lane / G;  // its forcing table (groups)
  R* seg = smem + (STAGE + FT) * IPW;
  const bool adjoint = S::ADJ && a.adjoint != 0ull;
  const u32 K = adjoint ? (u32)a.ck_every : 1u;
  constexpr bool EX = EXACT != 0 && S::ADJ;
  R* seg3 = seg + (K + 1u) * NSEG * WAVE;
  const u32 W = (u32)a.interp_w;
  const u32 Lx = (u32)a.exc_len;
  const u32 SE = (u32)a.save_every;
  const R h = R(a.h), h2 = R(a.h2), h6 = R(a.h6);
  GP(const u32) steps = (GP(const u32))a.steps;
  // instances are sorted by step count, descending: the wave's first
  // instance has the most steps, its last valid one the fewest
  const u32 T = uni(steps[wave * IPW]);
  const u32 Tmin = uni(steps[umin(wave * IPW + IPW, (u32)a.count) - 1u]);
  const u32 fsteps = steps[i];
  const i32 exc = ((GP(const i32))a.exc)[i];
  const bool forced = exc >= 0;
  const u32 fout = ((GP(const u32))a.out_f)[i];
  const u32 aout = ((GP(const u32))a.out_a)[i];
  const u32 nao = ((GP(const u32))a.na_out)[i];
  GP(const R) cf = (GP(const R))a.coef_f + col;
  GP(const R) ca = (GP(const R))a.coef_a + col;
  GP(const R) qi = (GP(const R))a.init_f + col;
  GP(const R) li = (GP(const R))a.init_a + col;
  GP(R) qs = (GP(R))a.qstate + col;
  GP(R) ms = (GP(R))a.mxstate + col;
  GP(R) ls = (GP(R))a.lstate + col;
  // checkpoints: the wave's block [nck_w][NF][WAVE] (sized by its longest instance)
  GP(R) ck = (GP(R))a.ckpt + (adjoint ? ((GP(const u64))a.ck_off)[wave] : 0ull) + lane;
  GP(R) sb = (GP(R))a.segbuf + col;
  GP(R) sb3 = sb + (u64)(K + 1u) * NF * st;
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
  // output row of lane slot r: band slots hold row rowf(r) + p stridef(r), a
  // shared slot's row is written by position 0 only (-1 elsewhere)
  auto orow_f = [&](int r) -> int {
    return r < S::RLF ? S::rowf(r) + (int)p * S::stridef(r) : (p == 0u ? S::rowf(r) : -1);
  };
  auto orow_a = [&](int r) -> int {
    return r < S::RLA ? S::rowa(r) + (int)p * S::stridea(r) : (p == 0u ? S::rowa(r) : -1);
  };

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
      for (int r = 0; r < NF; ++r)
        if (orow_f(r) >= 0) outq[fout + orow_f(r)] = q[r];
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
  // a launch that may stop mid-segment mirrors the segment to HBM so the
  // next launch can rebuild its LDS copy; one that runs to the end does not
  const bool mirror = a.budget != 0xFFFFFFFFull;

  while (budget > 0u && phase != kDone) {
    if (phase != kBackward) {
      const bool recompute = phase == kRecompute;
      R c[TF];
#pragma unroll
      for (int t = 0; t < S::TF; ++t) c[t] = cf[t * st];
      const u32 s0 = sg * K;
      // the checkpoint a recompute starts from and the samples of the
      // phase's first chunk are requested with the coefficients
      if (recompute && n == s0)
#pragma unroll
        for (int r = 0; r < NF; ++r) q[r] = ck[((u64)sg * NF + r) * WAVE];
      // forcing cursor: half-step position 2n = hi*W + hj, staged chunk
      u32 hi = W ? (2u * n) / W : 0u, hj = W ? 2u * n - hi * W : 0u;
      u32 ebase = 0, left = 0, nchunk = 0;
      u32 pbase = W ? hi : n;
      R pre[NPF];
      fetch_g<G, NPF>(pre, E, Lx, pbase, p);
      bool have = true;
      // the phase's loads land here: no wait for them inside the step loops,
      // where the in-order load counter would also hold back the prefetched
      // samples of the next chunk
      __builtin_amdgcn_s_waitcnt(0);
      if (recompute) {
        if (n == s0) {
#pragma unroll
          for (int r = 0; r < NF; ++r) {
            if (S::segr(r) >= 0) seg[S::segr(r) * WAVE + lane] = q[r];
            if (mirror) sb[r * st] = q[r];
          }
        } else if (!seg_here) {
          // resumed inside a recompute: the LDS copy of the steps recomputed
          // by earlier launches comes back from HBM
          for (u32 m = 0; m <= n - s0; ++m)
#pragma unroll
            for (int r = 0; r < NF; ++r)
              if (S::segr(r) >= 0) seg[(m * NSEG + S::segr(r)) * WAVE + lane] = sb[((u64)m * NF + r) * st];
          if constexpr (EX)
            for (u32 m = 0; m < 3u * (n - s0); ++m)
#pragma unroll
              for (int r = 0; r < NF; ++r)
                if (S::segr(r) >= 0)
                  seg3[(m * NSEG + S::segr(r)) * WAVE + lane] = sb3[((u64)m * NF + r) * st];
        }
        seg_here = true;
      }
      const u32 e = recompute ? umin(T, s0 + K) : T;
      const u32 nend = (e - n < budget) ? e : n + budget;
      budget -= nend - n;
      const u32 nfirst = n;
      // per-step bookkeeping only as far as the phase needs it
      const u32 mode = recompute ? mRec : ((SE != 0u || a.window != 0ull) ? mFull : (adjoint ? mCk : mPlain));
      u32 next_ck = ((n + K - 1u) / K) * K;
      u32 next_save = SE ? ((n + SE - 1u) / SE) * SE : 0u;
      auto steps_to = [&](auto all_t, auto mode_t, auto zoh_t, u32 stop) {
        constexpr bool ALL = decltype(all_t)::value;
        constexpr u32 MODE = decltype(mode_t)::value;
        constexpr bool ZOH = decltype(zoh_t)::value;
        for (; n < stop; ++n) {
          R u0, u1, u3;
          if (ZOH) {
            // 0 < n < L or 0: sample 0 and those past the end are staged as 0
            u0 = eb[(n - ebase) * IPW];
            u1 = u0;
            u3 = eb[(n + 1u - ebase) * IPW];
          } else {
            if constexpr (FT > 0u) {
              const u32 m = 3u * (n - nchunk);
              u0 = ft[m * IPW];
              u1 = ft[(m + 1u) * IPW];
              u3 = ft[(m + 2u) * IPW];
            } else {
              u0 = interp_at<IPW>(eb, ebase, wts, W, hi, hj, 0u);
              u1 = interp_at<IPW>(eb, ebase, wts, W, hi, hj, 1u);
              u3 = interp_at<IPW>(eb, ebase, wts, W, hi, hj, 2u);
            }
            hj += 2u;
            if (hj >= W) {
              hj -= W;
              hi += 1u;
            }
          }
          if ((MODE == mCk || MODE == mFull) && adjoint && !recompute && n == next_ck) {
#pragma unroll
            for (int r = 0; r < NF; ++r) ck[((u64)(n / K) * NF + r) * WAVE] = q[r];
            next_ck += K;
          }
          if constexpr (EX && MODE == mRec) {
            const u32 mb = 3u * (n - s0);
            rk4rec<NF, ALL, S>(q, [&](R* k, const R* y, int s) {
              S::rhs_f(k, y, c, s == 0 ? u0 : (s == 3 ? u3 : u1));
            }, [&](int s, const R* y) {
              const u32 m = mb + (u32)(s - 1);
#pragma unroll
              for (int r = 0; r < NF; ++r) {
                if (S::segr(r) >= 0) seg3[(m * NSEG + S::segr(r)) * WAVE + lane] = y[r];
                if (mirror) sb3[((u64)m * NF + r) * st] = y[r];
              }
            }, h, h2, h6, n < fsteps);
          } else {
            rk4<NF, ALL, S>(q, [&](R* k, const R* y, int s) {
              S::rhs_f(k, y, c, s == 0 ? u0 : (s == 3 ? u3 : u1));
            }, h, h2, h6, n < fsteps);
          }
          if constexpr (MODE == mRec) {
            const u32 m = n + 1u - s0;
#pragma unroll
            for (int r = 0; r < NF; ++r) {
              if (S::segr(r) >= 0) seg[(m * NSEG + S::segr(r)) * WAVE + lane] = q[r];
              if (mirror) sb[((u64)m * NF + r) * st] = q[r];
            }
          }
          if constexpr (MODE == mFull) {
            {
              if (SE && n == next_save) {
                const u32 slot = n / SE;
                if (slot < (u32)a.nslots && valid)
#pragma unroll
                  for (int r = 0; r < NF; ++r)
                    if (orow_f(r) >= 0) traj[(u64)slot * a.out_len + fout + orow_f(r)] = q[r];
                next_save += SE;
              }
              if (a.window && (ALL || n < fsteps) && n + 1u + (u32)a.window > fsteps)
#pragma unroll
                for (int r = 0; r < NF; ++r) mx[r] = fmaxR(mx[r], absR(q[r]));
              if (n + 1u == fsteps && valid) {
#pragma unroll
                for (int r = 0; r < NF; ++r) {
                  if (orow_f(r) < 0) continue;
                  outq[fout + orow_f(r)] = q[r];
                  if (a.window) outm[fout + orow_f(r)] = mx[r];
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
            put_g<G, NPF>(eb, pre, Lx, base, forced, p, W == 0u);
          } else {
            R v[NPF];
            fetch_g<G, NPF>(v, E, Lx, base, p);
            put_g<G, NPF>(eb, v, Lx, base, forced, p, W == 0u);
          }
          ebase = base;
          nchunk = n;
          if constexpr (FT > 0u) {
            if (W != 0u) {
              // value k of the table: step nchunk + k / 3, half step
              // 2 (nchunk + k / 3) + (0, 1, 2)[k % 3] = hi W + hj + ...
#pragma unroll
              for (u32 v = 0; v < NFT; ++v) {
                const u32 k = p + v * G;
                if (k < FT) {
                  const u32 m = k / 3u;
                  const u32 hs = hj + 2u * m + (k - 3u * m);
                  const u32 di = hs / W;
                  ft[k * IPW] = interp_at<IPW>(eb, ebase, wts, W, hi + di, hs - di * W, 0u);
                }
              }
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
              __builtin_amdgcn_wave_barrier();
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
            }
          }
          left = CHUNK;
          pbase = W ? (2u * (n + CHUNK)) / W : n + CHUNK;
          fetch_g<G, NPF>(pre, E, Lx, pbase, p);
          have = true;
        }
        u32 stop = umin(nend, n + left);
        if (n < Tmin && Tmin < stop) stop = Tmin;
        left -= stop - n;
        const bool all = n < Tmin;
        auto in_mode = [&](auto mode_t) {
#if FORCING == 0
          if (W == 0u) {
            if (all) steps_to(Flag<true>(), mode_t, Flag<true>(), stop);
            else steps_to(Flag<false>(), mode_t, Flag<true>(), stop);
          } else {
            if (all) steps_to(Flag<true>(), mode_t, Flag<false>(), stop);
            else steps_to(Flag<false>(), mode_t, Flag<false>(), stop);
          }
#elif FORCING == 2
          if (all) steps_to(Flag<true>(), mode_t, Flag<true>(), stop);
          else steps_to(Flag<false>(), mode_t, Flag<true>(), stop);
#else
          if (all) steps_to(Flag<true>(), mode_t, Flag<false>(), stop);
          else steps_to(Flag<false>(), mode_t, Flag<false>(), stop);
#endif
        };
        if (mode == mPlain) in_mode(Mode<mPlain>());
        else if (mode == mCk) in_mode(Mode<mCk>());
#if FULL
        else if (mode == mRec) in_mode(Mode<mRec>());
        else in_mode(Mode<mFull>());
#else
        else in_mode(Mode<mRec>());
#endif
      }
      // without per-step outputs: an instance whose last step fell in this
      // range kept its final state since (inactive steps leave it)
      if ((mode == mPlain || mode == mCk) && valid && fsteps > nfirst && fsteps <= nend)
#pragma unroll
        for (int r = 0; r < NF; ++r)
          if (orow_f(r) >= 0) outq[fout + orow_f(r)] = q[r];
      if (n == e && !recompute && SE && valid) {
        // trajectory slots after the wave's last step hold the final state
        for (u32 slot = (T + SE - 1u) / SE; slot < (u32)a.nslots; ++slot)
#pragma unroll
          for (int r = 0; r < NF; ++r)
            if (orow_f(r) >= 0) traj[(u64)slot * a.out_len + fout + orow_f(r)] = q[r];
      }
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
            for (int r = 0; r < S::NA; ++r)
              if (orow_a(r) >= 0 && (u32)orow_a(r) < nao) outg[aout + orow_a(r)] = li[r * st];
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
          for (int r = 0; r < NF; ++r)
            if (S::segr(r) >= 0) seg[(m * NSEG + S::segr(r)) * WAVE + lane] = sb[((u64)m * NF + r) * st];
        if constexpr (EX)
          for (u32 m = 0; m < 3u * (n - s0); ++m)
#pragma unroll
            for (int r = 0; r < NF; ++r)
              if (S::segr(r) >= 0)
                seg3[(m * NSEG + S::segr(r)) * WAVE + lane] = sb3[((u64)m * NF + r) * st];
        seg_here = true;
      }
      __builtin_amdgcn_s_waitcnt(0);
      const u32 nend = (n - s0 < budget) ? s0 : n - budget;
      budget -= n - nend;
      auto back_to = [&](auto all_t, u32 stop) {
        constexpr bool ALL = decltype(all_t)::value;
        for (; n > stop; --n) {
          if constexpr (EX) {
            // exact discrete adjoint of step n-1 -> n from its stage inputs
            // y1 (the post-step state of step n-1) .. y4:
            //   a4 = h6 lam,            mu4 = g(a4, y4)
            //   a3 = 2 h6 lam + h mu4,  mu3 = g(a3, y3)
            //   a2 = 2 h6 lam + h2 mu3, mu2 = g(a2, y2)
            //   a1 = h6 lam + h2 mu2,   mu1 = g(a1, y1)
            //   lam += ((mu4 + mu3) + mu2) + mu1
            R y1[NX], y2[NX], y3[NX], y4[NX];
            const u32 mb = 3u * (n - 1u - s0);
            S::xload(y4, y1, seg3 + (mb + 2u) * (NSEG * WAVE), seg + (n - 1u - s0) * (NSEG * WAVE), lane);
            S::xload(y3, y2, seg3 + (mb + 1u) * (NSEG * WAVE), seg3 + mb * (NSEG * WAVE), lane);
            const R h3 = R(2) * h6;
            R e[S::NE > 0 ? S::NE : 1];
            R av[NA], mu[NA], ac[NA];
#pragma unroll
            for (int r = 0; r < NA; ++r) av[r] = h6 * lam[r];
            S::eff_a(e, y4, c);
            S::lin_a(mu, av, e);
#pragma unroll
            for (int r = 0; r < NA; ++r) {
              ac[r] = mu[r];
              av[r] = fmaR(h, mu[r], h3 * lam[r]);
            }
            S::eff_a(e, y3, c);
            S::lin_a(mu, av, e);
#pragma unroll
            for (int r = 0; r < NA; ++r) {
              ac[r] = ac[r] + mu[r];
              av[r] = fmaR(h2, mu[r], h3 * lam[r]);
            }
            S::eff_a(e, y2, c);
            S::lin_a(mu, av, e);
#pragma unroll
            for (int r = 0; r < NA; ++r) {
              ac[r] = ac[r] + mu[r];
              av[r] = fmaR(h2, mu[r], h6 * lam[r]);
            }
            S::eff_a(e, y1, c);
            S::lin_a(mu, av, e);
            const bool act = ALL || n <= fsteps;
#pragma unroll
            for (int r = 0; r < NA; ++r) {
              ac[r] = ac[r] + mu[r];
              const R ln = lam[r] + ac[r];
              lam[r] = act ? ln : lam[r];
            }
            continue;
          }
          R xc[NX], xb[NX], xp[NX];
          S::xload(xc, xp, seg + (n - s0) * (NSEG * WAVE), seg + (n - 1u - s0) * (NSEG * WAVE), lane);
#pragma unroll
          for (int j = 0; j < NX; ++j) xb[j] = R(0.5) * (xc[j] + xp[j]);
          // effective coefficients per set point: x_n, the midpoint (stages
          // 2 and 3), x_{n-1}
          R e[S::NE > 0 ? S::NE : 1];
          rk4<NA, ALL, NoSink>(lam, [&](R* k, const R* y, int s) {
            if (s == 0) S::eff_a(e, xc, c);
            else if (s == 1) S::eff_a(e, xb, c);
            else if (s == 3) S::eff_a(e, xp, c);
            S::lin_a(k, y, e);
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
            for (int r = 0; r < S::NA; ++r)
              if (orow_a(r) >= 0 && (u32)orow_a(r) < nao) outg[aout + orow_a(r)] = lam[r];
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
// ---------------------------------------------------------------- multi-wave
// LDS-only workgroup barrier: global memory stays unordered, so loads in
// flight are not waited for.
__device__ __forceinline__ void bar() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
#define YSTR (NYMAX * LW)
// adjoint set point j at stage s: x_n, xbar = (x_n + x_{n-1}) / 2, x_{n-1}
__device__ __forceinline__ R xs(const R* Xc, const R* Xp, int j, int s) {
  return s == 0 ? Xc[j * LW] : (s == 3 ? Xp[j * LW] : R(0.5) * (Xc[j * LW] + Xp[j * LW]));
}

// One RK4 step of the NR rows a wave owns.  Stage values of every row go
// through the LDS buffers Y[2][NYMAX][LW]: stage s reads buffer s % 2 and
// writes buffer (s + 1) % 2; stage 4 writes the new state into buffer 0,
// which the next step's first stage reads.  One barrier per stage.
template <int NR, bool ALL, class F, class Row>
__device__ __forceinline__ void rk4_mw(R* q, R* Y, F rhs, Row row, R h, R h2, R h6, bool active) {
  R k[NR > 0 ? NR : 1], acc[NR > 0 ? NR : 1];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const R* Yc = Y + (s & 1) * YSTR;
    R* Yn = Y + ((s + 1) & 1) * YSTR;
    rhs(k, Yc, s);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      R v;
      if (s == 0) {
        acc[r] = k[r];
        v = fmaR(h2, k[r], q[r]);
      } else if (s == 1) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        v = fmaR(h2, k[r], q[r]);
      } else if (s == 2) {
        acc[r] = fmaR(R(2), k[r], acc[r]);
        v = fmaR(h, k[r], q[r]);
      } else {
        acc[r] = acc[r] + k[r];
        const R qn = fmaR(h6, acc[r], q[r]);
        q[r] = (ALL || active) ? qn : q[r];
        v = q[r];
      }
      if (row(r) >= 0) Yn[row(r) * LW] = v;
    }
    bar();
  }
}

// One wave of a multi-wave workgroup: the rows SW owns of LW instances, same
// phases, budget and resumable state as run().  Every wave of the workgroup
// follows the same control flow (it depends only on the workgroup's step
// counts and progress), so all reach the same barriers.  Lanes >= LW mirror
// lane % LW: they compute the same values and store them to the same
// addresses.  LDS: Y [2][NYMAX][LW], seg [SEGK+1][NFMAX][LW] (adjoint), and a
// [STAGE][LW] sample region per wave that drives forced rows.
template <class S, class SW>
__device__ __forceinline__ void run_mw(const LaneArgs& a, u32 wg, R* smem) {
  constexpr int NF = S::NF;
  constexpr int RF = SW::RF > 0 ? SW::RF : 1;
  constexpr int RA = SW::RA > 0 ? SW::RA : 1;
  constexpr int TFW = SW::TF > 0 ? SW::TF : 1;
  constexpr int TAW = SW::TA > 0 ? SW::TA : 1;
  constexpr bool DRIVES = SW::FSLOT >= 0;
  const u32 le = threadIdx.x & (LW - 1u);
  const u64 st = a.stride;
  const u32 i = wg * LW + le;
  const bool valid = i < (u32)a.count;
  R* Y = smem + le;
  R* seg = smem + 2 * YSTR + le;
  R* eb = smem + 2 * YSTR + SEGN + (DRIVES ? SW::FSLOT : 0) * (STAGE * LW) + le;
  const bool adjoint = S::ADJ && a.adjoint != 0ull;
  const u32 K = adjoint ? (u32)a.ck_every : 1u;
  const u32 W = (u32)a.interp_w;
  const u32 Lx = (u32)a.exc_len;
  const u32 SE = (u32)a.save_every;
  const R h = R(a.h), h2 = R(a.h2), h6 = R(a.h6);
  GP(const u32) steps = (GP(const u32))a.steps;
  const u32 T = uni(steps[wg * LW]);
  const u32 Tmin = uni(steps[umin(wg * LW + LW, (u32)a.count) - 1u]);
  const u32 fsteps = steps[i];
  const i32 exc = ((GP(const i32))a.exc)[i];
  const bool forced = exc >= 0;
  const u32 fout = ((GP(const u32))a.out_f)[i];
  const u32 aout = ((GP(const u32))a.out_a)[i];
  const u32 nao = ((GP(const u32))a.na_out)[i];
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
  GP(u32) pst = (GP(u32))a.progress + 4 * wg;
  u32 n = uni(pst[0]);
  u32 phase = uni(pst[1]);
  u32 sg = uni(pst[2]);
  u32 budget = (u32)a.budget;
  auto slotf = [](int r) { return SW::yf(r); };
  auto slota = [](int r) { return SW::ya(r); };

  R q[RF], mx[RF], lam[RA];
#pragma unroll
  for (int r = 0; r < RF; ++r) mx[r] = R(0);
#pragma unroll
  for (int r = 0; r < RA; ++r) lam[r] = R(0);
  if (n == 0u && phase == kForward) {
#pragma unroll
    for (int r = 0; r < SW::RF; ++r) q[r] = qi[SW::rf(r) * st];
    if (fsteps == 0u && valid)
#pragma unroll
      for (int r = 0; r < SW::RF; ++r) outq[fout + SW::rf(r)] = q[r];
  } else {
#pragma unroll
    for (int r = 0; r < SW::RF; ++r) {
      q[r] = qs[SW::rf(r) * st];
      if (a.window) mx[r] = ms[SW::rf(r) * st];
    }
  }
  if (adjoint && phase != kForward)
#pragma unroll
    for (int r = 0; r < SW::RA; ++r) lam[r] = ls[SW::ra(r) * st];
  bool seg_here = false;  // LDS holds the current segment up to step n
  // a launch that may stop mid-segment mirrors the segment to HBM so the
  // next launch can rebuild its LDS copy; one that runs to the end does not
  const bool mirror = a.budget != 0xFFFFFFFFull;

  while (budget > 0u && phase != kDone) {
    if (phase != kBackward) {
      const bool recompute = phase == kRecompute;
      R c[TFW];
      SW::load_f(c, cf, st);
      const u32 s0 = sg * K;
      if (recompute) {
        if (n == s0) {
#pragma unroll
          for (int r = 0; r < SW::RF; ++r) {
            q[r] = ck[((u64)sg * NF + SW::rf(r)) * st];
            seg[SW::rf(r) * LW] = q[r];
            if (mirror) sb[SW::rf(r) * st] = q[r];
          }
        } else if (!seg_here) {
          for (u32 m = 0; m <= n - s0; ++m)
#pragma unroll
            for (int r = 0; r < SW::RF; ++r)
              seg[(m * NFMAX + SW::rf(r)) * LW] = sb[((u64)m * NF + SW::rf(r)) * st];
        }
        seg_here = true;
      }
      const u32 e = recompute ? umin(T, s0 + K) : T;
      const u32 nend = (e - n < budget) ? e : n + budget;
      budget -= nend - n;
      const u32 nfirst = n;
      const bool gen = recompute || adjoint || SE != 0u || a.window != 0ull;
      u32 next_ck = ((n + K - 1u) / K) * K;
      u32 next_save = SE ? ((n + SE - 1u) / SE) * SE : 0u;
      u32 hi = W ? (2u * n) / W : 0u, hj = W ? 2u * n - hi * W : 0u;
      u32 ebase = 0, left = 0, pbase = 0;
      bool have = false;
      R pre[DRIVES ? STAGE : 1];
      // every row's state is visible before the first stage reads it
#pragma unroll
      for (int r = 0; r < SW::RF; ++r)
        if (SW::yf(r) >= 0) Y[SW::yf(r) * LW] = q[r];
      bar();
      auto steps_to = [&](auto all_t, auto gen_t, auto zoh_t, u32 stop) {
        constexpr bool ALL = decltype(all_t)::value;
        constexpr bool GEN = decltype(gen_t)::value;
        constexpr bool ZOH = decltype(zoh_t)::value;
        for (; n < stop; ++n) {
          R u0 = R(0), u1 = R(0), u3 = R(0);
          if constexpr (DRIVES) {
            if (ZOH) {
              const R e0 = eb[(n - ebase) * LW], e1 = eb[(n + 1u - ebase) * LW];
              u0 = (n > 0u && n < Lx) ? e0 : R(0);
              u1 = u0;
              u3 = e1;
            } else {
              u0 = interp_at<LW>(eb, ebase, wts, W, hi, hj, 0u);
              u1 = interp_at<LW>(eb, ebase, wts, W, hi, hj, 1u);
              u3 = interp_at<LW>(eb, ebase, wts, W, hi, hj, 2u);
            }
          }
          if (!ZOH) {
            hj += 2u;
            if (hj >= W) {
              hj -= W;
              hi += 1u;
            }
          }
          if (GEN && adjoint && !recompute && n == next_ck) {
#pragma unroll
            for (int r = 0; r < SW::RF; ++r) ck[((u64)(n / K) * NF + SW::rf(r)) * st] = q[r];
            next_ck += K;
          }
          rk4_mw<SW::RF, ALL>(q, Y, [&](R* k, const R* Yc, int s) {
            SW::rhs_f(k, Yc, c, s == 0 ? u0 : (s == 3 ? u3 : u1));
          }, slotf, h, h2, h6, n < fsteps);
          if (GEN) {
            if (recompute) {
              const u32 m = n + 1u - s0;
#pragma unroll
              for (int r = 0; r < SW::RF; ++r) {
                seg[(m * NFMAX + SW::rf(r)) * LW] = q[r];
                if (mirror) sb[((u64)m * NF + SW::rf(r)) * st] = q[r];
              }
            } else {
              if (SE && n == next_save) {
                const u32 slot = n / SE;
                if (slot < (u32)a.nslots && valid)
#pragma unroll
                  for (int r = 0; r < SW::RF; ++r)
                    traj[(u64)slot * a.out_len + fout + SW::rf(r)] = q[r];
                next_save += SE;
              }
              if (a.window && (ALL || n < fsteps) && n + 1u + (u32)a.window > fsteps)
#pragma unroll
                for (int r = 0; r < SW::RF; ++r) mx[r] = fmaxR(mx[r], absR(q[r]));
              if (n + 1u == fsteps && valid) {
#pragma unroll
                for (int r = 0; r < SW::RF; ++r) {
                  outq[fout + SW::rf(r)] = q[r];
                  if (a.window) outm[fout + SW::rf(r)] = mx[r];
                }
              }
            }
          }
        }
      };
      while (n < nend) {
        if (left == 0u) {
          if constexpr (DRIVES) {
            const u32 base = W ? hi : n;
            if (have && pbase == base) {
              put<LW>(eb, pre, Lx, base, forced);
            } else {
              R v[STAGE];
              fetch(v, E, Lx, base);
              put<LW>(eb, v, Lx, base, forced);
            }
            ebase = base;
            pbase = W ? (2u * (n + CHUNK)) / W : n + CHUNK;
            fetch(pre, E, Lx, pbase);
            have = true;
          }
          left = CHUNK;
        }
        u32 stop = umin(nend, n + left);
        if (n < Tmin && Tmin < stop) stop = Tmin;
        left -= stop - n;
        const bool all = n < Tmin;
        // one step is long here: per-step bookkeeping checks cost little
#if FORCING == 0
        if (W == 0u) {
          if (all) steps_to(Flag<true>(), Flag<true>(), Flag<true>(), stop);
          else steps_to(Flag<false>(), Flag<true>(), Flag<true>(), stop);
        } else {
          if (all) steps_to(Flag<true>(), Flag<true>(), Flag<false>(), stop);
          else steps_to(Flag<false>(), Flag<true>(), Flag<false>(), stop);
        }
#elif FORCING == 2
        if (all) steps_to(Flag<true>(), Flag<true>(), Flag<true>(), stop);
        else steps_to(Flag<false>(), Flag<true>(), Flag<true>(), stop);
#else
        if (all) steps_to(Flag<true>(), Flag<true>(), Flag<false>(), stop);
        else steps_to(Flag<false>(), Flag<true>(), Flag<false>(), stop);
#endif
      }
      if (!gen && valid && fsteps > nfirst && fsteps <= nend)
#pragma unroll
        for (int r = 0; r < SW::RF; ++r) outq[fout + SW::rf(r)] = q[r];
      if (n == e && !recompute && SE && valid) {
        // trajectory slots after the workgroup's last step hold the final state
        for (u32 slot = (T + SE - 1u) / SE; slot < (u32)a.nslots; ++slot)
#pragma unroll
          for (int r = 0; r < SW::RF; ++r)
            traj[(u64)slot * a.out_len + fout + SW::rf(r)] = q[r];
      }
      if (n == e) {
        if (recompute) {
          phase = kBackward;
        } else if (adjoint && T > 0u) {
          sg = (T - 1u) / K;
          n = sg * K;
          phase = kRecompute;
#pragma unroll
          for (int r = 0; r < SW::RA; ++r) lam[r] = li[SW::ra(r) * st];
        } else {
          phase = kDone;
          if (adjoint && valid)
#pragma unroll
            for (int r = 0; r < SW::RA; ++r)
              if ((u32)SW::ra(r) < nao) outg[aout + SW::ra(r)] = li[SW::ra(r) * st];
        }
      }
      bar();
    } else if (S::ADJ) {
      R c[TAW];
      SW::load_a(c, ca, st);
      const u32 s0 = sg * K;
      if (!seg_here) {
        for (u32 m = 0; m <= n - s0; ++m)
#pragma unroll
          for (int r = 0; r < SW::RF; ++r)
            seg[(m * NFMAX + SW::rf(r)) * LW] = sb[((u64)m * NF + SW::rf(r)) * st];
        seg_here = true;
      }
#pragma unroll
      for (int r = 0; r < SW::RA; ++r)
        if (SW::ya(r) >= 0) Y[SW::ya(r) * LW] = lam[r];
      // every wave's segment rows and adjoint state are visible
      bar();
      const u32 nend = (n - s0 < budget) ? s0 : n - budget;
      budget -= n - nend;
      auto back_to = [&](auto all_t, u32 stop) {
        constexpr bool ALL = decltype(all_t)::value;
        for (; n > stop; --n) {
          const R* Xc = seg + (n - s0) * (NFMAX * LW);
          const R* Xp = seg + (n - 1u - s0) * (NFMAX * LW);
          rk4_mw<SW::RA, ALL>(lam, Y, [&](R* k, const R* Yc, int s) {
            SW::rhs_a(k, Yc, Xc, Xp, c, s);
          }, slota, h, h2, h6, n <= fsteps);
        }
      };
      if (n > Tmin) back_to(Flag<false>(), Tmin > nend ? Tmin : nend);
      if (n > nend) back_to(Flag<true>(), nend);
      if (n == s0) {
        if (sg == 0u) {
          phase = kDone;
          if (valid)
#pragma unroll
            for (int r = 0; r < SW::RA; ++r)
              if ((u32)SW::ra(r) < nao) outg[aout + SW::ra(r)] = lam[r];
        } else {
          sg -= 1u;
          n = sg * K;
          phase = kRecompute;
          seg_here = false;
        }
      }
      bar();
    } else {
      phase = kDone;
    }
  }
#pragma unroll
  for (int r = 0; r < SW::RF; ++r) {
    qs[SW::rf(r) * st] = q[r];
    if (a.window) ms[SW::rf(r) * st] = mx[r];
  }
  if (adjoint)
#pragma unroll
    for (int r = 0; r < SW::RA; ++r) ls[SW::ra(r) * st] = lam[r];
  if (threadIdx.x == 0u) {
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

using Namer = std::function<std::string(uint32_t)>;

// Global coefficient slots of a side: terms row by row, then the force of
// every forced row (the order the host packs coefficients in).
struct Slots {
  std::vector<uint32_t> base;   // first term slot of each row
  std::vector<int> force;       // force slot of each row, -1 if unforced
};
Slots slots_of(const LaneSide& sd) {
  Slots sl;
  sl.base.resize(sd.n);
  sl.force.assign(sd.n, -1);
  uint32_t t = 0;
  for (uint32_t r = 0; r < sd.n; ++r) {
    sl.base[r] = t;
    t += uint32_t(sd.rows[r].size());
  }
  for (size_t i = 0; i < sd.forced.size(); ++i) sl.force[sd.forced[i]] = int(t + i);
  return sl;
}

// Statements computing k[kidx] for row r: each kind summed in term order
// (first term a product, then fused multiply-adds), the kinds added in
// order, then the force.  coef(g) names global coefficient slot g; S(j) and
// X(j) name factor j of the side's own state and of the set point.
void emit_row(std::ostringstream& o, const LaneSide& sd, const Slots& sl, uint32_t r,
              uint32_t kidx, bool forced_input, const Namer& coef, const Namer& S,
              const Namer& X) {
  std::string sums[4];
  uint32_t g = sl.base[r];
  for (int kind = 0; kind < 4; ++kind) {
    std::string acc;
    uint32_t gt = sl.base[r];
    for (const LaneTerm& tm : sd.rows[r]) {
      const uint32_t slot = gt++;
      if (int(tm.kind) != kind) continue;
      std::string prod;
      if (kind == 0) {
        prod = acc.empty() ? coef(slot) + " * " + S(tm.f[0])
                           : "fmaR(" + coef(slot) + ", " + S(tm.f[0]) + ", " + acc + ")";
      } else {
        std::string l = coef(slot) + " * " + S(tm.f[0]);
        for (int f = 1; f < int(kind); ++f) l = "(" + l + ") * " + X(tm.f[f]);
        const std::string last = X(tm.f[kind]);
        prod = acc.empty() ? "(" + l + ") * " + last : "fmaR(" + l + ", " + last + ", " + acc + ")";
      }
      const std::string name = "s" + std::to_string(kind) + "_" + std::to_string(r);
      if (acc.empty())
        o << "    R " << name << " = " << prod << ";\n";
      else
        o << "    " << name << " = " << prod << ";\n";
      acc = name;
    }
    sums[kind] = acc;
  }
  (void)g;
  std::string k;
  for (int kind = 0; kind < 4; ++kind) {
    if (sums[kind].empty()) continue;
    k = k.empty() ? sums[kind] : "(" + k + " + " + sums[kind] + ")";
  }
  if (k.empty()) k = "R(0)";
  if (forced_input && sl.force[r] >= 0) k = "fmaR(" + coef(uint32_t(sl.force[r])) + ", u, " + k + ")";
  o << "    k[" << kidx << "] = " << k << ";\n";
}

// Products of state factors emitted once per right-hand side evaluation
// (`m<k>` temporaries, keyed by the factor names in order), so terms sharing a
// monomial -- x0^2 in a Gamma and a Lambda term -- share its multiplies.
struct Monos {
  std::map<std::string, std::string> names;
  std::string get(std::ostringstream& o, const std::vector<std::string>& fs, size_t n) {
    if (n == 1) return fs[0];
    std::string key;
    for (size_t i = 0; i < n; ++i) key += fs[i] + ",";
    auto it = names.find(key);
    if (it != names.end()) return it->second;
    const std::string left = get(o, fs, n - 1);
    const std::string name = "m" + std::to_string(names.size());
    o << "    const R " << name << " = " << left << " * " << fs[n - 1] << ";\n";
    names.emplace(key, name);
    return name;
  }
};

// Single-wave right-hand side: every row, coefficients c[global slot].
void emit_rhs(std::ostringstream& o, const LaneSide& sd, bool forced_input, const char* S,
              const char* X) {
  const Slots sl = slots_of(sd);
  const std::string sn(S), xn(X);
  for (uint32_t r = 0; r < sd.n; ++r)
    emit_row(o, sd, sl, r, r, forced_input,
             [](uint32_t gs) { return "c[" + std::to_string(gs) + "]"; },
             [&](uint32_t j) { return sn + "[" + std::to_string(j) + "]"; },
             [&](uint32_t j) { return xn + "[" + std::to_string(j) + "]"; });
}

// Single-wave forward right-hand side with shared monomials: each kind
// summed in term order (coefficient times the cached product of its state
// factors), kinds added in order, then the force.
void emit_rhs_forward(std::ostringstream& o, const LaneSide& sd) {
  const Slots sl = slots_of(sd);
  Monos monos;
  auto y = [](uint32_t j) { return "y[" + std::to_string(j) + "]"; };
  for (uint32_t r = 0; r < sd.n; ++r) {
    std::string k;
    for (int kind = 0; kind < 4; ++kind) {
      std::string acc;
      uint32_t slot = sl.base[r];
      for (const LaneTerm& tm : sd.rows[r]) {
        const std::string coef = "c[" + std::to_string(slot++) + "]";
        if (int(tm.kind) != kind) continue;
        std::vector<std::string> fs;
        for (uint32_t a = 0; a <= tm.kind; ++a) fs.push_back(y(tm.f[a]));
        const std::string m = monos.get(o, fs, fs.size());
        const std::string prod = acc.empty() ? coef + " * " + m : "fmaR(" + coef + ", " + m + ", " + acc + ")";
        const std::string name = "s" + std::to_string(kind) + "_" + std::to_string(r);
        o << "    " << (acc.empty() ? "R " : "") << name << " = " << prod << ";\n";
        acc = name;
      }
      if (!acc.empty()) k = k.empty() ? acc : "(" + k + " + " + acc + ")";
    }
    if (k.empty()) k = "R(0)";
    if (sl.force[r] >= 0) k = "fmaR(c[" + std::to_string(sl.force[r]) + "], u, " + k + ")";
    o << "    k[" << r << "] = " << k << ";\n";
  }
}

// One wave's rows of a side in a multi-wave class: its own coefficient
// array (local slots, loaded from the global slots listed in *gslots).
void emit_wave_side(std::ostringstream& o, const LaneSide& sd, const std::vector<uint32_t>& rows,
                    bool forced_input, bool adjoint_side, std::vector<uint32_t>* gslots) {
  const Slots sl = slots_of(sd);
  std::map<uint32_t, uint32_t> local;
  auto coef = [&](uint32_t gs) {
    auto it = local.find(gs);
    if (it == local.end()) {
      it = local.emplace(gs, uint32_t(gslots->size())).first;
      gslots->push_back(gs);
    }
    return "c[" + std::to_string(it->second) + "]";
  };
  const std::vector<int> ys = read_slots(sd, !adjoint_side);
  auto S = [&](uint32_t j) { return "Y[" + std::to_string(ys[j]) + " * LW]"; };
  auto X = [&](uint32_t j) {
    return adjoint_side ? "xs(Xc, Xp, " + std::to_string(j) + ", s)"
                        : "Y[" + std::to_string(ys[j]) + " * LW]";
  };
  for (size_t lr = 0; lr < rows.size(); ++lr) {
    emit_row(o, sd, sl, rows[lr], uint32_t(lr), forced_input, coef, S, X);
    // factors are re-read from LDS per row rather than all kept in registers
    o << "    asm volatile(\"\" ::: \"memory\");\n";
  }
}

// The adjoint right-hand side is linear in the adjoint state: every row is
// sum_o e_o(x) y_o over its own factors o, with e_o(x) the sum of its
// terms' coefficients times their set-point factors.  eff_a evaluates the
// e_o once per set point (x_n, the midpoint -- shared by the two middle
// stages -- and x_{n-1}); lin_a combines them with the stage's values.
struct ETerm {
  uint32_t kind;
  std::string own;
  std::vector<std::string> xs;
  uint32_t coef;
};
struct ERow {
  std::vector<ETerm> terms;
  bool shared = false;
};
void emit_adjoint_linear(std::ostringstream& oe, std::ostringstream& ol, const std::vector<ERow>& rows,
                         uint32_t G, uint32_t* ne) {
  uint32_t j = 0;
  for (size_t r = 0; r < rows.size(); ++r) {
    std::vector<std::string> owns;
    std::vector<std::string> effs;
    for (const ETerm& t : rows[r].terms) {
      size_t o = 0;
      while (o < owns.size() && owns[o] != t.own) ++o;
      if (o == owns.size()) {
        owns.push_back(t.own);
        effs.push_back("");
      }
      const std::string c = "c[" + std::to_string(t.coef) + "]";
      std::string& e = effs[o];
      if (t.kind == 0) {
        e = e.empty() ? c : "(" + e + " + " + c + ")";
      } else {
        std::string l = c;
        for (uint32_t f = 0; f + 1 < t.kind; ++f) l = "(" + l + ") * " + t.xs[f];
        const std::string last = t.xs[t.kind - 1];
        e = e.empty() ? "(" + l + ") * " + last : "fmaR(" + l + ", " + last + ", " + e + ")";
      }
    }
    std::string k;
    for (size_t o = 0; o < owns.size(); ++o) {
      oe << "    e[" << j << "] = " << effs[o] << ";\n";
      const std::string ej = "e[" + std::to_string(j) + "]";
      k = k.empty() ? ej + " * " + owns[o] : "fmaR(" + ej + ", " + owns[o] + ", " + k + ")";
      ++j;
    }
    if (k.empty()) {
      k = "R(0)";
    } else if (rows[r].shared && G > 1) {
      k = "gsum<" + std::to_string(G) + "u>(" + k + ")";
    }
    ol << "    k[" << r << "] = " << k << ";\n";
  }
  *ne = j;
}

// Adjoint rows of a lane-group side for emit_adjoint_linear.
std::vector<ERow> group_adjoint_rows(const GSide& gs, const std::vector<GSrc>& gx) {
  std::vector<ERow> rows(gs.nl());
  uint32_t t = 0;
  for (uint32_t r = 0; r < gs.nl(); ++r) {
    const GRow& row = gs.rows[r];
    rows[r].shared = r >= gs.RL;
    for (size_t e = 0; e < row.terms.size(); ++e) {
      const GTerm& tm = row.terms[e];
      ETerm et;
      et.kind = tm.kind;
      et.coef = t + uint32_t(e);
      const GSrc& o = tm.f[0];
      et.own = (o.shared || o.d == 0) ? "y[" + std::to_string(o.slot) + "]"
                                      : "y" + std::to_string(o.slot) + "_" + std::to_string(o.d);
      for (uint32_t a = 1; a <= tm.kind; ++a) {
        const size_t jx = size_t(std::lower_bound(gx.begin(), gx.end(), tm.f[a]) - gx.begin());
        et.xs.push_back("x[" + std::to_string(jx) + "]");
      }
      rows[r].terms.push_back(et);
    }
    t += uint32_t(row.terms.size()) + (row.force.empty() ? 0u : 1u);
  }
  return rows;
}

// Adjoint rows of a one-lane side (coefficients in the side's slot order).
std::vector<ERow> lane_adjoint_rows(const LaneSide& sd) {
  std::vector<ERow> rows(sd.n);
  uint32_t t = 0;
  for (uint32_t r = 0; r < sd.n; ++r) {
    for (const LaneTerm& tm : sd.rows[r]) {
      ETerm et;
      et.kind = tm.kind;
      et.coef = t++;
      et.own = "y[" + std::to_string(tm.f[0]) + "]";
      for (uint32_t a = 1; a <= tm.kind; ++a) et.xs.push_back("x[" + std::to_string(tm.f[a]) + "]");
      rows[r].terms.push_back(et);
    }
  }
  return rows;
}

// Right-hand side of one side of a lane-group class: the own-state values
// the program reads from other positions of the group (rotated once per
// stage), then every lane slot: band slots directly, shared slots as
// partial sums added across the group.  Coefficients c[t] in per-lane slot
// order (row by row, terms then force); set points x[j] index `gx`.
void emit_group_side(std::ostringstream& o, const GSide& gs, uint32_t G, bool forward,
                     const std::vector<GSrc>& gx) {
  std::vector<GSrc> rem;
  for (const GRow& row : gs.rows)
    for (const GTerm& t : row.terms)
      for (uint32_t a = 0; a <= (forward ? t.kind : 0u); ++a)
        if (!t.f[a].shared && t.f[a].d != 0) rem.push_back(t.f[a]);
  std::sort(rem.begin(), rem.end());
  rem.erase(std::unique(rem.begin(), rem.end()), rem.end());
  for (const GSrc& r : rem)
    o << "    const R y" << r.slot << "_" << r.d << " = rot<" << G << "u, " << r.d << "u>(y["
      << r.slot << "]);\n";
  auto S = [&](const GSrc& r) {
    return (r.shared || r.d == 0) ? "y[" + std::to_string(r.slot) + "]"
                                  : "y" + std::to_string(r.slot) + "_" + std::to_string(r.d);
  };
  auto X = [&](const GSrc& r) {
    if (forward) return S(r);
    const size_t j = size_t(std::lower_bound(gx.begin(), gx.end(), r) - gx.begin());
    return "x[" + std::to_string(j) + "]";
  };
  Monos monos;
  uint32_t t = 0;
  for (uint32_t r = 0; r < gs.nl(); ++r) {
    const GRow& row = gs.rows[r];
    std::string k;
    for (int kind = 0; kind < 4; ++kind) {
      std::string acc;
      for (size_t e = 0; e < row.terms.size(); ++e) {
        const GTerm& tm = row.terms[e];
        if (int(tm.kind) != kind) continue;
        const std::string coef = "c[" + std::to_string(t + e) + "]";
        std::string prod;
        if (kind == 0) {
          prod = acc.empty() ? coef + " * " + S(tm.f[0])
                             : "fmaR(" + coef + ", " + S(tm.f[0]) + ", " + acc + ")";
        } else if (forward) {
          std::vector<std::string> fs{S(tm.f[0])};
          for (int f = 1; f <= kind; ++f) fs.push_back(X(tm.f[f]));
          const std::string m = monos.get(o, fs, fs.size());
          prod = acc.empty() ? coef + " * " + m : "fmaR(" + coef + ", " + m + ", " + acc + ")";
        } else {
          std::string l = coef + " * " + S(tm.f[0]);
          for (int f = 1; f < kind; ++f) l = "(" + l + ") * " + X(tm.f[f]);
          const std::string last = X(tm.f[kind]);
          prod = acc.empty() ? "(" + l + ") * " + last : "fmaR(" + l + ", " + last + ", " + acc + ")";
        }
        const std::string name = "s" + std::to_string(kind) + "_" + std::to_string(r);
        o << "    " << (acc.empty() ? "R " : "") << name << " = " << prod << ";\n";
        acc = name;
      }
      if (!acc.empty()) k = k.empty() ? acc : "(" + k + " + " + acc + ")";
    }
    const bool empty = k.empty() && (!forward || row.force.empty());
    const bool sink = r < gs.sink.size() && gs.sink[r];
    if (k.empty()) k = "R(0)";
    const uint32_t tf = t + uint32_t(row.terms.size());
    if (forward && !row.force.empty()) k = "fmaR(c[" + std::to_string(tf) + "], u, " + k + ")";
    // a sink's partial slope is summed across the group once per step (rk4)
    if (r >= gs.RL && G > 1 && !empty && !sink) k = "gsum<" + std::to_string(G) + "u>(" + k + ")";
    o << "    k[" << r << "] = " << k << ";\n";
    t = tf + (row.force.empty() ? 0u : 1u);
  }
}

std::string ternary_imap(const std::vector<int>& v) {
  std::string e = "-1";
  for (size_t i = v.size(); i-- > 0;)
    e = "r == " + std::to_string(i) + " ? " + std::to_string(v[i]) + " : (" + e + ")";
  return e;
}

std::string ternary_map(const std::vector<uint32_t>& v) {
  std::string e = "0";
  for (size_t i = v.size(); i-- > 0;)
    e = "r == " + std::to_string(i) + " ? " + std::to_string(v[i]) + " : (" + e + ")";
  return e;
}

std::string ternary_map_i(const std::vector<int>& v) {
  std::string e = "-1";
  for (size_t i = v.size(); i-- > 0;)
    e = "r == " + std::to_string(i) + " ? " + std::to_string(v[i]) + " : (" + e + ")";
  return e;
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

// Caller options come from PASIF_JIT_FLAGS (space separated).  hipRTC runs
// LLVM inside this process, and LLVM ends the process on option errors: an
// unknown `-mllvm` option exits, an option given twice in one compilation is
// refused, and the iterative-ilp machine scheduler crashes (SIGSEGV) on the
// one-lane two-DOF forward program.  So only options compiled cleanly on the
// generated programs are accepted, each replacing the library's own option of
// its kind, and anything else is an argument error returned to the caller:
//   -O0 .. -O3                          (replaces -O3)
//   -mllvm -amdgpu-sched-strategy=S     S in kJitSchedulers (replaces the
//                                       library's max-ilp choice)
const char* const kJitSchedulers[] = {"max-ilp", "max-memory-clause", "iterative-minreg",
                                      "iterative-maxocc"};

bool jit_user_options(std::string* opt_level, std::string* sched, std::string* err) {
  const char* e = std::getenv("PASIF_JIT_FLAGS");
  if (!e) return true;
  std::vector<std::string> tok;
  std::string t;
  for (const char* c = e;; ++c) {
    if (*c == ' ' || *c == '\t' || *c == '\0') {
      if (!t.empty()) tok.push_back(t);
      t.clear();
      if (*c == '\0') break;
    } else {
      t += *c;
    }
  }
  bool have_o = false, have_s = false;
  for (size_t i = 0; i < tok.size(); ++i) {
    const std::string& x = tok[i];
    if (x == "-O0" || x == "-O1" || x == "-O2" || x == "-O3") {
      if (have_o) {
        *err = "PASIF_JIT_FLAGS: more than one optimisation level";
        return false;
      }
      have_o = true;
      *opt_level = x;
      continue;
    }
    const std::string pre = "-amdgpu-sched-strategy=";
    if (x == "-mllvm" && i + 1 < tok.size() && tok[i + 1].compare(0, pre.size(), pre) == 0) {
      const std::string v = tok[i + 1].substr(pre.size());
      bool ok = false;
      for (const char* s : kJitSchedulers) ok = ok || v == s;
      if (!ok) {
        *err = "PASIF_JIT_FLAGS: scheduler '" + v +
               "' is not supported (accepted: max-ilp, max-memory-clause, iterative-minreg, "
               "iterative-maxocc)";
        return false;
      }
      if (have_s) {
        *err = "PASIF_JIT_FLAGS: -amdgpu-sched-strategy given more than once";
        return false;
      }
      have_s = true;
      *sched = v;
      ++i;
      continue;
    }
    *err = "PASIF_JIT_FLAGS: unsupported option '" + x +
           "' (accepted: -O0..-O3, -mllvm -amdgpu-sched-strategy=<scheduler>)";
    return false;
  }
  return true;
}

// Options of one lane-kernel compilation: the defaults, the machine scheduler
// (max-ilp for launches bound by a lone wave's dependency chains, `ilp`), and
// the caller's replacements; they are part of the cache key.
bool jit_options(bool ilp, std::vector<std::string>* o, std::string* err) {
  std::string opt_level = "-O3", sched = ilp ? "max-ilp" : "";
  if (!jit_user_options(&opt_level, &sched, err)) return false;
  o->assign(kJitOptions, kJitOptions + kJitOptionCount);
  for (std::string& x : *o)
    if (x == "-O3") x = opt_level;
  if (!sched.empty()) {
    o->push_back("-mllvm");
    o->push_back("-amdgpu-sched-strategy=" + sched);
  }
  return true;
}

std::mutex g_jit_mu;
std::map<std::pair<int, uint64_t>, LaneKernel> g_jit;

}  // namespace

namespace {

struct MwDims {
  uint32_t nymax = 1, nfmax = 1, nfslot = 0;
  bool adjoint = false;
};

MwDims mw_dims(const std::vector<LaneShape>& shapes, uint32_t waves) {
  MwDims d;
  for (const LaneShape& sh : shapes) {
    auto count = [](const std::vector<int>& v) {
      uint32_t k = 0;
      for (int x : v) k += x >= 0 ? 1 : 0;
      return k;
    };
    d.nymax = std::max(d.nymax, count(read_slots(sh.f, true)));
    if (sh.adjoint) d.nymax = std::max(d.nymax, count(read_slots(sh.a, false)));
    d.nfmax = std::max(d.nfmax, sh.f.n);
    d.adjoint = d.adjoint || sh.adjoint;
    std::vector<bool> drives(waves, false);
    for (uint32_t r : sh.f.forced) drives[sh.fwave[r]] = true;
    uint32_t k = 0;
    for (bool b : drives) k += b ? 1 : 0;
    d.nfslot = std::max(d.nfslot, k);
  }
  return d;
}

size_t mw_lds_values(const MwDims& d, const LaneLayout& lay) {
  return size_t(lay.lw) * (2 * size_t(d.nymax) + (d.adjoint ? size_t(lay.seg + 1) * d.nfmax : 0) +
                           size_t(d.nfslot) * kLaneStage);
}

}  // namespace

std::string lane_source(const std::vector<LaneShape>& shapes, int bits, const LaneLayout& lay) {
  const bool mw = lay.waves > 1;
  std::ostringstream o;
  o << "// generated lane kernel: " << shapes.size() << " structure class(es), FP" << bits
    << ", " << lay.waves << " wave(s) per workgroup, " << (mw ? lay.lw : kLaneWave)
    << " instances per workgroup, segment " << lay.seg << ", occupancy " << lay.occ << "\n";
  o << "typedef unsigned int u32;\ntypedef int i32;\ntypedef unsigned long long u64;\n";
  o << "typedef " << (bits == 64 ? "double" : "float") << " R;\n";
  o << "struct LaneArgs {\n";
#define PASIF_LANE_TEXT(T, name) o << "  " << type_text(#T) << " " << #name << ";\n";
  PASIF_LANE_FIELDS(PASIF_LANE_TEXT)
#undef PASIF_LANE_TEXT
  o << "};\n";
  o << "__device__ __forceinline__ u32 find_entry(const LaneArgs* cls, u32 n, u32 b) {\n"
       "  u32 lo = 0, hi = n;\n"
       "  while (hi - lo > 1u) {\n"
       "    const u32 mid = (lo + hi) >> 1;\n"
       "    if (b >= (u32)cls[mid].block_begin) lo = mid; else hi = mid;\n"
       "  }\n"
       "  return (u32)__builtin_amdgcn_readfirstlane((int)lo);\n"
       "}\n";
  o << "#define WAVE " << kLaneWave << "u\n#define CHUNK " << kLaneChunk << "u\n#define STAGE "
    << kLaneStage << "u\n";
  MwDims dims;
  if (mw) dims = mw_dims(shapes, lay.waves);
  o << "#define LW " << (mw ? lay.lw : kLaneWave) << "u\n#define NYMAX " << dims.nymax
    << "u\n#define NFMAX " << dims.nfmax << "u\n#define SEGK " << lay.seg << "u\n#define SEGN "
    << (dims.adjoint ? size_t(lay.seg + 1) * dims.nfmax * lay.lw : 0) << "u\n#define FORCING "
    << lay.forcing << "\n#define FULL " << (lay.full ? 1 : 0) << "\n#define EXACT "
    << (lay.exact && !mw ? 1 : 0) << "\n";
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
    const bool grp = !mw && sh.group > 1;
    const uint32_t G = grp ? sh.group : 1u;
    const uint32_t nf = grp ? sh.gf.nl() : sh.f.n;
    const uint32_t na = sh.adjoint ? (grp ? sh.ga.nl() : sh.a.n) : 0u;
    o << "struct C" << c << " {\n";
    o << "  enum { NF = " << nf << ", NA = " << na
      << ", TF = " << (grp ? sh.gf.coefs() : sh.f.coefs())
      << ", TA = " << (sh.adjoint ? (grp ? sh.ga.coefs() : sh.a.coefs()) : 0u)
      << ", ADJ = " << (sh.adjoint ? 1 : 0) << ", G = " << G
      << ", RLF = " << (grp ? sh.gf.RL : nf) << ", RLA = " << (grp ? sh.ga.RL : na)
      << ", NX = " << (grp ? uint32_t(sh.gx.size()) : nf);
    std::vector<int> segmap;
    if (!mw) o << ", NSEG = " << lane_seg_rows(sh, &segmap);
    o << " };\n";
    if (!mw) {
      o << "  static __device__ __forceinline__ constexpr int segr(int r) { return "
        << ternary_map_i(segmap) << "; }\n";
      std::vector<uint32_t> rf, ra;
      if (grp) {
        rf = sh.gf.band;
        rf.insert(rf.end(), sh.gf.shared.begin(), sh.gf.shared.end());
        ra = sh.ga.band;
        ra.insert(ra.end(), sh.ga.shared.begin(), sh.ga.shared.end());
        o << "  static __device__ __forceinline__ constexpr int rowf(int r) { return "
          << ternary_map(rf) << "; }\n";
        o << "  static __device__ __forceinline__ constexpr int rowa(int r) { return "
          << ternary_map(ra) << "; }\n";
        o << "  static __device__ __forceinline__ constexpr int stridef(int r) { return "
          << ternary_map(sh.gf.stride) << "; }\n";
        o << "  static __device__ __forceinline__ constexpr int stridea(int r) { return "
          << ternary_map(sh.ga.stride) << "; }\n";
        std::string sk = "false";
        for (uint32_t r = 0; r < sh.gf.sink.size(); ++r)
          if (sh.gf.sink[r]) sk = "r == " + std::to_string(r) + " || " + sk;
        o << "  static __device__ __forceinline__ constexpr bool sink(int r) { return " << sk
          << "; }\n";
      } else {
        o << "  static __device__ __forceinline__ constexpr int rowf(int r) { return r; }\n";
        o << "  static __device__ __forceinline__ constexpr int rowa(int r) { return r; }\n";
        o << "  static __device__ __forceinline__ constexpr int stridef(int) { return 1; }\n";
        o << "  static __device__ __forceinline__ constexpr int stridea(int) { return 1; }\n";
        o << "  static __device__ __forceinline__ constexpr bool sink(int) { return false; }\n";
      }
      o << "  static __device__ __forceinline__ void rhs_f(R* k, const R* y, const R* c, R u) {\n";
      if (grp) emit_group_side(o, sh.gf, G, true, sh.gx);
      else emit_rhs_forward(o, sh.f);
      o << "  }\n";
      // set points of the reverse step: forward values read from the
      // segment in LDS (another position's column for rotated sources)
      o << "  static __device__ __forceinline__ void xload(R* xc, R* xp, const R* sc, const R* sp, "
           "u32 lane) {\n";
      if (grp) {
        std::vector<uint32_t> ds;
        for (const GSrc& x : sh.gx)
          if (!x.shared && x.d != 0) ds.push_back(x.d);
        std::sort(ds.begin(), ds.end());
        ds.erase(std::unique(ds.begin(), ds.end()), ds.end());
        for (uint32_t dd : ds)
          o << "    const u32 l" << dd << " = glane<" << G << "u, " << dd << "u>(lane);\n";
        for (size_t j = 0; j < sh.gx.size(); ++j) {
          const GSrc& x = sh.gx[j];
          const std::string col = (x.shared || x.d == 0) ? "lane" : "l" + std::to_string(x.d);
          const int r = segmap[x.slot];
          o << "    xc[" << j << "] = sc[" << r << "u * WAVE + " << col << "];\n";
          o << "    xp[" << j << "] = sp[" << r << "u * WAVE + " << col << "];\n";
        }
      } else if (sh.adjoint) {
        // rows the reverse step does not read are not in the segment
        for (uint32_t j = 0; j < nf; ++j) {
          if (segmap[j] < 0) {
            o << "    xc[" << j << "] = R(0);\n    xp[" << j << "] = R(0);\n";
          } else {
            o << "    xc[" << j << "] = sc[" << segmap[j] << "u * WAVE + lane];\n";
            o << "    xp[" << j << "] = sp[" << segmap[j] << "u * WAVE + lane];\n";
          }
        }
      }
      o << "  }\n";
      std::ostringstream oe, ol;
      uint32_t ne = 0;
      if (sh.adjoint) {
        const std::vector<ERow> rows = grp ? group_adjoint_rows(sh.ga, sh.gx) : lane_adjoint_rows(sh.a);
        if (grp) {
          // own-state values of other positions, rotated once per stage
          std::vector<GSrc> rem;
          for (const GRow& row : sh.ga.rows)
            for (const GTerm& t : row.terms)
              if (!t.f[0].shared && t.f[0].d != 0) rem.push_back(t.f[0]);
          std::sort(rem.begin(), rem.end());
          rem.erase(std::unique(rem.begin(), rem.end()), rem.end());
          for (const GSrc& r : rem)
            ol << "    const R y" << r.slot << "_" << r.d << " = rot<" << G << "u, " << r.d
               << "u>(y[" << r.slot << "]);\n";
        }
        std::ostringstream ol2;
        emit_adjoint_linear(oe, ol2, rows, G, &ne);
        ol << ol2.str();
      }
      o << "  enum { NE = " << ne << " };\n";
      o << "  static __device__ __forceinline__ void eff_a(R* e, const R* x, const R* c) {\n"
        << oe.str() << "  }\n";
      o << "  static __device__ __forceinline__ void lin_a(R* k, const R* y, const R* e) {\n"
        << ol.str() << "  }\n";
    }
    o << "};\n";
    if (!mw) continue;
    int fslot = 0;
    for (uint32_t w = 0; w < lay.waves; ++w) {
      std::vector<uint32_t> rf, ra;
      for (uint32_t r = 0; r < sh.f.n; ++r)
        if (sh.fwave[r] == w) rf.push_back(r);
      if (sh.adjoint)
        for (uint32_t r = 0; r < sh.a.n; ++r)
          if (sh.awave[r] == w) ra.push_back(r);
      bool drives = false;
      for (uint32_t r : sh.f.forced) drives = drives || sh.fwave[r] == w;
      std::ostringstream bf, ba;
      std::vector<uint32_t> gf, ga;
      emit_wave_side(bf, sh.f, rf, true, false, &gf);
      if (sh.adjoint) emit_wave_side(ba, sh.a, ra, false, true, &ga);
      o << "struct C" << c << "W" << w << " {\n";
      o << "  enum { RF = " << rf.size() << ", RA = " << ra.size() << ", TF = " << gf.size()
        << ", TA = " << ga.size() << ", FSLOT = " << (drives ? fslot : -1) << " };\n";
      if (drives) ++fslot;
      o << "  static __device__ __forceinline__ constexpr int rf(int r) { return "
        << ternary_map(rf) << "; }\n";
      o << "  static __device__ __forceinline__ constexpr int ra(int r) { return "
        << ternary_map(ra) << "; }\n";
      {
        const std::vector<int> yfs = read_slots(sh.f, true);
        std::vector<int> yf, ya;
        for (uint32_t r : rf) yf.push_back(yfs[r]);
        if (sh.adjoint) {
          const std::vector<int> yas = read_slots(sh.a, false);
          for (uint32_t r : ra) ya.push_back(yas[r]);
        }
        o << "  static __device__ __forceinline__ constexpr int yf(int r) { return "
          << ternary_imap(yf) << "; }\n";
        o << "  static __device__ __forceinline__ constexpr int ya(int r) { return "
          << ternary_imap(ya) << "; }\n";
      }
      o << "  static __device__ __forceinline__ void load_f(R* c, GP(const R) cf, u64 st) {\n";
      for (size_t t = 0; t < gf.size(); ++t)
        o << "    c[" << t << "] = cf[" << gf[t] << "ull * st];\n";
      o << "  }\n";
      o << "  static __device__ __forceinline__ void load_a(R* c, GP(const R) ca, u64 st) {\n";
      for (size_t t = 0; t < ga.size(); ++t)
        o << "    c[" << t << "] = ca[" << ga[t] << "ull * st];\n";
      o << "  }\n";
      o << "  static __device__ __forceinline__ void rhs_f(R* k, const R* Y, const R* c, R u) {\n"
        << bf.str() << "  }\n";
      o << "  static __device__ __forceinline__ void rhs_a(R* k, const R* Y, const R* Xc, "
           "const R* Xp, const R* c, int s) {\n"
        << ba.str() << "  }\n";
      o << "};\n";
    }
  }
  if (!mw) {
    o << "extern \"C\" __global__ void __launch_bounds__(" << kLaneWave << ") ";
    if (lay.occ > 1) o << "__attribute__((amdgpu_waves_per_eu(" << lay.occ << "))) ";
    o << "pasif_lane_kernel(const LaneArgs* __restrict__ cls, u32 ncls) {\n"
         "  extern __shared__ R smem[];\n"
         "  const u32 b = blockIdx.x;\n"
         "  const u32 e = find_entry(cls, ncls, b);\n"
         "  const LaneArgs& a = cls[e];\n"
         "  switch (uni((u32)a.klass)) {\n";
    for (size_t c = 0; c < shapes.size(); ++c)
      o << "    case " << c << ": run<C" << c << ">(a, b - (u32)a.block_begin, smem); break;\n";
    o << "    default: break;\n  }\n}\n";
  } else {
    o << "extern \"C\" __global__ void __launch_bounds__(" << kLaneWave * lay.waves
      << ") pasif_lane_kernel(const LaneArgs* __restrict__ cls, u32 ncls) {\n"
      << "  __shared__ R smem[" << mw_lds_values(dims, lay) << "];\n"
      << "  const u32 b = blockIdx.x;\n"
         "  const u32 wid = uni(threadIdx.x >> 6);\n"
         "  const u32 e = find_entry(cls, ncls, b);\n"
         "  const LaneArgs& a = cls[e];\n"
         "  switch (uni((u32)a.klass)) {\n";
    for (size_t c = 0; c < shapes.size(); ++c) {
      o << "    case " << c << ":\n      switch (wid) {\n";
      for (uint32_t w = 0; w < lay.waves; ++w)
        o << "        case " << w << ": run_mw<C" << c << ", C" << c << "W" << w
          << ">(a, b - (u32)a.block_begin, smem); break;\n";
      o << "        default: break;\n      }\n      break;\n";
    }
    o << "    default: break;\n  }\n}\n";
  }
  return o.str();
}

size_t lane_lds_bytes(const std::vector<LaneShape>& shapes, int bits, const LaneLayout& lay) {
  const size_t el = bits == 64 ? 8 : 4;
  if (lay.waves > 1) return mw_lds_values(mw_dims(shapes, lay.waves), lay) * el;
  // per class: the sample stage (a column per instance), then the segment
  size_t v = 0;
  for (const LaneShape& s : shapes) {
    const uint32_t G = s.group > 1 ? s.group : 1u;
    const uint32_t K = s.seg ? s.seg : lay.seg;
    size_t c = size_t(kLaneStage + (G > 1 ? 3 * kLaneChunk : 0)) * (kLaneWave / G);
    if (s.adjoint) c += lane_seg_values(K, lay.exact) * lane_seg_rows(s) * kLaneWave;
    v = std::max(v, c);
  }
  return v * el;
}

namespace {
// hipRTC compilation of `src` to a gfx950 code object
bool rtc_compile(const std::string& src, std::vector<char>* code, std::string* err,
                 bool ilp = false) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "pasif_lane.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    *err = "hipRTC: cannot create the lane kernel program";
    return false;
  }
  std::vector<std::string> opt;
  if (!jit_options(ilp, &opt, err)) {
    hiprtcDestroyProgram(&prog);
    return false;
  }
  std::vector<const char*> optp;
  for (const std::string& x : opt) optp.push_back(x.c_str());
  const hiprtcResult rc = hiprtcCompileProgram(prog, int(optp.size()), optp.data());
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

bool lane_jit_flags_ok(std::string* err) {
  std::string o, s;
  return jit_user_options(&o, &s, err);
}

bool lane_compile_only(const std::string& src, size_t* code_bytes, std::string* err, bool ilp) {
  std::vector<char> code;
  if (!rtc_compile(src, &code, err, ilp)) return false;
  if (code_bytes) *code_bytes = code.size();
  return true;
}

namespace {
std::string jit_cache_path(uint64_t key) {
  const std::string dir = cache_dir();
  if (dir.empty()) return std::string();
  char name[64];
  snprintf(name, sizeof name, "lane_%016llx.co", (unsigned long long)key);
  return dir + "/" + name;
}

void read_cached(const std::string& path, std::vector<char>* code) {
  if (path.empty()) return;
  if (FILE* f = fopen(path.c_str(), "rb")) {
    fseek(f, 0, SEEK_END);
    const long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz > 0) {
      code->resize(size_t(sz));
      if (fread(code->data(), 1, code->size(), f) != code->size()) code->clear();
    }
    fclose(f);
  }
}

void write_cached(const std::string& path, const std::vector<char>& code) {
  if (path.empty()) return;
  mkdirs(cache_dir());
  const std::string tmp = path + ".tmp" + std::to_string(long(getpid()));
  if (FILE* f = fopen(tmp.c_str(), "wb")) {
    const bool ok = fwrite(code.data(), 1, code.size(), f) == code.size();
    fclose(f);
    if (ok)
      std::rename(tmp.c_str(), path.c_str());
    else
      std::remove(tmp.c_str());
  }
}
}  // namespace

bool lane_compile_many(const std::vector<std::string>& srcs, const std::vector<bool>& ilp,
                       int dev, const std::vector<LaneKernel*>& outs, std::string* err) {
  const size_t n = srcs.size();
  std::vector<uint64_t> keys(n, 0);
  std::vector<std::vector<char>> code(n);
  std::vector<char> done(n, 0);
  std::vector<size_t> todo;
  bool mllvm = false;
  std::lock_guard<std::mutex> lk(g_jit_mu);
  for (size_t i = 0; i < n; ++i) {
    std::vector<std::string> optv;
    if (!jit_options(ilp[i], &optv, err)) return false;
    std::string opts;
    for (const std::string& x : optv) {
      opts += x + " ";
      mllvm = mllvm || x == "-mllvm";
    }
    keys[i] = fnv1a(srcs[i] + opts);
    if (dev >= 0) {
      auto it = g_jit.find({dev, keys[i]});
      if (it != g_jit.end()) {
        *outs[i] = it->second;
        done[i] = 1;
        continue;
      }
      read_cached(jit_cache_path(keys[i]), &code[i]);
    }
    if (code[i].empty()) todo.push_back(i);
  }
  std::vector<std::string> errs(n);
  std::vector<char> ok(n, 1);
  auto compile = [&](size_t j) {
    const size_t i = todo[j];
    ok[i] = rtc_compile(srcs[i], &code[i], &errs[i], ilp[i]) ? 1 : 0;
  };
  const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
  const size_t nt = mllvm ? 1 : std::min<size_t>({todo.size(), hw, 16});
  if (nt > 1) {
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nt; ++t)
      pool.emplace_back([&] {
        for (size_t j; (j = next.fetch_add(1)) < todo.size();) compile(j);
      });
    for (std::thread& t : pool) t.join();
  } else {
    for (size_t j = 0; j < todo.size(); ++j) compile(j);
  }
  for (size_t i : todo)
    if (!ok[i]) {
      *err = errs[i];
      return false;
    }
  if (dev < 0) return true;
  for (size_t i : todo) write_cached(jit_cache_path(keys[i]), code[i]);
  for (size_t i = 0; i < n; ++i) {
    if (done[i]) continue;
    LaneKernel k;
    hipError_t e = hipModuleLoadData(&k.mod, code[i].data());
    if (e != hipSuccess) {
      *err = std::string("loading the lane kernel module failed: ") + hipGetErrorName(e);
      return false;
    }
    e = hipModuleGetFunction(&k.fn, k.mod, "pasif_lane_kernel");
    if (e != hipSuccess) {
      *err = std::string("lane kernel not found in its module: ") + hipGetErrorName(e);
      return false;
    }
    g_jit[{dev, keys[i]}] = k;
    *outs[i] = k;
  }
  return true;
}

bool lane_compile(const std::string& src, int dev, LaneKernel* out, std::string* err, bool ilp) {
  return lane_compile_many({src}, {ilp}, dev, {out}, err);
}

namespace {
uint64_t jit_key(const std::string& src, bool ilp, std::string* err, bool* ok) {
  std::vector<std::string> optv;
  *ok = jit_options(ilp, &optv, err);
  std::string opts;
  for (const std::string& x : optv) opts += x + " ";
  return fnv1a(src + opts);
}
}  // namespace

bool lane_compile_code(const std::string& src, bool ilp, std::vector<char>* code,
                       std::string* err) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  bool ok = false;
  const uint64_t key = jit_key(src, ilp, err, &ok);
  if (!ok) return false;
  code->clear();
  read_cached(jit_cache_path(key), code);
  if (!code->empty()) return true;
  if (!rtc_compile(src, code, err, ilp)) return false;
  write_cached(jit_cache_path(key), *code);
  return true;
}

bool lane_load_code(const std::string& src, bool ilp, int dev, const std::vector<char>& code,
                    LaneKernel* out, std::string* err) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  bool ok = false;
  const uint64_t key = jit_key(src, ilp, err, &ok);
  if (!ok) return false;
  auto it = g_jit.find({dev, key});
  if (it != g_jit.end()) {
    *out = it->second;
    return true;
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
