// Host runtime of libpasif_hip.so: the C ABI declared in include/pasif.h.
//
// It replaces the reference's __GpuDriver (src/__GpuDriver.cu, bound by
// src/PASIfgpu.cpp): it owns the excitation set, the forward and adjoint
// systems as the Python layer assembles them (scaled CSR of B+K, COO tensors,
// force pattern, initial / terminal conditions), splits the block-diagonal
// global system back into its independent systems, packs (system,
// excitation) instances into tiles (pasif_tiles.hpp) and runs the tile kernel
// (pasif_kernels.hip) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <map>
#include <set>
#include <thread>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/pasif.h"
#include "pasif_host.hpp"
#include "pasif_lane.hpp"
#include "pasif_tiles.hpp"

using namespace pasif;

namespace {

struct Fail {
  int code;
};

// ------------------------------------------------------------------ device
// Sub-allocator of a pool's device memory: chunks taken with hipMalloc once
// and carved into 256-byte aligned blocks (first fit over an offset-ordered,
// coalescing free list).  Admitting or retiring a job then neither allocates
// nor frees device memory (hipFree synchronises the whole device).  A block
// is only handed out again after the job that held it left the pool, i.e.
// after the last launch that could read it completed.
struct Arena {
  // Free blocks of a chunk, indexed by offset (to merge neighbours when a
  // block comes back) and by size (best fit in O(log n): with thousands of
  // resident jobs the free list is long and a first-fit scan dominated
  // admission).
  struct Chunk {
    char* base;
    size_t cap;
    std::map<size_t, size_t> free_;           // offset -> bytes
    std::multimap<size_t, size_t> by_size;    // bytes -> offset
    void add(size_t off, size_t len) {
      free_.emplace(off, len);
      by_size.emplace(len, off);
    }
    void drop_size(size_t len, size_t off) {
      auto r = by_size.equal_range(len);
      for (auto it = r.first; it != r.second; ++it)
        if (it->second == off) {
          by_size.erase(it);
          return;
        }
    }
  };
  std::vector<Chunk> chunks;
  size_t chunk_bytes = size_t(1) << 30;
  size_t reserved = 0;
  static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~size_t(255); }
  void* take(size_t bytes) {
    const size_t need = round(bytes);
    for (Chunk& c : chunks) {
      auto it = c.by_size.lower_bound(need);
      if (it == c.by_size.end()) continue;
      const size_t len = it->first, off = it->second;
      c.by_size.erase(it);
      c.free_.erase(off);
      if (len > need) c.add(off + need, len - need);
      return c.base + off;
    }
    Chunk c{};
    c.cap = std::max(chunk_bytes, need);
    void* base = nullptr;
    if (hipMalloc(&base, c.cap) != hipSuccess) return nullptr;
    c.base = static_cast<char*>(base);
    reserved += c.cap;
    if (c.cap > need) c.add(need, c.cap - need);
    chunks.push_back(std::move(c));
    return chunks.back().base;
  }
  void give(void* p, size_t bytes) {
    char* q = static_cast<char*>(p);
    for (Chunk& c : chunks) {
      if (q < c.base || q >= c.base + c.cap) continue;
      size_t off = size_t(q - c.base), len = round(bytes);
      auto nx = c.free_.lower_bound(off);
      if (nx != c.free_.end() && off + len == nx->first) {
        len += nx->second;
        c.drop_size(nx->second, nx->first);
        nx = c.free_.erase(nx);
      }
      if (nx != c.free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
          off = pv->first;
          len += pv->second;
          c.drop_size(pv->second, pv->first);
          c.free_.erase(pv);
        }
      }
      c.add(off, len);
      return;
    }
  }
  ~Arena() {
    for (Chunk& c : chunks) (void)hipFree(c.base);
  }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;  // block of a pool's arena (else a hipMalloc)
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      if (arena)
        arena->give(p, bytes);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    arena = nullptr;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// ------------------------------------------------------------------ inputs
struct CsrIn {
  bool set = false;
  uint32_t n0 = 0, n1 = 0;
  std::vector<double> val;
  std::vector<uint32_t> idx, ptr;
};
struct CooIn {
  bool set = false;
  int order = 0;
  std::array<uint32_t, 5> dims{};
  std::vector<double> val;
  std::vector<uint32_t> idx;  // interleaved
};
struct SideIn {
  CsrIn K;
  CooIn G, L, P;
  bool force_set = false, init_set = false;
  std::vector<double> force, init;
  SideIn() {
    G.order = 3;
    L.order = 4;
    P.order = 5;
  }
};

// One structure class of the lane kernel: its instances (one per lane,
// sorted by step count, descending) and their device buffers.  `waves`
// counts the class's workgroups (one progress record each).
struct LaneCls {
  uint32_t count = 0, stride = 0, waves = 0, block_begin = 0, tmax = 0;
  uint32_t nf = 0, na = 0;
  uint32_t klass = 0;  // the class's case in its kernel's dispatch
  uint32_t seg = 1;    // checkpoint interval (adjoint)
  // single-wave adjoint: checkpoint elements, each wave's block sized by its
  // longest instance (offsets in ckoff)
  uint64_t ck_elems = 0;
  // the longest instance of each workgroup (instances are sorted by step
  // count, descending, so this is its first; 0 past the last instance) and
  // the workgroups a pool tick still launches: a prefix, since a workgroup
  // whose longest instance's work the completed ticks cover has finished
  std::vector<uint32_t> wave_t;
  uint32_t live = 0;
  // program arrays: views into the part's one program buffer
  const void *coef_f = nullptr, *coef_a = nullptr, *init_f = nullptr, *init_a = nullptr,
             *steps = nullptr, *exc = nullptr, *out_f = nullptr, *out_a = nullptr,
             *na_out = nullptr, *ckoff = nullptr;
  DevBuf qstate, lstate, mxstate, progress, ckpt, segbuf;  // run
};

// One lane kernel's share of a packed problem: the instances of each of its
// structure classes and the kernel (with its layout) that runs them.  A
// driver's plan has one part; a pool job has one per pool lane module it uses.
struct LanePart {
  std::vector<std::unique_ptr<LaneCls>> cls;
  DevBuf prog;  // every class's program arrays, uploaded in one copy
  LaneKernel kern;
  LaneLayout lay;
  uint32_t blocks = 0, tmax = 0;
  size_t lds = 0;
  DevBuf args;
  int module = -1;  // pool lane module (pool jobs)
};

struct Packed {
  bool valid = false;
  int precision = 0;
  bool adjoint = false;
  // lane kernel parts (and tile part) of the packed problem
  bool lane = false, tile = false;
  uint32_t ltmax = 0;
  std::vector<std::unique_ptr<LanePart>> lparts;
  // the driver's own lane plan (diagnostics): its classes, layout, checkpoint
  std::vector<LaneShape> lshapes;
  std::vector<uint32_t> lgroups;  // lanes per instance of each class in lshapes
  std::vector<std::pair<std::vector<LaneShape>, LaneLayout>> lmods;  // per module
  uint32_t lseg = 1;
  LaneLayout llay;
  uint32_t lanes = 0, ntiles = 0, slot_cfg = 0, tile_target = 0;
  uint64_t key = 0;
  uint32_t out_len_f = 0, out_len_a = 0, tmax = 0;
  DevBuf hdr;
  DevBuf f_info, f_out, f_src, f_steps, f_exc, f_fslot, f_force, f_val, f_idx, f_init;
  DevBuf a_info, a_out, a_src, a_steps, a_exc, a_fslot, a_force, a_val, a_idx, a_init;
  // adjoint: checkpoint interval, first checkpoint slot of each tile (a tile
  // holds ceil(its steps / interval) slots of `lanes` values), slots in all
  DevBuf ck_tile;
  uint32_t ck_every = 0;
  uint64_t ck_slots = 0;
  // row-parallel path (systems larger than a workgroup's lanes): instances,
  // packed sides, per-instance scratch sizes (elements)
  bool big = false;
  uint32_t nbig = 0;
  DevBuf b_inst, b_rf, b_vf, b_xf, b_ff, b_if, b_ra, b_va, b_xa, b_ia;
  DevBuf b_sgf, b_spf, b_sga, b_spa;  // segment lists and their (row, kind) ranges
  uint64_t b_part = 0;                 // segment partial sums, all instances
  uint32_t b_part_max = 0;             // most segments of one instance's side
  uint64_t b_vec_f = 0, b_vec_a = 0, b_ck = 0, b_seg = 0;
  uint32_t b_tmax = 0, b_every = 1;
  uint32_t max_forced = 0;
  size_t program_bytes = 0;
  void release() {
    DevBuf* all[] = {&hdr,     &f_info,  &f_out,  &f_src,   &f_steps, &f_exc, &f_fslot,
                     &f_force, &f_val,   &f_idx,  &f_init,  &a_info,  &a_out, &a_src,
                     &a_steps, &a_exc,   &a_fslot, &a_force, &a_val,  &a_idx, &a_init,
                     &ck_tile, &b_inst,  &b_rf,    &b_vf,    &b_xf,    &b_ff,  &b_if,
                     &b_ra,    &b_va,    &b_xa,    &b_ia,    &b_sgf,   &b_spf, &b_sga,
                     &b_spa};
    for (DevBuf* b : all) b->release();
    lparts.clear();
    lshapes.clear();
    lgroups.clear();
    lmods.clear();
    lane = false;
    tile = false;
    big = false;
    nbig = 0;
    b_part = 0;
    b_part_max = 0;
    ltmax = 0;
    valid = false;
  }
};

}  // namespace

struct pasif_driver {
  int dev = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  bool d_compute = false, d_system = false, d_solver = false;

  // excitations
  std::vector<double> exc;
  uint32_t n_exc = 0, exc_len = 0, sample_rate = 1, numsteps = 0;
  std::vector<double> interp;  // rows 0 and 1, [2][W]
  uint32_t interp_w = 0;
  std::vector<double> modulation;

  SideIn fin, bin;
  bool sys_ready = false, adj_ready = false;
  RowTerms fwd, adj;

  // layout extension
  std::vector<uint32_t> hint_f, hint_a;
  std::vector<int32_t> map_exc;
  std::vector<uint32_t> map_steps;
  bool mapped = false;
  uint32_t layout_systems = 0;

  int precision = 64;
  // adjoint recursion: 0 the reference's (continuous adjoint, interpolated
  // set points), 1 the exact discrete adjoint of the forward RK4 map
  int adj_mode = 0;
  int solver = PASIF_SOLVER_PERSISTENT;
  uint32_t tile_lanes = 0, slot_cfg = 0, ck_every = 0, window = 0;
  int kernel_path = PASIF_KERNEL_AUTO;
  // lanes per instance requested for lane-kernel classes (0: chosen per class)
  uint32_t lane_group = 0;
  // steps along each class's feasible group sizes from the cost model's
  // choice (the autotuner probes -1, 0, +1)
  int lane_bias = 0;
  // lane class of every system in the packed problem, -1: tile kernel
  std::vector<int32_t> lane_cls;
  // systems on the row-parallel path (too large for a tile), by system
  std::vector<uint8_t> big_sys;
  // structure of every system in the lane plan before classes were merged
  std::vector<uint32_t> lane_struct;
  // the last lane plan of each mode (forward, adjoint) by the signature of
  // everything it depends on but coefficient values: a reload of new values
  // into the same structure reuses its classes, layouts and kernels
  struct LanePlanCache {
    bool valid = false;
    uint64_t sig = 0;
    std::vector<int32_t> lane_cls;
    std::vector<uint32_t> lane_struct;
    std::vector<std::vector<LaneShape>> shapes;
    std::vector<LaneLayout> lays;
    std::vector<LaneKernel> kerns;
  } lcache[2];
  // second stream and fork / join events when lane and tile kernels run together
  // side streams (lane modules beside each other and beside the tiles), the
  // fork event and one join event per side stream
  static constexpr int kSide = 3;
  hipStream_t aux[kSide] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[kSide] = {nullptr, nullptr, nullptr};
  uint32_t graph_slice = 256, graph_launches = 16;

  Packed pk;
  // device copies of excitations / interpolation for the packed precision
  DevBuf d_exc, d_interp;
  int d_exc_prec = 0;
  bool d_exc_dirty = true;
  // run buffers
  DevBuf qstate, lstate, mxstate, progress, out_q, out_max, traj, out_grad, ckpt, segbuf;
  // run buffers of the row-parallel path
  DevBuf bvec_f, bvec_a, bsetp, bckpt, bseg, bprog, bpart;
  // results of the last solve (host, double)
  std::vector<double> res_q, res_max;
  bool have_max = false;
  double last_ms = 0.0;
  uint32_t last_launches = 0;
  // a job of a pool: uses the pool's device and upload stream
  pasif_pool* pool = nullptr;
  // destroyed by its owner while a tick in flight still ran it: the pool
  // deletes it once the tick is over
  bool orphan = false;
  uint32_t pool_ck = 0;
  // a pool job's lane kernel arguments, one vector per lane part (ticks
  // copy them into the pool's per-module tables)
  std::vector<std::vector<LaneArgs>> pool_largs;
  // excitations / interpolation rows owned by the pool (shared signal bank)
  const void* ext_exc = nullptr;
  const void* ext_interp = nullptr;
};

// Continuous-batching pool: jobs (each a driver holding one request's
// systems) are packed into tiles of a common shape; every tick advances all
// unfinished jobs' tiles by `budget` steps in one launch.
struct pasif_pool {
  int dev = 0;
  hipStream_t stream = nullptr;   // ticks
  hipStream_t upload = nullptr;   // admission (packing, uploads)
  hipEvent_t ev0 = nullptr, ev1 = nullptr, uploaded = nullptr;
  std::string err;
  int precision = 64;
  uint32_t lanes = 256, slot_cfg = 1, ck_every = 256, budget = 1000;
  double hbm_budget = 0.0;
  double used = 0.0;
  struct Job {
    pasif_driver* d;
    bool adjoint;
    uint64_t work, done;
    double bytes;
    bool finished;
    uint32_t tiles;
  };
  std::vector<Job> jobs;
  // host staging of the lane programs of admitted jobs (reused)
  std::vector<unsigned char> stage_host;
  // every job handle the pool created and that is not destroyed yet (when
  // the pool goes first, their device memory is freed and they are detached)
  std::set<pasif_driver*> handles;
  // shared signal bank, uploaded once in the pool's precision
  DevBuf bank, bank_interp;
  uint32_t bank_n = 0, bank_len = 0, bank_w = 0, bank_rate = 0;
  std::vector<unsigned char> table_host;
  std::vector<uint32_t> map_host;
  void* table_dev = nullptr;
  size_t table_bytes = 0;
  void* map_dev = nullptr;
  size_t map_bytes = 0;
  bool in_flight = false;
  double last_ms = 0.0;
  uint32_t last_blocks = 0;
  // slot cap: resident tiles (0 = only the HBM budget limits admission)
  uint32_t max_tiles = 0, tiles_used = 0;
  // admission time by phase (PASIF_POOL_TIMING set: printed when the pool is
  // destroyed): concatenation, packing, run buffers and arguments, all of
  // admission, the lane programs' uploads (part of packing)
  bool timing = std::getenv("PASIF_POOL_TIMING") != nullptr;
  double tm[5] = {0, 0, 0, 0, 0};
  uint64_t tm_n = 0;
  // device memory of the jobs
  std::unique_ptr<Arena> arena;
  // registered systems: one forward block each (local indices) and, when
  // given, its adjoint block (factor 0 local adjoint rows, factors 1..3
  // local forward rows)
  struct Entry {
    RowTerms f, a;
    bool has_adj = false;
  };
  std::vector<Entry> reg;
  // Lane modules over the registered structures, built when lane work is
  // first submitted: 0 = classes that fit one lane, 1 = classes spread over
  // the waves of a workgroup.  reg_lane[id] = (module << 16 | class) or -1.
  struct Module {
    std::vector<LaneShape> shapes;
    LaneLayout lay;
    LaneKernel kern;
    size_t lds = 0;
    std::vector<LaneArgs> table;  // this tick's entries
    void* table_dev = nullptr;
    size_t table_bytes = 0;
    uint32_t blocks = 0;
  };
  std::vector<Module> mods;
  bool mods_built = false;
  std::map<std::vector<uint32_t>, int32_t> key_lane;
  // forward structure alone -> a class whose forward side it is (forward-only
  // jobs carrying their own arrays)
  std::map<std::vector<uint32_t>, int32_t> fkey_lane;
  std::vector<int32_t> reg_lane;
  // Structures of array-carrying jobs that no module runs (compile_new set):
  // queued, then compiled into further single-wave modules on a thread of
  // their own, a batch at a time; a finished batch is installed by the next
  // submit or tick, and later jobs of those structures run on it.
  struct NewClass {
    LaneShape sh;
    std::vector<uint32_t> key, fkey;
    uint32_t waves = 1;  // 1: single-wave module; >= 2: multi-wave, that many waves
  };
  struct Build {
    std::thread th;
    std::atomic<bool> done{false};
    bool ok = false;
    std::string src, err;
    std::vector<char> code;
    Module M;
    std::vector<NewClass> cls;
    bool merge = false;  // the classes of earlier modules, recompiled as one
  };
  bool compile_new = false;
  uint32_t mods_forcing = 0;  // the forcing rule the modules were compiled for
  // raw structure of an array job's system -> its class key (and the forward
  // part's length): repeated structures skip the lane-side analysis
  std::map<std::vector<uint32_t>, std::pair<std::vector<uint32_t>, uint32_t>> raw_key;
  std::vector<NewClass> pending;
  std::set<std::vector<uint32_t>> queued;  // keys pending, compiling or tried
  // single-wave classes of background modules not merged yet, with the
  // module each runs in: once the queue is empty and they span
  // kPoolMergeModules modules, they are compiled again as one module (one
  // launch per tick instead of several)
  std::vector<NewClass> loose;
  std::vector<int> loose_mod;
  std::unique_ptr<Build> build;
  uint32_t built_modules = 0, failed_builds = 0;
  // one stream (and join event) per lane module
  std::vector<hipStream_t> aux;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  // pinned staging of batched collects (grows, freed with the pool)
  unsigned char* pinned = nullptr;
  size_t pinned_bytes = 0;
};

namespace {

void set_err(pasif_driver* d, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  d->err = buf;
}

#define PASIF_HIP(call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      set_err(d, "HIP error %s (%d) at %s:%d: %s", hipGetErrorName(e_), (int)e_, __FILE__, \
              __LINE__, #call);                                                          \
      throw Fail{PASIF_ERR_DEVICE};                                                      \
    }                                                                                    \
  } while (0)

#define PASIF_REQUIRE(cond, code, ...) \
  do {                                 \
    if (!(cond)) {                     \
      set_err(d, __VA_ARGS__);         \
      throw Fail{code};                \
    }                                  \
  } while (0)

template <class F>
int guarded(pasif_driver* d, F&& f) {
  if (!d) return PASIF_ERR_ARGUMENT;
  try {
    d->err.clear();
    f();
    return PASIF_OK;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    d->err = "host memory exhausted";
    return PASIF_ERR_MEMORY;
  } catch (const std::exception& e) {
    d->err = e.what();
    return PASIF_ERR_ARGUMENT;
  }
}

static_assert(kScratchBytes == PASIF_NOALLOC_SCRATCH_BYTES, "scratch size stated in pasif.h");

// set while pasif_solve_noalloc runs on this thread: any device allocation
// is then a library bug, reported instead of performed
thread_local bool t_no_alloc = false;

void dev_alloc(pasif_driver* d, DevBuf& b, size_t bytes) {
  PASIF_REQUIRE(!t_no_alloc, PASIF_ERR_STATE,
                "internal error: device allocation of %zu bytes inside pasif_solve_noalloc",
                bytes);
  b.release();
  if (bytes == 0) bytes = 16;
  if (d->pool && d->pool->arena) {
    b.p = d->pool->arena->take(bytes);
    if (!b.p) {
      set_err(d, "device allocation of %zu bytes for the pool failed", bytes);
      throw Fail{PASIF_ERR_MEMORY};
    }
    b.arena = d->pool->arena.get();
    b.bytes = bytes;
    return;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    set_err(d, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorName(e));
    throw Fail{PASIF_ERR_MEMORY};
  }
  b.bytes = bytes;
}

// keep an existing allocation that is large enough (run buffers are reused
// across solves so a repeated solve allocates nothing)
void dev_reserve(pasif_driver* d, DevBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return;
  dev_alloc(d, b, bytes);
}

template <typename T>
void upload(pasif_driver* d, DevBuf& b, const std::vector<T>& h) {
  if (d->dev < 0) {  // host-only handle: packing and validation without a GPU
    b.release();
    b.bytes = h.size() * sizeof(T);
    return;
  }
  dev_alloc(d, b, h.size() * sizeof(T));
  if (!h.empty())
    PASIF_HIP(hipMemcpyAsync(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice,
                             d->stream));
}

void upload_bytes(pasif_driver* d, DevBuf& b, const void* h, size_t bytes) {
  if (d->dev < 0) {  // host-only handle: packing and validation without a GPU
    b.release();
    b.bytes = bytes;
    return;
  }
  dev_alloc(d, b, bytes);
  if (bytes)
    PASIF_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, d->stream));
}

void check_rows(pasif_driver* d, const CsrIn& K, const char* what) {
  PASIF_REQUIRE(K.set, PASIF_ERR_STATE, "%s: K (CSR of B+K) not set", what);
  PASIF_REQUIRE(K.n0 == K.n1, PASIF_ERR_ARGUMENT, "%s: K must be square (%u x %u)", what, K.n0,
                K.n1);
  PASIF_REQUIRE(K.ptr.size() == size_t(K.n0) + 1, PASIF_ERR_ARGUMENT, "%s: indptr length", what);
  PASIF_REQUIRE(K.ptr[0] == 0 && K.ptr.back() == K.val.size(), PASIF_ERR_ARGUMENT,
                "%s: indptr does not span the values", what);
  for (uint32_t r = 0; r < K.n0; ++r)
    PASIF_REQUIRE(K.ptr[r] <= K.ptr[r + 1], PASIF_ERR_ARGUMENT, "%s: indptr not monotone", what);
  for (uint32_t c : K.idx)
    PASIF_REQUIRE(c < K.n1, PASIF_ERR_ARGUMENT, "%s: column index %u out of range", what, c);
}

// Build per-row term lists (global indices) of one side.
void build_rows(pasif_driver* d, const SideIn& in, uint32_t nset, bool adjoint, RowTerms& out,
                const char* what) {
  check_rows(d, in.K, what);
  const uint32_t n = in.K.n0;
  PASIF_REQUIRE(in.init_set && in.init.size() == n, PASIF_ERR_ARGUMENT,
                "%s: %s has %zu entries, the system has %u DOFs", what,
                adjoint ? "terminal condition" : "initial condition", in.init.size(), n);
  PASIF_REQUIRE(!in.force_set || in.force.size() == n, PASIF_ERR_ARGUMENT,
                "%s: force pattern has %zu entries, the system has %u DOFs", what, in.force.size(),
                n);
  out = RowTerms{};
  out.n = n;
  out.nset = nset;
  out.init = in.init;
  out.force = in.force_set ? in.force : std::vector<double>(n, 0.0);
  // kind 1: CSR rows as given (the reference's SpMM order)
  out.ptr[0] = in.K.ptr;
  out.val[0] = in.K.val;
  out.idx[0] = in.K.idx;
  const CooIn* tens[3] = {&in.G, &in.L, &in.P};
  for (int k = 1; k < 4; ++k) {
    const CooIn& t = *tens[k - 1];
    const int order = k + 2;
    const size_t nnz = t.set ? t.val.size() : 0;
    if (t.set) {
      PASIF_REQUIRE(t.idx.size() == nnz * order, PASIF_ERR_ARGUMENT,
                    "%s: tensor of order %d has %zu indices for %zu values", what, order,
                    t.idx.size(), nnz);
      if (nnz) {
        PASIF_REQUIRE(t.dims[0] == n && t.dims[1] == n, PASIF_ERR_ARGUMENT,
                      "%s: tensor of order %d has leading dims %u x %u, expected %u", what, order,
                      t.dims[0], t.dims[1], n);
        for (int a = 2; a < order; ++a)
          PASIF_REQUIRE(t.dims[a] == nset, PASIF_ERR_ARGUMENT,
                        "%s: tensor of order %d axis %d has size %u, expected %u", what, order,
                        a, t.dims[a], nset);
      }
    }
    std::vector<uint32_t>& ptr = out.ptr[k];
    ptr.assign(n + 1, 0);
    for (size_t e = 0; e < nnz; ++e) {
      const uint32_t o = t.idx[e * order];
      PASIF_REQUIRE(o < n, PASIF_ERR_ARGUMENT, "%s: tensor output index %u out of range", what, o);
      for (int a = 1; a < order; ++a) {
        const uint32_t lim = a == 1 ? n : nset;
        PASIF_REQUIRE(t.idx[e * order + a] < lim, PASIF_ERR_ARGUMENT,
                      "%s: tensor index %u on axis %d out of range", what, t.idx[e * order + a], a);
      }
      ptr[o + 1]++;
    }
    for (uint32_t r = 0; r < n; ++r) ptr[r + 1] += ptr[r];
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    const int nf = k + 1;  // state factors of this kind
    out.val[k].assign(nnz, 0.0);
    out.idx[k].assign(nnz * nf, 0);
    for (size_t e = 0; e < nnz; ++e) {  // stable: keeps COO order within a row
      const uint32_t o = t.idx[e * order];
      const uint32_t pos = fill[o]++;
      out.val[k][pos] = t.val[e];
      for (int a = 0; a < nf; ++a) out.idx[k][size_t(pos) * nf + a] = t.idx[e * order + 1 + a];
    }
  }
}

// Finest partition of [0, n) into contiguous blocks closed under every
// term's indices (the block-diagonal structure PASIf.py:475-543 builds).
std::vector<uint32_t> detect_blocks(const RowTerms& rt, bool use_x) {
  const uint32_t n = rt.n;
  std::vector<uint32_t> reach(n);
  std::iota(reach.begin(), reach.end(), 0u);
  for (int k = 0; k < 4; ++k) {
    const int nidx = k + 1;
    for (uint32_t r = 0; r < n; ++r)
      for (uint32_t e = rt.ptr[k][r]; e < rt.ptr[k][r + 1]; ++e) {
        uint32_t lo = r, hi = r;
        const int lim = use_x ? nidx : 1;
        for (int a = 0; a < lim; ++a) {
          const uint32_t v = rt.idx[k][size_t(e) * nidx + a];
          lo = std::min(lo, v);
          hi = std::max(hi, v);
        }
        reach[lo] = std::max(reach[lo], hi);
      }
  }
  std::vector<uint32_t> sizes;
  uint32_t start = 0, far = 0;
  for (uint32_t i = 0; i < n; ++i) {
    far = std::max(far, reach[i]);
    if (far == i) {
      sizes.push_back(i + 1 - start);
      start = i + 1;
    }
  }
  return sizes;
}

void apply_partition(pasif_driver* d, RowTerms& rt, const std::vector<uint32_t>& sizes,
                     const char* what) {
  uint64_t tot = 0;
  for (uint32_t s : sizes) {
    PASIF_REQUIRE(s > 0, PASIF_ERR_ARGUMENT, "%s: empty system in the layout", what);
    tot += s;
  }
  PASIF_REQUIRE(tot == rt.n, PASIF_ERR_ARGUMENT, "%s: system sizes add up to %llu, not %u", what,
                (unsigned long long)tot, rt.n);
  rt.size = sizes;
  rt.off.assign(sizes.size(), 0);
  for (size_t s = 1; s < sizes.size(); ++s) rt.off[s] = rt.off[s - 1] + sizes[s - 1];
}

// every term of system s must stay inside its block (S) and inside the
// matching set-point block (X)
void validate_blocks(pasif_driver* d, const RowTerms& rt, const RowTerms* xside, const char* what) {
  for (size_t s = 0; s < rt.size.size(); ++s) {
    const uint32_t o0 = rt.off[s], o1 = o0 + rt.size[s];
    const uint32_t x0 = xside ? xside->off[s] : o0;
    const uint32_t x1 = xside ? x0 + xside->size[s] : o1;
    for (uint32_t r = o0; r < o1; ++r)
      for (int k = 0; k < 4; ++k) {
        const int nidx = k + 1;
        for (uint32_t e = rt.ptr[k][r]; e < rt.ptr[k][r + 1]; ++e)
          for (int a = 0; a < nidx; ++a) {
            const uint32_t v = rt.idx[k][size_t(e) * nidx + a];
            const bool ok = a == 0 ? (v >= o0 && v < o1) : (v >= x0 && v < x1);
            PASIF_REQUIRE(ok, PASIF_ERR_ARGUMENT,
                          "%s: row %u couples system %zu to index %u outside its block", what, r,
                          s, v);
          }
      }
  }
}

uint32_t seg_count(const RowTerms& rt, uint32_t r, const SlotConfig& c) {
  const int cap[4] = {c.c1, c.c2, c.c3, c.c4};
  uint32_t ns = 1;
  for (int k = 0; k < 4; ++k) {
    const uint32_t cnt = rt.ptr[k][r + 1] - rt.ptr[k][r];
    if (cnt) {
      if (cap[k] == 0) return 0;
      ns = std::max<uint32_t>(ns, (cnt + cap[k] - 1) / cap[k]);
    }
  }
  return ns;
}

uint32_t system_lanes(const RowTerms& rt, uint32_t s, const SlotConfig& c) {
  uint32_t l = 0;
  for (uint32_t r = rt.off[s]; r < rt.off[s] + rt.size[s]; ++r) {
    const uint32_t ns = seg_count(rt, r, c);
    if (ns == 0 || ns > 1023) return 0;
    l += ns;
  }
  return l;
}

template <typename R>
struct HostProgram {
  std::vector<uint32_t> info, out, src, steps, fslot;
  std::vector<int32_t> exc;
  std::vector<R> force, val;
  std::vector<uint32_t> idx;
  // Padding slots (value 0) point every factor at the zero row L.
  void resize(size_t ntiles, uint32_t L, const SlotConfig& c, uint32_t spad, uint32_t xpad) {
    const size_t nl = ntiles * L;
    info.assign(nl, make_lane(kRowIdle, 0, false));
    out.assign(nl, 0);
    src.assign(nl, 0);
    steps.assign(nl, 0);
    exc.assign(nl, -1);
    fslot.assign(nl, 0);
    force.assign(nl, R(0));
    val.assign(nl * slot_values(c), R(0));
    const int CW = slot_words(c);
    idx.resize(nl * CW);
    for (size_t t = 0; t < ntiles; ++t)
      for (int w = 0; w < CW; ++w) {
        const int rel = w - (c.c1 + c.c2);
        const bool second = rel >= 0 && (rel % 2) == 1;
        const uint32_t pad = second ? (xpad | (xpad << 16)) : (spad | (xpad << 16));
        std::fill_n(idx.begin() + (t * CW + w) * L, L, pad);
      }
  }
};

// Place one instance's rows of one side into tile `t` starting at lane `lo`
// and tile row `ro`.  X indices are offset by `xro` (the instance's forward
// rows in the tile).  Indices are stored as LDS byte offsets: rows of the
// stage buffer are `sstride` bytes apart, rows of the set-point buffer
// `xstride`.  Returns lanes used; updates the tile's trip counts.
template <typename R>
uint32_t place(const RowTerms& rt, uint32_t sys, const Instance& in, bool adjoint,
               HostProgram<R>& hp, size_t t, uint32_t L, uint32_t lo, uint32_t ro, uint32_t xro,
               const SlotConfig& c, uint32_t cnt[4], bool& split, uint32_t xoff,
               uint32_t sstride, uint32_t xstride, uint32_t& nforced) {
  const int cap[4] = {c.c1, c.c2, c.c3, c.c4};
  int vbase[4], wbase[4];
  vbase[0] = 0;
  vbase[1] = c.c1;
  vbase[2] = c.c1 + c.c2;
  vbase[3] = c.c1 + c.c2 + c.c3;
  wbase[0] = 0;
  wbase[1] = c.c1;
  wbase[2] = c.c1 + c.c2;
  wbase[3] = c.c1 + c.c2 + 2 * c.c3;
  const int CT = slot_values(c), CW = slot_words(c);
  const uint32_t off = rt.off[sys], n = rt.size[sys];
  const double eps = std::numeric_limits<R>::epsilon();
  uint32_t lane = lo;
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t g = off + r;
    const uint32_t ns = seg_count(rt, g, c);
    if (ns > 1) split = true;
    const size_t owner = t * L + lane;
    hp.info[owner] = make_lane(ro + r, ns, true);
    for (uint32_t sgm = 1; sgm < ns; ++sgm) hp.info[owner + sgm] = make_lane(ro + r, 0, false);
    hp.src[owner] = g;
    hp.steps[owner] = in.steps;
    if (!adjoint) {
      hp.out[owner] = in.out_f + r;
      const double f = rt.force[g];
      if (std::fabs(f) > eps && in.exc >= 0) {  // COOVector eps drop (helpers.cu:676)
        hp.force[owner] = R(f);
        hp.exc[owner] = in.exc;
        hp.fslot[owner] = nforced++;
      }
    } else {
      hp.out[owner] = in.out_a + r;
    }
    for (int k = 0; k < 4; ++k) {
      const uint32_t b = rt.ptr[k][g], e = rt.ptr[k][g + 1];
      const int nidx = k + 1;
      for (uint32_t i = 0; i < e - b; ++i) {
        const uint32_t sgm = i / cap[k], slot = i % cap[k];
        cnt[k] = std::max(cnt[k], slot + 1);
        const size_t ln = lane + sgm;
        const uint32_t* ix = &rt.idx[k][size_t(b + i) * nidx];
        uint32_t loc[4] = {L * xstride, L * xstride, L * xstride, L * xstride};
        loc[0] = (ro + (ix[0] - off)) * sstride;
        for (int a = 1; a < nidx; ++a) loc[a] = (xro + (ix[a] - xoff)) * xstride;
        hp.val[(t * CT + vbase[k] + slot) * L + ln] = R(rt.val[k][b + i]);
        if (k < 2) {
          hp.idx[(t * CW + wbase[k] + slot) * L + ln] = loc[0] | (loc[1] << 16);
        } else {
          hp.idx[(t * CW + wbase[k] + 2 * slot) * L + ln] = loc[0] | (loc[1] << 16);
          hp.idx[(t * CW + wbase[k] + 2 * slot + 1) * L + ln] = loc[2] | (loc[3] << 16);
        }
      }
    }
    lane += ns;
  }
  return lane - lo;
}

SlotConfig slot_of(uint32_t cfg) {
  return cfg == 2 ? kSlotWide : (cfg == 3 ? kSlotSmall : kSlotNarrow);
}

// Check, before anything reaches the GPU, that every index the kernel will
// follow stays inside the buffer it addresses: LDS byte offsets inside the
// stage / set-point buffers, rows inside the tile, segments inside the
// workgroup, initial-condition and output indices inside their arrays.
template <typename R>
void validate_program(pasif_driver* d, const HostProgram<R>& hp, const std::vector<TileHeader>& hdr,
                      bool adjoint_side, uint32_t L, const SlotConfig& c, uint32_t ss,
                      uint32_t xs, size_t n_init, size_t out_len, uint32_t n_exc) {
  const int CW = slot_words(c);
  const uint32_t smax = L * ss, xmax = L * xs;
  const size_t ntiles = hdr.size();
  for (size_t t = 0; t < ntiles; ++t) {
    const uint32_t rows = adjoint_side ? hdr[t].rows_a : hdr[t].rows_f;
    PASIF_REQUIRE(rows <= L, PASIF_ERR_UNSUPPORTED, "tile %zu: %u rows > %u lanes", t, rows, L);
    for (uint32_t l = 0; l < L; ++l) {
      const size_t i = t * L + l;
      const uint32_t w = hp.info[i];
      const uint32_t row = lane_row(w);
      if (row == kRowIdle) continue;
      PASIF_REQUIRE(row < rows, PASIF_ERR_UNSUPPORTED, "tile %zu lane %u: row %u of %u", t, l,
                    row, rows);
      if (lane_owner(w)) {
        PASIF_REQUIRE(lane_nseg(w) >= 1 && l + lane_nseg(w) <= L, PASIF_ERR_UNSUPPORTED,
                      "tile %zu lane %u: %u segments overflow the tile", t, l, lane_nseg(w));
        PASIF_REQUIRE(hp.src[i] < n_init, PASIF_ERR_UNSUPPORTED, "initial index %u of %zu",
                      hp.src[i], n_init);
        PASIF_REQUIRE(hp.out[i] < out_len, PASIF_ERR_UNSUPPORTED, "output index %u of %zu",
                      hp.out[i], out_len);
        PASIF_REQUIRE(hp.exc[i] < int32_t(n_exc), PASIF_ERR_UNSUPPORTED, "excitation %d of %u",
                      hp.exc[i], n_exc);
        PASIF_REQUIRE(hp.exc[i] < 0 || hp.fslot[i] < hdr[t].nforced, PASIF_ERR_UNSUPPORTED,
                      "tile %zu lane %u: sample slot %u of %u", t, l, hp.fslot[i],
                      hdr[t].nforced);
      }
    }
    for (int wd = 0; wd < CW; ++wd) {
      const int rel = wd - (c.c1 + c.c2);
      const bool second = rel >= 0 && (rel % 2) == 1;
      for (uint32_t l = 0; l < L; ++l) {
        const uint32_t w = hp.idx[(t * CW + wd) * L + l];
        const uint32_t lo = w & 0xFFFFu, hi = w >> 16;
        const bool ok = second ? (lo <= xmax && hi <= xmax) : (lo <= smax && hi <= xmax);
        PASIF_REQUIRE(ok && (lo % sizeof(R)) == 0 && (hi % sizeof(R)) == 0, PASIF_ERR_UNSUPPORTED,
                      "tile %zu word %d lane %u: LDS offsets %u/%u beyond %u/%u", t, wd, l, lo, hi,
                      smax, xmax);
      }
    }
  }
}

// Instances of the problem; `part` keeps those of the lane kernel's systems
// (kPartLane), of the tile kernel's (kPartTile) or all (kPartAll).
enum { kPartAll = 0, kPartLane = 1, kPartTile = 2, kPartBig = 3 };
bool in_part(const pasif_driver* d, uint32_t s, int part) {
  if (part == kPartAll) return true;
  const bool big = s < d->big_sys.size() && d->big_sys[s] != 0;
  if (part == kPartBig || big) return part == kPartBig && big;
  const bool lane = s < d->lane_cls.size() && d->lane_cls[s] >= 0;
  return part == kPartLane ? lane : !lane;
}
// lane_cls[s] = (lane part << 16) | class index within the part, -1: tiles
inline int32_t lane_code(int part, uint32_t cls) { return int32_t((uint32_t(part) << 16) | cls); }
inline int lane_part_of(int32_t code) { return int(uint32_t(code) >> 16); }
inline uint32_t lane_class_of(int32_t code) { return uint32_t(code) & 0xFFFFu; }

std::vector<Instance> make_instances(pasif_driver* d, bool adjoint, int part = kPartAll) {
  std::vector<Instance> v;
  const uint32_t S = d->fwd.size.size();
  const uint32_t N = d->fwd.n, Na = adjoint ? d->adj.n : 0;
  if (d->mapped) {
    v.reserve(S);
    for (uint32_t s = 0; s < S; ++s) {
      if (!in_part(d, s, part)) continue;
      Instance in{};
      in.sys = s;
      in.exc = d->map_exc.empty() ? 0 : d->map_exc[s];
      PASIF_REQUIRE(in.exc < (int32_t)d->n_exc, PASIF_ERR_ARGUMENT,
                    "system %u maps to excitation %d of %u", s, in.exc, d->n_exc);
      in.steps = d->map_steps.empty() ? d->numsteps : d->map_steps[s];
      in.out_f = d->fwd.off[s];
      in.out_a = adjoint ? d->adj.off[s] : 0;
      v.push_back(in);
    }
  } else {
    v.reserve(size_t(S) * d->n_exc);
    for (uint32_t p = 0; p < d->n_exc; ++p)
      for (uint32_t s = 0; s < S; ++s) {
        if (!in_part(d, s, part)) continue;
        Instance in{};
        in.sys = s;
        in.exc = int32_t(p);
        in.steps = d->map_steps.empty() ? d->numsteps : d->map_steps[s];
        in.out_f = p * N + d->fwd.off[s];
        in.out_a = adjoint ? p * Na + d->adj.off[s] : 0;
        v.push_back(in);
      }
  }
  return v;
}

template <typename R>
void pack_impl(pasif_driver* d, bool adjoint) {
  Packed& pk = d->pk;
  const uint32_t S = d->fwd.size.size();
  std::vector<Instance> inst = make_instances(d, adjoint, kPartTile);
  // per-system instance counts (Cartesian: every excitation)
  std::vector<uint64_t> mult(S, 0);
  for (const Instance& in : inst) mult[in.sys]++;
  // kinds in use (a kind no row uses is skipped by the kernel)
  bool kind_used[4] = {false, false, false, false};
  for (int k = 0; k < 4; ++k) {
    kind_used[k] = !d->fwd.val[k].empty();
    if (adjoint) kind_used[k] = kind_used[k] || !d->adj.val[k].empty();
  }
  // Slot configuration: the given one, or the one with the fewest lane-slot
  // LDS reads per stage (lanes x sum over used kinds of slots x factors).
  const uint32_t cands[3] = {3, 1, 2};
  uint32_t cfg = d->slot_cfg;
  SlotConfig c{};
  std::vector<uint32_t> lanes_f(S), lanes_a(S, 0);
  double best = -1.0;
  for (uint32_t cand : cands) {
    if (d->slot_cfg && cand != d->slot_cfg) continue;
    const SlotConfig cc = slot_of(cand);
    std::vector<uint32_t> lf(S), la(S, 0);
    uint32_t maxl = 0;
    uint64_t total = 0;
    bool ok = true;
    for (uint32_t s = 0; s < S && ok; ++s) {
      if (!in_part(d, s, kPartTile)) continue;
      lf[s] = system_lanes(d->fwd, s, cc);
      if (adjoint) la[s] = system_lanes(d->adj, s, cc);
      if (lf[s] == 0 || (adjoint && la[s] == 0)) ok = false;
      const uint32_t l = std::max(lf[s], la[s]);
      maxl = std::max(maxl, l);
      total += uint64_t(l) * mult[s];
    }
    if (!ok || maxl > kMaxLanes || (maxl > 256 && cand != 1)) continue;
    const int cap[4] = {cc.c1, cc.c2, cc.c3, cc.c4};
    double reads = 0;
    for (int k = 0; k < 4; ++k)
      if (kind_used[k]) reads += cap[k] * (k + 1);
    const double cost = double(total) * reads;
    if (best < 0 || cost < best) {
      best = cost;
      cfg = cand;
      c = cc;
      lanes_f = lf;
      lanes_a = la;
    }
  }
  PASIF_REQUIRE(best >= 0, PASIF_ERR_UNSUPPORTED,
                "no slot configuration fits these systems (a system needs more than %u lanes, or "
                "slot configuration %u cannot hold it)",
                kMaxLanes, d->slot_cfg);
  uint32_t maxl = 0;
  for (uint32_t s = 0; s < S; ++s) maxl = std::max(maxl, std::max(lanes_f[s], lanes_a[s]));
  uint64_t total_lanes = 0;
  for (Instance& in : inst) {
    in.lanes = std::max(lanes_f[in.sys], lanes_a[in.sys]);
    total_lanes += in.lanes;
  }
  uint32_t L = d->tile_lanes;
  if (L == 0) {
    const uint64_t per = (total_lanes + 511) / 512;  // aim for >= 512 tiles
    L = uint32_t(std::min<uint64_t>(256, std::max<uint64_t>(64, ((per + 63) / 64) * 64)));
  }
  L = std::max(L, ((maxl + 63) / 64) * 64);
  PASIF_REQUIRE(L <= kMaxLanes, PASIF_ERR_UNSUPPORTED, "tile of %u lanes", L);
  // longest instances first (their tiles start first), stable otherwise
  std::stable_sort(inst.begin(), inst.end(),
                   [](const Instance& a, const Instance& b) { return a.steps > b.steps; });
  // greedy fill
  std::vector<uint32_t> tile_of(inst.size());
  uint32_t ntiles = 0, used = L;
  for (size_t i = 0; i < inst.size(); ++i) {
    if (used + inst[i].lanes > L) {
      ++ntiles;
      used = 0;
    }
    tile_of[i] = ntiles - 1;
    used += inst[i].lanes;
  }
  // LDS row strides: stage buffer [row][2], adjoint set points [row][4]
  const uint32_t ss = 2 * sizeof(R), xs_f = 2 * sizeof(R), xs_a = 4 * sizeof(R);
  HostProgram<R> hf, ha;
  hf.resize(ntiles, L, c, L * ss, L * xs_f);
  if (adjoint) ha.resize(ntiles, L, c, L * ss, L * xs_a);
  std::vector<TileHeader> hdr(ntiles);
  std::memset(hdr.data(), 0, hdr.size() * sizeof(TileHeader));
  std::vector<uint32_t> lane_used(ntiles, 0), rf(ntiles, 0), ra(ntiles, 0);
  std::vector<bool> split_f(ntiles, false), split_a(ntiles, false);
  uint32_t tmax = 0;
  for (size_t i = 0; i < inst.size(); ++i) {
    const Instance& in = inst[i];
    const uint32_t t = tile_of[i];
    TileHeader& H = hdr[t];
    bool sf = split_f[t];
    place(d->fwd, in.sys, in, false, hf, t, L, lane_used[t], rf[t], rf[t], c, H.cnt_f, sf,
          d->fwd.off[in.sys], ss, xs_f, H.nforced);
    split_f[t] = sf;
    if (adjoint) {
      bool sa = split_a[t];
      uint32_t no_forcing = 0;
      place(d->adj, in.sys, in, true, ha, t, L, lane_used[t], ra[t], rf[t], c, H.cnt_a, sa,
            d->adj.xoff[in.sys], ss, xs_a, no_forcing);
      split_a[t] = sa;
      ra[t] += d->adj.size[in.sys];
    }
    rf[t] += d->fwd.size[in.sys];
    lane_used[t] += in.lanes;
    H.steps = std::max(H.steps, in.steps);
    tmax = std::max(tmax, in.steps);
  }
  uint32_t max_forced = 0;
  for (uint32_t t = 0; t < ntiles; ++t) {
    max_forced = std::max(max_forced, hdr[t].nforced);
    hdr[t].rows_f = rf[t];
    hdr[t].rows_a = ra[t];
    hdr[t].split_f = split_f[t] ? 1 : 0;
    hdr[t].split_a = split_a[t] ? 1 : 0;
  }
  const uint32_t out_len_f = d->mapped ? d->fwd.n : d->fwd.n * d->n_exc;
  const uint32_t out_len_a = adjoint ? (d->mapped ? d->adj.n : d->adj.n * d->n_exc) : 0;
  validate_program(d, hf, hdr, false, L, c, ss, xs_f, d->fwd.init.size(), out_len_f, d->n_exc);
  {
    uint32_t mf = 0;
    for (const TileHeader& h : hdr) mf = std::max(mf, h.nforced);
    const size_t lds = (2 * (L + 1) + L + (adjoint ? 4 * (L + 1) : 0) + 2 * kMaxInterpWindow +
                        size_t(mf) * kForceStage) * sizeof(R);
    PASIF_REQUIRE(lds <= kMaxLdsBytes, PASIF_ERR_UNSUPPORTED,
                  "tiles of %u lanes with %u forced rows need %zu bytes of LDS (> %zu); use "
                  "smaller tiles",
                  L, mf, lds, kMaxLdsBytes);
  }
  if (adjoint)
    validate_program(d, ha, hdr, true, L, c, ss, xs_a, d->adj.init.size(), out_len_a, d->n_exc);
  // initial conditions in the working precision
  std::vector<R> initf(d->fwd.init.begin(), d->fwd.init.end());
  upload(d, pk.hdr, hdr);
  upload(d, pk.f_info, hf.info);
  upload(d, pk.f_out, hf.out);
  upload(d, pk.f_src, hf.src);
  upload(d, pk.f_steps, hf.steps);
  upload(d, pk.f_exc, hf.exc);
  upload(d, pk.f_fslot, hf.fslot);
  upload(d, pk.f_force, hf.force);
  upload(d, pk.f_val, hf.val);
  upload(d, pk.f_idx, hf.idx);
  upload(d, pk.f_init, initf);
  pk.program_bytes = pk.f_info.bytes * 4 + pk.f_force.bytes * 2 + pk.f_val.bytes + pk.f_idx.bytes;
  if (adjoint) {
    // checkpoint blocks sized by each tile's own longest instance
    uint32_t K = d->ck_every ? d->ck_every : 256;
    K = std::max<uint32_t>(1, std::min<uint32_t>(K, std::max<uint32_t>(tmax, 1)));
    std::vector<uint32_t> ck_tile(ntiles);
    uint64_t slots = 0;
    for (uint32_t t = 0; t < ntiles; ++t) {
      ck_tile[t] = uint32_t(slots);
      slots += std::max<uint32_t>(1, (hdr[t].steps + K - 1) / K);
    }
    PASIF_REQUIRE(slots <= 0xFFFFFFFFull, PASIF_ERR_UNSUPPORTED,
                  "%llu checkpoint slots (more than 2^32): use a longer checkpoint interval",
                  (unsigned long long)slots);
    upload(d, pk.ck_tile, ck_tile);
    pk.ck_every = K;
    pk.ck_slots = slots;
    std::vector<R> inita(d->adj.init.begin(), d->adj.init.end());
    upload(d, pk.a_info, ha.info);
    upload(d, pk.a_out, ha.out);
    upload(d, pk.a_src, ha.src);
    upload(d, pk.a_steps, ha.steps);
    upload(d, pk.a_exc, ha.exc);
    upload(d, pk.a_fslot, ha.fslot);
    upload(d, pk.a_force, ha.force);
    upload(d, pk.a_val, ha.val);
    upload(d, pk.a_idx, ha.idx);
    upload(d, pk.a_init, inita);
    pk.program_bytes += pk.a_info.bytes * 4 + pk.a_force.bytes * 2 + pk.a_val.bytes +
                        pk.a_idx.bytes;
  }
  if (d->dev >= 0) PASIF_HIP(hipStreamSynchronize(d->stream));
  pk.lanes = L;
  pk.ntiles = ntiles;
  pk.slot_cfg = cfg;
  pk.adjoint = adjoint;
  pk.precision = int(sizeof(R) * 8);
  pk.tmax = tmax;
  pk.max_forced = max_forced;
  pk.out_len_f = out_len_f;
  pk.out_len_a = out_len_a;
  pk.valid = true;
  if (d->d_compute)
    printf("[pasif] %zu instances in %u tiles of %u lanes (slot config %u), T max %u, %.1f MB "
           "program\n",
           inst.size(), ntiles, L, cfg, tmax,
           pk.program_bytes / 1048576.0);
}

// Row-parallel packing: every instance with its own copy of its system's
// terms (local rows, kinds in order per row), force and initial values in
// the working precision, and its scratch offsets.
template <typename R>
void pack_big(pasif_driver* d, bool adjoint) {
  Packed& pk = d->pk;
  std::vector<Instance> inst = make_instances(d, adjoint, kPartBig);
  const double eps = std::numeric_limits<R>::epsilon();
  struct Side {
    std::vector<uint32_t> rptr, idx, seg, sptr;
    std::vector<R> val, force, init;
  };
  Side sf, sa;
  const uint32_t K0 = d->ck_every ? d->ck_every : 256;
  uint64_t vf = 0, va = 0, ck = 0, sg = 0;
  uint32_t tmax = 0, K = 1;
  for (const Instance& in : inst) tmax = std::max(tmax, in.steps);
  if (adjoint) K = std::max<uint32_t>(1, std::min<uint32_t>(K0, std::max<uint32_t>(tmax, 1)));
  std::vector<BigInst> bi(inst.size());
  // every (row, kind) also as segments of at most kBigSegTerms terms, so a
  // row with many terms is summed by several threads
  auto add_side = [&](const RowTerms& rt, uint32_t s, bool adj, const Instance& in, Side& out,
                      uint64_t& rptr_off, uint64_t& term_off, uint64_t& row_off,
                      uint64_t& sptr_off, uint64_t& seg_off, uint32_t& nseg, uint32_t& split) {
    const uint32_t off = rt.off[s], n = rt.size[s];
    const uint32_t xo = adj ? rt.xoff[s] : off;
    rptr_off = out.rptr.size();
    term_off = out.val.size();
    row_off = out.init.size();
    sptr_off = out.sptr.size();
    seg_off = out.seg.size() / 3;
    nseg = 0;
    split = 0;
    uint32_t t = 0;
    for (uint32_t r = 0; r < n; ++r) {
      const uint32_t g = off + r;
      for (int k = 0; k < 4; ++k) {
        out.rptr.push_back(t);
        out.sptr.push_back(nseg);
        const uint32_t cnt = rt.ptr[k][g + 1] - rt.ptr[k][g];
        if (cnt > kBigSegTerms) split = 1;
        for (uint32_t b = 0; b < cnt; b += kBigSegTerms) {
          out.seg.push_back(t + b);
          out.seg.push_back(t + std::min(cnt, b + kBigSegTerms));
          out.seg.push_back(uint32_t(k));
          ++nseg;
        }
        const int nf = k + 1;
        for (uint32_t e = rt.ptr[k][g]; e < rt.ptr[k][g + 1]; ++e, ++t) {
          out.val.push_back(R(rt.val[k][e]));
          uint32_t w[4] = {0, 0, 0, 0};
          w[0] = rt.idx[k][size_t(e) * nf] - off;
          for (int a = 1; a < nf; ++a) w[a] = rt.idx[k][size_t(e) * nf + a] - xo;
          for (int a = 0; a < 4; ++a) out.idx.push_back(w[a]);
        }
      }
      const double f = adj ? 0.0 : rt.force[g];
      out.force.push_back((std::fabs(f) > eps && in.exc >= 0) ? R(f) : R(0));
      out.init.push_back(R(rt.init[g]));
    }
    out.rptr.push_back(t);
    out.sptr.push_back(nseg);
  };
  uint64_t part = 0;
  for (size_t i = 0; i < inst.size(); ++i) {
    const Instance& in = inst[i];
    BigInst& b = bi[i];
    add_side(d->fwd, in.sys, false, in, sf, b.rptr_f, b.term_f, b.row_f, b.sptr_f, b.sseg_f,
             b.nseg_f, b.split_f);
    b.part_f = part;
    part += b.nseg_f;
    pk.b_part_max = std::max(pk.b_part_max, b.nseg_f);
    b.n_f = d->fwd.size[in.sys];
    b.steps = in.steps;
    b.exc = in.exc;
    b.out_f = in.out_f;
    b.scr_f = vf;
    vf += 4ull * b.n_f;
    if (adjoint) {
      add_side(d->adj, in.sys, true, in, sa, b.rptr_a, b.term_a, b.row_a, b.sptr_a, b.sseg_a,
               b.nseg_a, b.split_a);
      b.part_a = part;
      part += b.nseg_a;
      pk.b_part_max = std::max(pk.b_part_max, b.nseg_a);
      b.n_a = d->adj.size[in.sys];
      b.out_a = in.out_a;
      b.scr_a = va;
      va += 4ull * b.n_a;
      b.ck = ck;
      ck += uint64_t(std::max<uint32_t>(1, (in.steps + K - 1) / K)) * b.n_f;
      b.seg = sg;
      sg += uint64_t(d->adj_mode ? 4ull * K : K + 1ull) * b.n_f;
    }
  }
  PASIF_REQUIRE(sf.val.size() < 0xFFFFFFFFull && sa.val.size() < 0xFFFFFFFFull,
                PASIF_ERR_UNSUPPORTED, "more than 2^32 terms");
  upload(d, pk.b_inst, bi);
  upload(d, pk.b_rf, sf.rptr);
  upload(d, pk.b_vf, sf.val);
  upload(d, pk.b_xf, sf.idx);
  upload(d, pk.b_ff, sf.force);
  upload(d, pk.b_if, sf.init);
  upload(d, pk.b_sgf, sf.seg);
  upload(d, pk.b_spf, sf.sptr);
  pk.b_part = part;
  if (adjoint) {
    upload(d, pk.b_ra, sa.rptr);
    upload(d, pk.b_va, sa.val);
    upload(d, pk.b_xa, sa.idx);
    upload(d, pk.b_ia, sa.init);
    upload(d, pk.b_sga, sa.seg);
    upload(d, pk.b_spa, sa.sptr);
  }
  if (d->dev >= 0) PASIF_HIP(hipStreamSynchronize(d->stream));
  pk.big = true;
  pk.nbig = uint32_t(inst.size());
  pk.b_vec_f = vf;
  pk.b_vec_a = va;
  pk.b_ck = ck;
  pk.b_seg = sg;
  pk.b_tmax = tmax;
  pk.b_every = K;
  pk.out_len_f = d->mapped ? d->fwd.n : d->fwd.n * d->n_exc;
  pk.out_len_a = adjoint ? (d->mapped ? d->adj.n : d->adj.n * d->n_exc) : 0;
  if (d->d_compute)
    printf("[pasif] %zu instances on the row-parallel path, T max %u\n", inst.size(), tmax);
}

// Whether system s fits one tile under some slot configuration the tile
// packer may choose (the same test pack_impl applies to every system).
bool fits_tile(const pasif_driver* d, uint32_t s, bool adjoint) {
  const uint32_t cands[3] = {3, 1, 2};
  for (uint32_t cand : cands) {
    if (d->slot_cfg && cand != d->slot_cfg) continue;
    const SlotConfig c = slot_of(cand);
    const uint32_t lf = system_lanes(d->fwd, s, c);
    const uint32_t la = adjoint ? system_lanes(d->adj, s, c) : 0u;
    if (lf == 0 || (adjoint && la == 0)) continue;
    const uint32_t l = std::max(lf, la);
    if (l <= kMaxLanes && (l <= 256 || cand == 1)) return true;
  }
  return false;
}

// 0 for an automatic group size, log2(lanes) + 1 otherwise (3 bits)
uint32_t lane_group_code(uint32_t lanes) {
  uint32_t c = 0;
  while (lanes) {
    ++c;
    lanes >>= 1;
  }
  return c;
}

uint64_t pack_key(pasif_driver* d, bool adjoint) {
  return (uint64_t(d->precision) << 1) ^ (uint64_t(adjoint) << 8) ^
         (uint64_t(adjoint && d->adj_mode) << 9) ^ (uint64_t(d->tile_lanes) << 10) ^
         (uint64_t(d->slot_cfg) << 24) ^ (uint64_t(d->kernel_path) << 28) ^
         (uint64_t(d->ck_every) << 32) ^ (uint64_t(lane_group_code(d->lane_group)) << 56) ^
         (uint64_t(uint32_t(d->lane_bias + 8) & 15u) << 59);
}

// ------------------------------------------------------------ lane plan
// Largest state + coefficient count (values of the working type) a class may
// keep live: 2 VGPRs per FP64 value, leaving room for temporaries in 256.
constexpr uint32_t kLaneMaxLive = 96;
uint32_t lane_max_live() {
  if (const char* e = std::getenv("PASIF_LANE_MAX_LIVE")) {
    const long v = std::strtol(e, nullptr, 10);
    if (v > 0 && v < 100000) return uint32_t(v);
  }
  return kLaneMaxLive;
}
// structure classes the lane kernel takes (the rest run on tiles), and per
// module (one generated kernel)
constexpr uint32_t kLaneMaxClasses = 64;
constexpr uint32_t kLaneModuleClasses = 8;
// Multi-wave classes compile one kernel body per wave (about 5 s of hipRTC
// time each): the lane kernel takes them, heaviest first, up to this many
// bodies in all; the rest run on tiles.
constexpr uint32_t kLaneMultiWaveBodies = 16;
// multi-wave: waves per workgroup and values one wave's share may keep live
constexpr uint32_t kLaneMaxWaves = 16;
constexpr uint32_t kLaneMaxWaveLive = 100;
constexpr size_t kLaneMaxLds = 64 * 1024;

// Terms of row r of system s with identical monomials merged: for the
// forward side every factor is the state, so a term is identified by its
// kind and its sorted factor rows (Gamma[o,j,k] and Gamma[o,k,j] are one
// monomial); for the adjoint side the own factor (lambda) stays apart from
// the set-point factors, which are sorted.  `which[e]` is the merged term of
// the row's e-th entry (entries in kind order, then entry order).  Rows have
// few terms: a linear scan, no allocation beyond the reused outputs.
void merged_row(const RowTerms& rt, uint32_t s, uint32_t r, bool adjoint,
                std::vector<LaneTerm>& out, std::vector<uint32_t>& which) {
  out.clear();
  which.clear();
  const uint32_t off = rt.off[s];
  const uint32_t xo = adjoint ? rt.xoff[s] : off;
  const uint32_t g = off + r;
  for (int k = 0; k < 4; ++k) {
    const int nf = k + 1;
    for (uint32_t e = rt.ptr[k][g]; e < rt.ptr[k][g + 1]; ++e) {
      LaneTerm t{};
      t.kind = uint32_t(k);
      t.f[0] = rt.idx[k][size_t(e) * nf] - off;
      for (int a = 1; a < nf; ++a) t.f[a] = rt.idx[k][size_t(e) * nf + a] - xo;
      if (!adjoint) std::sort(t.f, t.f + nf);
      else std::sort(t.f + 1, t.f + nf);
      uint32_t m = 0;
      for (; m < out.size(); ++m) {
        const LaneTerm& o = out[m];
        if (o.kind != t.kind) continue;
        bool same = true;
        for (int a = 0; a < nf && same; ++a) same = o.f[a] == t.f[a];
        if (same) break;
      }
      if (m == out.size()) out.push_back(t);
      which.push_back(m);
    }
  }
}

// Structure of system s: its merged terms in coefficient order with local
// factor rows, and (forward) the rows with a force above the working type's
// eps.
LaneSide lane_side(const RowTerms& rt, uint32_t s, bool adjoint, double eps) {
  LaneSide sd;
  const uint32_t off = rt.off[s], n = rt.size[s];
  sd.n = n;
  sd.rows.resize(n);
  std::vector<LaneTerm> mt;
  std::vector<uint32_t> which;
  for (uint32_t r = 0; r < n; ++r) {
    merged_row(rt, s, r, adjoint, mt, which);
    sd.rows[r] = mt;
    if (!adjoint && std::fabs(rt.force[off + r]) > eps) sd.forced.push_back(r);
  }
  return sd;
}

void shape_key(const LaneSide& sd, std::vector<uint32_t>& key) {
  key.push_back(sd.n);
  for (const auto& row : sd.rows) {
    key.push_back(uint32_t(row.size()));
    for (const LaneTerm& t : row) {
      key.push_back(t.kind);
      for (uint32_t a = 0; a <= t.kind; ++a) key.push_back(t.f[a]);
    }
  }
  key.push_back(uint32_t(sd.forced.size()));
  for (uint32_t r : sd.forced) key.push_back(r);
}

bool same_term(const LaneTerm& x, const LaneTerm& y) {
  if (x.kind != y.kind) return false;
  for (uint32_t a = 0; a <= x.kind; ++a)
    if (x.f[a] != y.f[a]) return false;
  return true;
}

// Position of merged term t in a class row (the class may be the union of
// several structures: a system fills its own terms, the others stay zero).
uint32_t term_slot(const std::vector<LaneTerm>& row, const LaneTerm& t) {
  for (uint32_t m = 0; m < row.size(); ++m)
    if (same_term(row[m], t)) return m;
  throw Fail{PASIF_ERR_UNSUPPORTED};
}

// The union of two structures of one size: every row's terms of `a` in
// order, then those of `b` that `a` lacks; the rows forced in either.
LaneSide side_union(const LaneSide& a, const LaneSide& b) {
  LaneSide u = a;
  u.n = std::max(a.n, b.n);
  u.rows.resize(u.n);
  for (uint32_t r = 0; r < b.n; ++r)
    for (const LaneTerm& t : b.rows[r]) {
      bool have = false;
      for (const LaneTerm& x : u.rows[r]) have = have || same_term(x, t);
      if (!have) u.rows[r].push_back(t);
    }
  std::vector<uint32_t> f;
  std::set_union(a.forced.begin(), a.forced.end(), b.forced.begin(), b.forced.end(),
                 std::back_inserter(f));
  u.forced = f;
  return u;
}

// Coefficients of system s in the class's slot order (merged terms row by
// row, each the sum of its entries in entry order, at its place in the
// class row, zero for class terms the system lacks; then the force of each
// forced row of the class, zero where this system's is below eps).
void lane_coefs(const RowTerms& rt, uint32_t s, const LaneSide& sd, bool adjoint, double eps,
                std::vector<double>& out) {
  out.clear();
  const uint32_t off = rt.off[s];
  const uint32_t n = rt.size[s];
  if (n > sd.n) throw Fail{PASIF_ERR_UNSUPPORTED};  // not a system of this class
  std::vector<LaneTerm> mt;
  std::vector<uint32_t> which;
  std::vector<double> v;
  for (uint32_t r = 0; r < n; ++r) {
    merged_row(rt, s, r, adjoint, mt, which);
    v.assign(mt.size(), 0.0);
    // merged terms appear in the order of their first entries
    uint32_t q = 0, fresh = 0;
    const uint32_t g = off + r;
    for (int k = 0; k < 4; ++k)
      for (uint32_t e = rt.ptr[k][g]; e < rt.ptr[k][g + 1]; ++e, ++q) {
        const uint32_t m = which[q];
        if (m == fresh) {
          v[m] = rt.val[k][e];
          ++fresh;
        } else {
          v[m] += rt.val[k][e];
        }
      }
    const size_t base = out.size();
    out.resize(base + sd.rows[r].size(), 0.0);
    for (uint32_t m = 0; m < mt.size(); ++m) out[base + term_slot(sd.rows[r], mt[m])] = v[m];
  }
  // rows a padded adjoint class has beyond this system's
  for (uint32_t r = n; r < sd.n; ++r) out.resize(out.size() + sd.rows[r].size(), 0.0);
  for (uint32_t r : sd.forced) {
    const double f = rt.force[off + r];
    out.push_back(std::fabs(f) > eps ? f : 0.0);
  }
}

// lane_coefs through a merge map cached from one system of a class: the raw
// entries of every row and kind (factor rows relative to the system) and the
// merged slot each entry adds into.  A system whose raw entries are the same
// takes its coefficients by one pass over its values (same slots, same
// summation order as lane_coefs); any other goes through lane_coefs.
struct CoefMap {
  bool valid = false;
  uint32_t n = 0, nslots = 0;
  std::vector<uint32_t> cnt;    // [n][4] entries per row and kind
  std::vector<uint32_t> rel;    // every entry's factor rows, relative, in order
  std::vector<uint32_t> slot;   // every entry's merged slot
  std::vector<uint8_t> first;   // 1: the first entry of its slot (assigned, not added)
};

void lane_coefs_mapped(const RowTerms& rt, uint32_t s, const LaneSide& sd, bool adjoint,
                       double eps, CoefMap& cm, std::vector<double>& out) {
  const uint32_t off = rt.off[s];
  const uint32_t xo = adjoint ? rt.xoff[s] : off;
  const uint32_t n = rt.size[s];
  if (!cm.valid) {
    std::vector<LaneTerm> mt;
    std::vector<uint32_t> which;
    cm.n = n;
    cm.cnt.assign(size_t(n) * 4, 0);
    cm.rel.clear();
    cm.slot.clear();
    cm.first.clear();
    if (n > sd.n) throw Fail{PASIF_ERR_UNSUPPORTED};  // not a system of this class
    uint32_t base = 0;
    std::vector<uint32_t> pos;
    for (uint32_t r = 0; r < n; ++r) {
      merged_row(rt, s, r, adjoint, mt, which);
      pos.resize(mt.size());
      for (uint32_t m = 0; m < mt.size(); ++m) pos[m] = term_slot(sd.rows[r], mt[m]);
      const uint32_t g = off + r;
      uint32_t q = 0, fresh = 0;
      for (int k = 0; k < 4; ++k) {
        const int nf = k + 1;
        cm.cnt[size_t(r) * 4 + k] = rt.ptr[k][g + 1] - rt.ptr[k][g];
        for (uint32_t e = rt.ptr[k][g]; e < rt.ptr[k][g + 1]; ++e, ++q) {
          cm.rel.push_back(rt.idx[k][size_t(e) * nf] - off);
          for (int a = 1; a < nf; ++a) cm.rel.push_back(rt.idx[k][size_t(e) * nf + a] - xo);
          cm.slot.push_back(base + pos[which[q]]);
          cm.first.push_back(which[q] == fresh ? 1 : 0);
          if (which[q] == fresh) ++fresh;
        }
      }
      base += uint32_t(sd.rows[r].size());
    }
    for (uint32_t r = n; r < sd.n; ++r) base += uint32_t(sd.rows[r].size());
    cm.nslots = base;
    cm.valid = true;
  } else {
    // the same raw entries as the cached system?
    bool same = n == cm.n;
    size_t x = 0;
    for (uint32_t r = 0; r < n && same; ++r) {
      const uint32_t g = off + r;
      for (int k = 0; k < 4 && same; ++k) {
        const int nf = k + 1;
        if (rt.ptr[k][g + 1] - rt.ptr[k][g] != cm.cnt[size_t(r) * 4 + k]) {
          same = false;
          break;
        }
        for (uint32_t e = rt.ptr[k][g]; e < rt.ptr[k][g + 1] && same; ++e) {
          same = rt.idx[k][size_t(e) * nf] - off == cm.rel[x++];
          for (int a = 1; a < nf && same; ++a)
            same = rt.idx[k][size_t(e) * nf + a] - xo == cm.rel[x++];
        }
      }
    }
    if (!same) {
      lane_coefs(rt, s, sd, adjoint, eps, out);
      return;
    }
  }
  out.assign(cm.nslots, 0.0);
  size_t q = 0;
  for (uint32_t r = 0; r < n; ++r) {
    const uint32_t g = off + r;
    for (int k = 0; k < 4; ++k)
      for (uint32_t e = rt.ptr[k][g]; e < rt.ptr[k][g + 1]; ++e, ++q) {
        if (cm.first[q])
          out[cm.slot[q]] = rt.val[k][e];
        else
          out[cm.slot[q]] += rt.val[k][e];
      }
  }
  for (uint32_t r : sd.forced) {
    const double f = rt.force[off + r];
    out.push_back(std::fabs(f) > eps ? f : 0.0);
  }
}

// Lanes per instance (powers of two up to a wave) at which a class fits the
// register budget, with the step cost of each.
std::vector<std::pair<uint32_t, double>> group_options(const LaneShape& sh, int bits) {
  std::vector<std::pair<uint32_t, double>> out;
  for (uint32_t G = 1; G <= kLaneWave; G *= 2) {
    LaneShape t = sh;
    if (!lane_group(t, G)) continue;
    if (lane_live(t) > lane_max_live()) continue;
    out.push_back({G, lane_step_cost(t, bits)});
  }
  return out;
}

struct ClassLoad {
  double inst = 0, steps_sum = 0, steps_max = 0;
};

// FNV-1a over 32-bit words
struct Sig {
  uint64_t h = 1469598103934665603ull;
  void word(uint64_t w) {
    for (int b = 0; b < 8; ++b) {
      h ^= (w >> (8 * b)) & 0xFFu;
      h *= 1099511628211ull;
    }
  }
  // eight bytes at a time (a signature, not a checksum of the bytes' order
  // within a word)
  template <typename T>
  void vec(const std::vector<T>& v) {
    word(v.size());
    const size_t nb = v.size() * sizeof(T);
    const unsigned char* p = reinterpret_cast<const unsigned char*>(v.data());
    size_t i = 0;
    for (; i + 8 <= nb; i += 8) {
      uint64_t w;
      std::memcpy(&w, p + i, 8);
      h = (h ^ w) * 1099511628211ull;
      h ^= h >> 29;
    }
    for (; i < nb; ++i) h = (h ^ p[i]) * 1099511628211ull;
  }
};

// What a lane plan depends on besides coefficient values and initial
// states: the systems' structure (term indices, sizes, force pattern), the
// instances (excitations, step counts), the settings and the device.
uint64_t lane_plan_sig(const pasif_driver* d, bool adjoint, int bits) {
  Sig g;
  g.word(pack_key(const_cast<pasif_driver*>(d), adjoint));
  g.word(uint64_t(bits));
  g.word(uint64_t(int64_t(d->dev)));
  g.word(lane_max_live());
  if (const char* f = std::getenv("PASIF_JIT_FLAGS"))
    for (const char* c = f; *c; ++c) g.word(uint64_t(uint8_t(*c)));
  auto rows = [&](const RowTerms& rt) {
    g.word(rt.n);
    g.vec(rt.off);
    g.vec(rt.size);
    g.vec(rt.xoff);
    for (int k = 0; k < 4; ++k) {
      g.vec(rt.ptr[k]);
      g.vec(rt.idx[k]);
    }
  };
  rows(d->fwd);
  g.vec(d->fwd.force);
  if (adjoint) rows(d->adj);
  g.word(d->mapped ? 1 : 0);
  g.vec(d->map_exc);
  g.vec(d->map_steps);
  g.word(d->n_exc);
  g.word(d->numsteps);
  return g.h;
}

// Per-step cost of a class (coefficient slots and state values of both
// sides), the unit of the merge decisions below.
double class_cost(const LaneShape& sh) {
  double c = sh.f.coefs() + 2.0 * sh.f.n;
  if (sh.adjoint) c += sh.a.coefs() + 2.0 * sh.a.n + sh.f.n;
  return c;
}

// Coefficient slots of the union of two sides of one size.
uint32_t union_coefs(const LaneSide& a, const LaneSide& b) {
  size_t t = 0;
  for (uint32_t r = a.n; r < b.n; ++r) t += b.rows[r].size();
  for (uint32_t r = 0; r < a.n; ++r) {
    t += a.rows[r].size();
    if (r >= b.n) continue;
    for (const LaneTerm& x : b.rows[r]) {
      bool have = false;
      for (const LaneTerm& y : a.rows[r]) have = have || same_term(x, y);
      t += have ? 0 : 1;
    }
  }
  size_t i = 0, j = 0, f = 0;
  while (i < a.forced.size() || j < b.forced.size()) {
    if (j == b.forced.size() || (i < a.forced.size() && a.forced[i] < b.forced[j])) ++i;
    else if (i == a.forced.size() || b.forced[j] < a.forced[i]) ++j;
    else ++i, ++j;
    ++f;
  }
  return uint32_t(t + f);
}

// one generated lane kernel of a plan: its classes (indices into the plan's
// classes), their shapes and loads, and its layout
struct LaneModule {
  std::vector<uint32_t> cls;
  std::vector<LaneShape> shapes;
  std::vector<ClassLoad> load;
  LaneLayout lay;
};

// Lanes per instance of every class of a single-wave lane kernel.  A class's
// work is its step cost x instance-steps x G / 64 (wave instructions); the
// kernel takes about max(total work / SIMDs, longest chain), a chain being a
// class's step cost x its longest step count (one wave's sequence of steps).
// Every class starts at its cheapest layout that fits the registers; the
// class with the longest chain is widened while that shortens the estimate.
// `want` > 0 requests that many lanes for every class where it fits.
void choose_groups(std::vector<LaneShape>& shapes, const std::vector<ClassLoad>& load, int bits,
                   double simds, uint32_t want, int bias) {
  const size_t C = shapes.size();
  std::vector<std::vector<std::pair<uint32_t, double>>> opt(C);
  std::vector<size_t> sel(C, 0);
  for (size_t c = 0; c < C; ++c) {
    opt[c] = group_options(shapes[c], bits);
    if (std::getenv("PASIF_LANE_DEBUG")) {
      for (uint32_t G = 1; G <= kLaneWave; G *= 2) {
        LaneShape t = shapes[c];
        const bool ok = lane_group(t, G);
        fprintf(stderr, "[pasif] class %zu G %u: %s live %u cost %.0f nl %u/%u coefs %u/%u nx %zu\n", c, G,
                ok ? "ok" : "no", ok ? lane_live(t) : 0u, ok ? lane_step_cost(t, bits) : 0.0,
                t.gf.nl(), t.ga.nl(), t.gf.coefs(), t.ga.coefs(), t.gx.size());
      }
    }
    if (opt[c].empty()) continue;
    size_t b = 0;
    for (size_t k = 0; k < opt[c].size(); ++k) {
      if (want && opt[c][k].first == want) {
        b = k;
        break;
      }
      if (opt[c][k].second * opt[c][k].first < opt[c][b].second * opt[c][b].first) b = k;
    }
    sel[c] = b;
  }
  auto work = [&](size_t c, size_t k) {
    return opt[c][k].second * load[c].steps_sum * opt[c][k].first / double(kLaneWave);
  };
  auto chain = [&](size_t c, size_t k) { return opt[c][k].second * load[c].steps_max; };
  auto estimate = [&](const std::vector<size_t>& s, size_t* crit) {
    double w = 0, ch = 0;
    for (size_t c = 0; c < C; ++c) {
      if (opt[c].empty()) continue;
      w += work(c, s[c]);
      if (chain(c, s[c]) > ch) {
        ch = chain(c, s[c]);
        if (crit) *crit = c;
      }
    }
    return std::max(w / simds, ch);
  };
  for (int it = 0; it < 64 && !want; ++it) {
    size_t m = 0;
    const double est = estimate(sel, &m);
    if (opt[m].empty() || sel[m] + 1 >= opt[m].size()) break;
    std::vector<size_t> t = sel;
    t[m] += 1;
    if (estimate(t, nullptr) >= 0.99 * est) break;
    sel = t;
  }
  for (size_t c = 0; c < C; ++c) {
    if (opt[c].empty()) continue;
    // the bias moves each class along its feasible sizes (clamped)
    const long k = std::min<long>(long(opt[c].size()) - 1, std::max<long>(0, long(sel[c]) + (want ? 0 : bias)));
    (void)lane_group(shapes[c], opt[c][size_t(k)].first);
  }
}

// SIMDs of the device (4 per CU), 1024 when planning without one
double device_simds(int dev) {
  int cus = 0;
  if (dev >= 0 && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    return 4.0 * cus;
  return 1024.0;
}

// Waves per SIMD a single-wave kernel should keep resident: enough for all
// its waves to run at once, as far as the largest class's live values
// (plus temporaries) leave registers for (512 VGPRs per SIMD lane).
uint32_t single_wave_occupancy(const std::vector<LaneShape>& shapes,
                               const std::vector<ClassLoad>& load, int bits, int dev) {
  double waves = 0;
  uint32_t live = 0;
  for (size_t c = 0; c < shapes.size(); ++c) {
    waves += std::ceil(load[c].inst * shapes[c].group / double(kLaneWave));
    live = std::max(live, lane_live(shapes[c]));
  }
  const uint32_t regs = live * uint32_t(bits / 32) + 48;
  const uint32_t by_regs = std::max<uint32_t>(1, 512 / regs);
  const uint32_t by_waves = uint32_t(std::min(8.0, std::max(1.0, std::ceil(waves / device_simds(dev)))));
  return std::min(by_regs, by_waves);
}

// Checkpoint intervals of the classes of a single-wave adjoint kernel: the
// requested one, shortened until a class's segment [K+1][slots][64] and its
// sample stage fit the LDS one wave may take while every wave of the launch
// stays resident (160 KB per CU shared by ceil(waves / CUs) waves, at most
// 64 KB, at least 5 steps of segment).  When the checkpoints of all waves
// (a state every K steps of each wave's longest instance) would take more
// than half the free device memory, the LDS share grows (fewer resident
// waves, longer intervals) until they fit.  Returns the shortest interval.
uint32_t single_wave_segs(std::vector<LaneShape>& shapes, const std::vector<ClassLoad>& load,
                          size_t el, uint32_t Kreq, int dev, bool exact = false) {
  double waves = 0;
  for (size_t c = 0; c < shapes.size(); ++c)
    waves += std::ceil(load[c].inst * shapes[c].group / double(kLaneWave));
  const double cus = device_simds(dev) / 4.0;
  const double per_cu = std::max(1.0, std::ceil(waves / cus));
  double cap = 64e9;
  size_t fr = 0, tot = 0;
  if (dev >= 0 && hipMemGetInfo(&fr, &tot) == hipSuccess) cap = 0.5 * double(fr);
  size_t budget = std::min<size_t>(kLaneMaxLds, size_t(double(kMaxLdsBytes) / per_cu));
  uint32_t kmin = Kreq;
  for (;;) {
    double bytes = 0;
    kmin = Kreq;
    for (size_t c = 0; c < shapes.size(); ++c) {
      LaneShape& sh = shapes[c];
      const uint32_t G = sh.group > 1 ? sh.group : 1u;
      const size_t stage = size_t(kLaneStage + (G > 1 ? 3 * kLaneChunk : 0)) * (kLaneWave / G) * el;
      // a step of the segment in LDS: the rows the reverse step reads; a
      // checkpoint: the whole forward state
      const size_t step = size_t(lane_seg_rows(sh)) * kLaneWave * el;
      const size_t ckstep = size_t(G > 1 ? sh.gf.nl() : sh.f.n) * kLaneWave * el;
      // K + 1 post-step states (and 3 K stage inputs for the exact adjoint)
      const size_t per = exact ? 4 : 1;
      const size_t b = std::min<size_t>(kLaneMaxLds, std::max(budget, stage + (6 * per) * step));
      const size_t fit = b > stage ? (b - stage) / step : 0;
      sh.seg = uint32_t(std::max<size_t>(1, std::min<size_t>(Kreq, fit > 0 ? (fit - 1) / per : 0)));
      kmin = std::min(kmin, sh.seg);
      // a checkpoint of every wave's lanes every seg steps of its longest instance
      bytes += (load[c].steps_sum * G / double(kLaneWave) / sh.seg + load[c].inst * G / kLaneWave) *
               double(ckstep);
    }
    if (std::getenv("PASIF_LANE_DEBUG"))
      fprintf(stderr, "[pasif] LDS share %zu B: checkpoints %.3g GB (cap %.3g GB), shortest interval %u\n",
              budget, bytes / 1e9, cap / 1e9, kmin);
    if (bytes <= cap || budget >= kLaneMaxLds) break;
    budget = std::min<size_t>(kLaneMaxLds, budget + budget / 4 + 1024);
  }
  return kmin;
}

// Plan the lane kernel for the loaded problem: true if every system fits a
// lane and the systems form at most kLaneMaxClasses structure classes.
template <typename R>
LanePart* lane_build(pasif_driver* d, bool adjoint, int part,
                     const std::vector<LaneShape>& shapes, const std::vector<uint32_t>& klass,
                     const LaneLayout& lay);

template <typename R>
bool lane_plan(pasif_driver* d, bool adjoint) {
  const auto tp0 = std::chrono::steady_clock::now();
  Packed& pk = d->pk;
  const uint32_t S = uint32_t(d->fwd.size.size());
  const double eps = std::numeric_limits<R>::epsilon();
  const bool forced_lane = d->kernel_path == PASIF_KERNEL_LANE;
  const bool exact = adjoint && d->adj_mode != 0;
  const int bits = int(sizeof(R) * 8);
  auto& cache = d->lcache[adjoint ? 1 : 0];
  const uint64_t sig = d->pool ? 0 : lane_plan_sig(d, adjoint, bits);
  if (!d->pool && cache.valid && cache.sig == sig) {
    // the same structure with new values: the cached classes, layouts and
    // kernels, parts rebuilt with the current coefficients
    d->lane_cls = cache.lane_cls;
    d->lane_struct = cache.lane_struct;
    if (cache.lays.empty()) return false;
    pk.lshapes.clear();
    pk.lgroups.clear();
    pk.lmods.clear();
    pk.lseg = 0;
    for (size_t m = 0; m < cache.lays.size(); ++m) {
      const auto& shapes = cache.shapes[m];
      const LaneLayout& lk = cache.lays[m];
      std::vector<uint32_t> klass(shapes.size());
      std::iota(klass.begin(), klass.end(), 0u);
      LanePart* part = lane_build<R>(d, adjoint, int(m), shapes, klass, lk);
      part->lds = lane_lds_bytes(shapes, bits, lk);
      if (m < cache.kerns.size()) part->kern = cache.kerns[m];
      for (const LaneShape& x : shapes) {
        pk.lshapes.push_back(x);
        pk.lgroups.push_back(lk.waves > 1 ? 1u : x.group);
      }
      pk.lmods.push_back({shapes, lk});
      pk.lseg = pk.lseg ? std::min(pk.lseg, lk.seg) : lk.seg;
    }
    pk.llay = cache.lays[0];
    return true;
  }
  cache.valid = false;
  d->lane_cls.assign(S, -1);
  // structure classes of every system
  std::map<std::vector<uint32_t>, uint32_t> cls_of_key;
  std::vector<uint32_t> cls_of(S);
  std::vector<LaneShape> all;
  std::vector<uint32_t> key;
  for (uint32_t s = 0; s < S; ++s) {
    LaneShape sh;
    sh.f = lane_side(d->fwd, s, false, eps);
    sh.adjoint = adjoint;
    if (adjoint) sh.a = lane_side(d->adj, s, true, eps);
    key.clear();
    shape_key(sh.f, key);
    if (adjoint) shape_key(sh.a, key);
    auto it = cls_of_key.find(key);
    if (it == cls_of_key.end()) {
      it = cls_of_key.emplace(key, uint32_t(all.size())).first;
      all.push_back(std::move(sh));
    }
    cls_of[s] = it->second;
  }
  // what each class needs: 1 = one lane, w >= 2 = a workgroup of w waves, 0 = too large
  std::vector<uint32_t> need(all.size(), 0);
  std::vector<double> work(all.size(), 0.0);
  for (size_t c = 0; c < all.size(); ++c) {
    LaneShape& sh = all[c];
    // lane_group == 1: one lane per instance only (workgroups of waves for
    // classes that do not fit one lane)
    const bool fits = d->lane_group == 1 ? sh.live_values() <= lane_max_live()
                                         : !group_options(sh, bits).empty();
    if (fits) {
      need[c] = 1;  // one lane or a group of lanes per instance
    } else {
      for (uint32_t w = 2; w <= kLaneMaxWaves && !need[c]; ++w)
        if (lane_partition(sh, w) <= kLaneMaxWaveLive) need[c] = w;
    }
  }
  std::vector<ClassLoad> loads(all.size());
  for (const Instance& in : make_instances(d, adjoint)) {
    const uint32_t c = cls_of[in.sys];
    work[c] += double(in.steps) * (all[c].f.n + all[c].a.n);
    loads[c].inst += 1;
    loads[c].steps_sum += in.steps;
    loads[c].steps_max = std::max(loads[c].steps_max, double(in.steps));
  }
  // Structures of one size merge into union classes (a system's coefficient
  // is zero where it lacks a term of the union) while more classes of one
  // layout (one wave per instance group, or a workgroup of waves per
  // instance) remain than one module takes: cheapest merge first, as long as
  // a merge adds at most half the work of the two classes it joins and the
  // union still fits that layout.  Among the 256 heaviest classes.
  d->lane_struct = cls_of;
  std::vector<uint32_t> into(all.size());  // class a class merged into (itself if kept)
  std::iota(into.begin(), into.end(), 0u);
  {
    const size_t C = all.size();
    std::vector<uint32_t> cand;
    for (uint32_t c = 0; c < C; ++c)
      if (need[c] >= 1) cand.push_back(c);
    std::stable_sort(cand.begin(), cand.end(),
                     [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
    if (cand.size() > 256) cand.resize(256);
    // classes of each layout (0: one wave, 1: several) among the candidates
    size_t live[2] = {0, 0};
    for (uint32_t c : cand) live[need[c] == 1 ? 0 : 1]++;
    std::vector<double> cost(C), steps(C);
    for (uint32_t c = 0; c < C; ++c) {
      cost[c] = class_cost(all[c]);
      steps[c] = loads[c].steps_sum;
    }
    const size_t M = cand.size();
    // added work of merging candidates x and y (infinite: not mergeable)
    const double kNo = std::numeric_limits<double>::infinity();
    std::vector<double> gain(M * M, kNo);
    std::vector<char> alive(M, 1);
    auto eval = [&](size_t x, size_t y) {
      const LaneShape& A = all[cand[x]];
      const LaneShape& B = all[cand[y]];
      // the forward sides must match in size; an adjoint side pads to the larger
      if (A.f.n != B.f.n || A.adjoint != B.adjoint) return kNo;
      const int la = need[cand[x]] == 1 ? 0 : 1, lb = need[cand[y]] == 1 ? 0 : 1;
      if (la != lb || live[la] <= kLaneModuleClasses) return kNo;
      double cu = union_coefs(A.f, B.f) + 2.0 * A.f.n;
      if (A.adjoint) cu += union_coefs(A.a, B.a) + 2.0 * A.a.n + A.f.n;
      const double wa = steps[cand[x]] * cost[cand[x]], wb = steps[cand[y]] * cost[cand[y]];
      const double added = steps[cand[x]] * (cu - cost[cand[x]]) + steps[cand[y]] * (cu - cost[cand[y]]);
      return added <= 0.5 * (wa + wb) ? added / std::max(wa + wb, 1e-300) : kNo;
    };
    for (size_t x = 0; x < M; ++x)
      for (size_t y = x + 1; y < M; ++y) gain[x * M + y] = eval(x, y);
    while (live[0] > kLaneModuleClasses || live[1] > kLaneModuleClasses) {
      size_t bx = M, by = M;
      double best = kNo;
      for (size_t x = 0; x < M; ++x)
        for (size_t y = x + 1; y < M; ++y)
          if (alive[x] && alive[y] && gain[x * M + y] < best) {
            best = gain[x * M + y];
            bx = x;
            by = y;
          }
      if (bx == M) break;
      const uint32_t cx = cand[bx], cy = cand[by];
      LaneShape U = all[cx];
      U.f = side_union(all[cx].f, all[cy].f);
      if (U.adjoint) U.a = side_union(all[cx].a, all[cy].a);
      const int lay = need[cx] == 1 ? 0 : 1;
      uint32_t wu = 0;  // waves the union needs
      if (lay == 0) {
        wu = (d->lane_group == 1 ? U.live_values() <= lane_max_live()
                                 : !group_options(U, bits).empty())
                 ? 1u
                 : 0u;
      } else {
        for (uint32_t w = 2; w <= kLaneMaxWaves && !wu; ++w) {
          LaneShape t = U;
          if (lane_partition(t, w) <= kLaneMaxWaveLive) wu = w;
        }
      }
      const bool fits = wu != 0;
      if (std::getenv("PASIF_LANE_DEBUG"))
        fprintf(stderr, "[pasif] merge classes %u (n %u/%u) + %u (n %u/%u): +%.1f %% work%s\n", cx,
                all[cx].f.n, all[cx].a.n, cy, all[cy].f.n, all[cy].a.n, 100.0 * best,
                fits ? "" : ", union does not fit a lane: kept apart");
      if (!fits) {
        gain[bx * M + by] = kNo;
        continue;
      }
      all[cx] = U;
      need[cx] = wu;
      cost[cx] = class_cost(U);
      steps[cx] += steps[cy];
      loads[cx].inst += loads[cy].inst;
      loads[cx].steps_sum += loads[cy].steps_sum;
      loads[cx].steps_max = std::max(loads[cx].steps_max, loads[cy].steps_max);
      work[cx] += work[cy];
      into[cy] = cx;
      alive[by] = 0;
      const size_t before = live[lay]--;
      if (before - 1 == kLaneModuleClasses) {
        // this layout is down to one module: no further merges in it
        for (size_t x = 0; x < M; ++x)
          for (size_t y = x + 1; y < M; ++y)
            if (alive[x] && alive[y] && (need[cand[x]] == 1 ? 0 : 1) == lay) gain[x * M + y] = kNo;
        continue;
      }
      for (size_t z = 0; z < M; ++z) {
        if (!alive[z] || z == bx) continue;
        const double g = eval(std::min(bx, z), std::max(bx, z));
        gain[std::min(bx, z) * M + std::max(bx, z)] = g;
      }
    }
    for (uint32_t s2 = 0; s2 < S; ++s2) {
      uint32_t c = cls_of[s2];
      while (into[c] != c) c = into[c];
      cls_of[s2] = c;
    }
    if (std::getenv("PASIF_LANE_DEBUG"))
      for (uint32_t c = 0; c < C; ++c)
        if (into[c] == c)
          fprintf(stderr, "[pasif] class %u: n %u/%u, %.0f instances, needs %u wave(s)%s\n", c,
                  all[c].f.n, all[c].a.n, loads[c].inst, need[c],
                  need[c] ? "" : " (too large for lanes)");
  }
  // the lane kernel takes the classes that fit, heaviest first, up to kLaneMaxClasses
  std::vector<uint32_t> order;
  size_t kept = 0;
  for (uint32_t c = 0; c < all.size(); ++c) {
    if (into[c] != c) continue;
    ++kept;
    if (need[c]) order.push_back(c);
  }
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t x, uint32_t y) { return work[x] > work[y]; });
  if (order.size() > kLaneMaxClasses) order.resize(kLaneMaxClasses);
  const char* why = nullptr;
  if (order.size() < kept) {
    bool too_large = false;
    for (uint32_t c = 0; c < all.size(); ++c) too_large = too_large || (into[c] == c && !need[c]);
    why = too_large ? "a system needs more registers than a workgroup of lanes holds"
                    : "the systems have too many distinct structures";
    PASIF_REQUIRE(!forced_lane, PASIF_ERR_UNSUPPORTED, "lane kernel: %s", why);
    why = nullptr;
  }
  if (order.empty()) {
    if (!d->pool) {
      cache.lane_cls = d->lane_cls;
      cache.lane_struct = d->lane_struct;
      cache.shapes.clear();
      cache.lays.clear();
      cache.kerns.clear();
      cache.sig = sig;
      cache.valid = true;
    }
    return false;
  }
  // Modules: the classes that fit one wave per instance group, and those
  // that spread an instance over a workgroup of waves, each set split into
  // modules of at most kLaneModuleClasses classes (dealt out in order of
  // work, so the modules carry similar work); every module is one kernel,
  // compiled on its own and launched beside the others.
  const size_t el = sizeof(R);
  const uint32_t Kreq = d->ck_every ? d->ck_every : 256;
  std::vector<uint32_t> one, many;
  uint32_t bodies = 0;
  for (uint32_t c : order) {
    if (need[c] == 1) {
      one.push_back(c);
    } else if (!exact && (forced_lane || bodies + need[c] <= kLaneMultiWaveBodies)) {
      // (the exact discrete adjoint runs on single-wave kernels only)
      many.push_back(c);
      bodies += need[c];
    }
  }
  std::vector<LaneModule> mods;
  auto deal = [&](const std::vector<uint32_t>& cls, const LaneLayout& base) {
    const size_t nm = (cls.size() + kLaneModuleClasses - 1) / kLaneModuleClasses;
    const size_t first = mods.size();
    for (size_t m = 0; m < nm; ++m) mods.push_back(LaneModule{{}, {}, {}, base});
    for (size_t i = 0; i < cls.size(); ++i) {
      LaneModule& M = mods[first + i % nm];
      M.cls.push_back(cls[i]);
      M.shapes.push_back(all[cls[i]]);
      M.load.push_back(loads[cls[i]]);
    }
  };
  if (!one.empty()) {
    // group sizes, checkpoint intervals and the ILP choice over every
    // single-wave class together (their modules share the GPU)
    std::vector<LaneShape> sh;
    std::vector<ClassLoad> ld;
    for (uint32_t c : one) {
      sh.push_back(all[c]);
      ld.push_back(loads[c]);
    }
    choose_groups(sh, ld, bits, device_simds(d->dev), d->lane_group, d->lane_bias);
    if (adjoint) (void)single_wave_segs(sh, ld, el, Kreq, d->dev, exact);
    double waves = 0;
    for (size_t c = 0; c < sh.size(); ++c)
      waves += std::ceil(ld[c].inst * sh[c].group / double(kLaneWave));
    for (size_t i = 0; i < one.size(); ++i) all[one[i]] = sh[i];
    LaneLayout lay;
    lay.ilp = waves <= device_simds(d->dev);
    lay.exact = exact;
    const uint32_t by_waves =
        uint32_t(std::min(8.0, std::max(1.0, std::ceil(waves / device_simds(d->dev)))));
    const size_t m0 = mods.size();
    deal(one, lay);
    for (size_t m = m0; m < mods.size(); ++m) {
      LaneModule& M = mods[m];
      uint32_t live = 0;
      M.lay.seg = adjoint ? Kreq : 1u;
      for (const LaneShape& x : M.shapes) {
        live = std::max(live, lane_live(x));
        if (adjoint) M.lay.seg = std::min(M.lay.seg, x.seg ? x.seg : Kreq);
      }
      // waves per SIMD: the registers of this module's widest class, the
      // residency of every single-wave module together
      const uint32_t by_regs = std::max<uint32_t>(1, 512 / (live * uint32_t(bits / 32) + 48));
      M.lay.occ = std::min(by_regs, by_waves);
    }
  }
  if (!many.empty()) {
    const size_t m0 = mods.size();
    deal(many, LaneLayout{});
    double total = 0;
    for (uint32_t c : many) total += loads[c].inst;
    for (size_t m = m0; m < mods.size(); ++m) {
      LaneModule& M = mods[m];
      LaneLayout& lay = M.lay;
      uint32_t waves = 2;
      for (uint32_t c : M.cls) waves = std::max(waves, need[c]);
      for (LaneShape& x : M.shapes) (void)lane_partition(x, waves);
      lay.waves = waves;
      // instances per workgroup: enough workgroups for every CU, then the
      // longest checkpoint interval (up to the requested one) that fits LDS
      lay.lw = total >= 256.0 * 64 ? 64 : (total >= 256.0 * 32 ? 32 : 16);
      bool fits = false;
      for (;; lay.lw /= 2) {
        lay.seg = 1;
        if (adjoint) {
          uint32_t lo = 1, hi = Kreq;
          lay.seg = 0;
          while (lo <= hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            LaneLayout t = lay;
            t.seg = mid;
            if (lane_lds_bytes(M.shapes, int(el * 8), t) <= kMaxLdsBytes) {
              lay.seg = mid;
              lo = mid + 1;
            } else {
              hi = mid - 1;
            }
          }
        }
        fits = lay.seg >= 1 && lane_lds_bytes(M.shapes, int(el * 8), lay) <= kMaxLdsBytes;
        if (fits && (!adjoint || lay.seg >= std::min<uint32_t>(Kreq, 8) || lay.lw == 16)) break;
        if (lay.lw == 16) break;
      }
      if (!fits) {
        // these classes stay on tiles
        PASIF_REQUIRE(!forced_lane, PASIF_ERR_UNSUPPORTED,
                      "lane kernel: the systems' stage buffers do not fit LDS");
        M.cls.clear();
        M.shapes.clear();
      }
    }
    mods.erase(std::remove_if(mods.begin(), mods.end(),
                              [](const LaneModule& M) { return M.cls.empty(); }),
               mods.end());
  }
  // a module whose LDS does not fit leaves its classes to the tiles
  for (LaneModule& M : mods) {
    const size_t lds = lane_lds_bytes(M.shapes, bits, M.lay);
    if (lds <= (M.lay.waves > 1 ? kMaxLdsBytes : kLaneMaxLds)) continue;
    PASIF_REQUIRE(!forced_lane, PASIF_ERR_UNSUPPORTED, "lane kernel LDS %zu bytes", lds);
    M.cls.clear();
    M.shapes.clear();
  }
  mods.erase(std::remove_if(mods.begin(), mods.end(),
                            [](const LaneModule& M) { return M.cls.empty(); }),
             mods.end());
  if (mods.empty()) {
    d->lane_cls.assign(S, -1);
    if (!d->pool) {
      cache.lane_cls = d->lane_cls;
      cache.lane_struct = d->lane_struct;
      cache.shapes.clear();
      cache.lays.clear();
      cache.kerns.clear();
      cache.sig = sig;
      cache.valid = true;
    }
    return false;
  }
  // lane_cls of every system before the parts are built (a part gathers its
  // instances through it)
  std::vector<int32_t> code_of(all.size(), -1);
  for (size_t m = 0; m < mods.size(); ++m)
    for (size_t k = 0; k < mods[m].cls.size(); ++k)
      code_of[mods[m].cls[k]] = lane_code(int(m), uint32_t(k));
  for (uint32_t s = 0; s < S; ++s) d->lane_cls[s] = code_of[cls_of[s]];
  const auto tp1 = std::chrono::steady_clock::now();
  std::vector<std::string> srcs;
  pk.lshapes.clear();
  pk.lgroups.clear();
  pk.lmods.clear();
  pk.lseg = 0;
  for (size_t m = 0; m < mods.size(); ++m) {
    LaneModule& M = mods[m];
    LaneLayout lk = M.lay;
    lk.seg = adjoint ? lk.seg : 1u;
    std::vector<uint32_t> klass(M.shapes.size());
    for (size_t c = 0; c < M.shapes.size(); ++c) klass[c] = uint32_t(c);
    LanePart* part = lane_build<R>(d, adjoint, int(m), M.shapes, klass, lk);
    const size_t lds = lane_lds_bytes(M.shapes, int(sizeof(R) * 8), lk);
    PASIF_REQUIRE(lds <= (lk.waves > 1 ? kMaxLdsBytes : kLaneMaxLds), PASIF_ERR_UNSUPPORTED,
                  "lane kernel LDS %zu bytes", lds);
    part->lds = lds;
    srcs.push_back(lane_source(M.shapes, int(sizeof(R) * 8), lk));
    for (const LaneShape& x : M.shapes) {
      pk.lshapes.push_back(x);
      pk.lgroups.push_back(lk.waves > 1 ? 1u : x.group);
    }
    pk.lmods.push_back({M.shapes, lk});
    pk.lseg = pk.lseg ? std::min(pk.lseg, lk.seg) : lk.seg;
    M.lay = lk;
  }
  if (d->dev >= 0) {
    std::string err;
    if (!lane_jit_flags_ok(&err)) {
      set_err(d, "%s", err.c_str());
      throw Fail{PASIF_ERR_ARGUMENT};
    }
    std::vector<LaneKernel*> outs;
    std::vector<bool> ilp;
    for (size_t m = 0; m < mods.size(); ++m) {
      outs.push_back(&pk.lparts[pk.lparts.size() - mods.size() + m]->kern);
      ilp.push_back(mods[m].lay.ilp);
    }
    if (!lane_compile_many(srcs, ilp, d->dev, outs, &err)) {
      set_err(d, "%s", err.c_str());
      throw Fail{PASIF_ERR_DEVICE};
    }
  }
  if (std::getenv("PASIF_PACK_TIMING"))
    fprintf(stderr, "[pasif] lane plan %.1f ms, lane build and compile %.1f ms\n",
            std::chrono::duration<double, std::milli>(tp1 - tp0).count(),
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tp1)
                .count());
  pk.llay = mods[0].lay;
  if (!d->pool) {
    cache.lane_cls = d->lane_cls;
    cache.lane_struct = d->lane_struct;
    cache.shapes.clear();
    cache.lays.clear();
    cache.kerns.clear();
    for (size_t m = 0; m < mods.size(); ++m) {
      cache.shapes.push_back(mods[m].shapes);
      cache.lays.push_back(mods[m].lay);
      if (d->dev >= 0) cache.kerns.push_back(pk.lparts[pk.lparts.size() - mods.size() + m]->kern);
    }
    cache.sig = sig;
    cache.valid = true;
  }
  if (d->d_compute)
    for (size_t m = 0; m < mods.size(); ++m) {
      const LanePart& P = *pk.lparts[pk.lparts.size() - mods.size() + m];
      printf("[pasif] lane module %zu: %zu structure class(es), %u workgroups of %u wave(s), "
             "T max %u, checkpoint %u\n",
             m, mods[m].shapes.size(), P.blocks, mods[m].lay.waves, P.tmax, mods[m].lay.seg);
    }
  return true;
}

// Build lane part `part` of the packed problem: the instances of the systems
// whose lane_cls names this part, one LaneCls per class with instances
// (sorted by step count, descending), coefficients / initial states /
// indices uploaded.  shapes[c] is the structure of class c, klass[c] its case
// in the kernel that will run the part (laid out as `lay`).
template <typename R>
LanePart* lane_build(pasif_driver* d, bool adjoint, int part,
                     const std::vector<LaneShape>& shapes, const std::vector<uint32_t>& klass,
                     const LaneLayout& lay) {
  Packed& pk = d->pk;
  const uint32_t unit = lay.waves > 1 ? lay.lw : kLaneWave;  // instances per workgroup
  std::vector<Instance> inst = make_instances(d, adjoint, kPartLane);
  std::vector<std::vector<const Instance*>> members(shapes.size());
  for (const Instance& in : inst) {
    const int32_t code = d->lane_cls[in.sys];
    if (lane_part_of(code) == part) members[lane_class_of(code)].push_back(&in);
  }
  std::unique_ptr<LanePart> P(new LanePart());
  P->lay = lay;
  uint32_t block = 0, tmax = 0;
  // the part's program arrays, 256-byte aligned in one host buffer (a pool
  // job reuses the pool's): one upload for the whole part.  Every class's
  // fields are placed first, the buffer is cleared once, then the fields are
  // written in place; (field, offset) is patched in after the upload.
  std::vector<unsigned char> own;
  std::vector<unsigned char>& host = d->pool ? d->pool->stage_host : own;
  std::vector<std::pair<const void**, size_t>> views;
  size_t top = 0;
  auto place = [&](const void** where, size_t bytes) {
    const size_t off = (top + 255) / 256 * 256;
    top = off + std::max<size_t>(bytes, 1);
    views.push_back({where, off});
    return off;
  };
  const uint32_t out_len_f = d->mapped ? d->fwd.n : d->fwd.n * d->n_exc;
  const uint32_t out_len_a = adjoint ? (d->mapped ? d->adj.n : d->adj.n * d->n_exc) : 0;
  // classes and their fields' places
  struct Fields {
    size_t cf, ca, qf, qa, stp, ex, of, oa, na, ck;
  };
  std::vector<Fields> fld(shapes.size());
  for (size_t c = 0; c < shapes.size(); ++c) {
    auto& mem = members[c];
    if (mem.empty()) continue;
    std::stable_sort(mem.begin(), mem.end(),
                     [](const Instance* x, const Instance* y) { return x->steps > y->steps; });
    std::unique_ptr<LaneCls> L(new LaneCls());
    const LaneShape& sh = shapes[c];
    // lane groups: G lanes per instance, per-lane arrays over lane columns
    // (instance i, position p at column i * G + p), per-instance arrays over
    // instances
    const bool grp = lay.waves == 1 && sh.group > 1;
    const uint32_t G = grp ? sh.group : 1u;
    L->count = uint32_t(mem.size());
    if (grp) {
      const uint32_t ipw = kLaneWave / G;
      L->waves = std::max<uint32_t>(1, (L->count + ipw - 1) / ipw);
      L->stride = L->waves * kLaneWave;
    } else {
      L->stride = std::max<uint32_t>(kLaneWave, (L->count + kLaneWave - 1) / kLaneWave * kLaneWave);
      L->waves = L->stride / unit;
    }
    L->block_begin = block;
    L->nf = grp ? sh.gf.nl() : sh.f.n;
    L->na = adjoint ? (grp ? sh.ga.nl() : sh.a.n) : 0;
    L->klass = klass[c];
    L->seg = adjoint ? ((lay.waves == 1 && sh.seg) ? sh.seg : lay.seg) : 1u;
    block += L->waves;
    const uint32_t ipb = grp ? kLaneWave / G : unit;
    L->wave_t.assign(L->waves, 0u);
    for (uint32_t w = 0; w < L->waves; ++w)
      if (size_t(w) * ipb < mem.size()) L->wave_t[w] = mem[size_t(w) * ipb]->steps;
    L->live = L->waves;
    const size_t st = L->stride;
    const uint32_t TF = grp ? sh.gf.coefs() : sh.f.coefs();
    const uint32_t TA = adjoint ? (grp ? sh.ga.coefs() : sh.a.coefs()) : 0;
    Fields& f = fld[c];
    f.ck = (adjoint && lay.waves == 1) ? place(&L->ckoff, L->waves * sizeof(uint64_t)) : 0;
    f.cf = place(&L->coef_f, size_t(std::max<uint32_t>(TF, 1)) * st * sizeof(R));
    f.ca = place(&L->coef_a, size_t(std::max<uint32_t>(TA, 1)) * st * sizeof(R));
    f.qf = place(&L->init_f, size_t(L->nf) * st * sizeof(R));
    f.qa = place(&L->init_a, size_t(std::max<uint32_t>(L->na, 1)) * st * sizeof(R));
    f.stp = place(&L->steps, st * sizeof(uint32_t));
    f.ex = place(&L->exc, st * sizeof(int32_t));
    f.of = place(&L->out_f, st * sizeof(uint32_t));
    f.oa = place(&L->out_a, st * sizeof(uint32_t));
    f.na = place(&L->na_out, st * sizeof(uint32_t));
    P->cls.push_back(std::move(L));
  }
  if (host.size() < top) host.resize(top);
  std::memset(host.data(), 0, top);
  size_t k = 0;
  for (size_t c = 0; c < shapes.size(); ++c) {
    auto& mem = members[c];
    if (mem.empty()) continue;
    LaneCls* L = P->cls[k++].get();
    const LaneShape& sh = shapes[c];
    const Fields& f = fld[c];
    const bool grp = lay.waves == 1 && sh.group > 1;
    const uint32_t G = grp ? sh.group : 1u;
    const size_t st = L->stride;
    const uint32_t TF = grp ? sh.gf.coefs() : sh.f.coefs();
    const uint32_t TA = adjoint ? (grp ? sh.ga.coefs() : sh.a.coefs()) : 0;
    R* cf = reinterpret_cast<R*>(host.data() + f.cf);
    R* ca = reinterpret_cast<R*>(host.data() + f.ca);
    R* qf = reinterpret_cast<R*>(host.data() + f.qf);
    R* qa = reinterpret_cast<R*>(host.data() + f.qa);
    uint32_t* stp = reinterpret_cast<uint32_t*>(host.data() + f.stp);
    int32_t* ex = reinterpret_cast<int32_t*>(host.data() + f.ex);
    uint32_t* of = reinterpret_cast<uint32_t*>(host.data() + f.of);
    uint32_t* oa = reinterpret_cast<uint32_t*>(host.data() + f.oa);
    uint32_t* nao = reinterpret_cast<uint32_t*>(host.data() + f.na);
    std::fill(ex, ex + st, -1);
    const double eps = std::numeric_limits<R>::epsilon();
    // lane slot r of position p of a group side -> the system's row
    auto slot_row = [](const GSide& gs, uint32_t r, uint32_t p) {
      return r < gs.RL ? gs.band[r] + p * gs.stride[r] : gs.shared[r - gs.RL];
    };
    // bounds and per-instance words, in order
    for (uint32_t i = 0; i < L->count; ++i) {
      const Instance& in = *mem[i];
      // every output index the kernel writes stays inside the output arrays
      PASIF_REQUIRE(uint64_t(in.out_f) + sh.f.n <= out_len_f, PASIF_ERR_UNSUPPORTED,
                    "lane plan: output index out of range");
      // adjoint rows of this instance's own system (a merged class may pad
      // its adjoint side beyond them; padded rows are never stored)
      const uint32_t own_a = adjoint ? d->adj.size[in.sys] : 0u;
      PASIF_REQUIRE(!adjoint || (own_a <= sh.a.n && uint64_t(in.out_a) + own_a <= out_len_a),
                    PASIF_ERR_UNSUPPORTED, "lane plan: gradient index out of range");
      PASIF_REQUIRE(in.exc < int32_t(d->n_exc), PASIF_ERR_UNSUPPORTED,
                    "lane plan: excitation out of range");
      stp[i] = in.steps;
      ex[i] = in.exc;
      of[i] = in.out_f;
      oa[i] = in.out_a;
      nao[i] = own_a;
      L->tmax = std::max(L->tmax, in.steps);
    }
    // coefficients and initial states of instances [i0, i1) (instances
    // write disjoint columns: ranges run on threads of their own, each with
    // its own coefficient maps); false on a structure mismatch
    auto fill = [&](uint32_t i0, uint32_t i1) -> bool {
      std::map<uint32_t, CoefMap> map_f, map_a;
      std::vector<double> cv, lv;
      try {
        for (uint32_t i = i0; i < i1; ++i) {
          const Instance& in = *mem[i];
          const uint32_t own_a = nao[i];
          const uint32_t sk = in.sys < d->lane_struct.size() ? d->lane_struct[in.sys] : 0u;
          lane_coefs_mapped(d->fwd, in.sys, sh.f, false, eps, map_f[sk], cv);
          if (cv.size() != sh.f.coefs()) return false;
          if (grp) {
            for (uint32_t p = 0; p < G; ++p) {
              const size_t col = size_t(i) * G + p;
              lane_group_coefs(sh.gf, cv, p, lv);
              for (uint32_t t = 0; t < TF; ++t) cf[size_t(t) * st + col] = R(lv[t]);
              for (uint32_t r = 0; r < L->nf; ++r)
                qf[size_t(r) * st + col] =
                    R(d->fwd.init[d->fwd.off[in.sys] + slot_row(sh.gf, r, p)]);
            }
          } else {
            for (uint32_t t = 0; t < TF; ++t) cf[size_t(t) * st + i] = R(cv[t]);
            for (uint32_t r = 0; r < L->nf; ++r)
              qf[size_t(r) * st + i] = R(d->fwd.init[d->fwd.off[in.sys] + r]);
          }
          if (!adjoint) continue;
          lane_coefs_mapped(d->adj, in.sys, sh.a, true, eps, map_a[sk], cv);
          if (cv.size() != sh.a.coefs()) return false;
          if (grp) {
            for (uint32_t p = 0; p < G; ++p) {
              const size_t col = size_t(i) * G + p;
              lane_group_coefs(sh.ga, cv, p, lv);
              for (uint32_t t = 0; t < TA; ++t) ca[size_t(t) * st + col] = R(lv[t]);
              for (uint32_t r = 0; r < L->na; ++r) {
                const uint32_t row = slot_row(sh.ga, r, p);
                qa[size_t(r) * st + col] =
                    row < own_a ? R(d->adj.init[d->adj.off[in.sys] + row]) : R(0);
              }
            }
          } else {
            for (uint32_t t = 0; t < TA; ++t) ca[size_t(t) * st + i] = R(cv[t]);
            for (uint32_t r = 0; r < L->na; ++r)
              qa[size_t(r) * st + i] = r < own_a ? R(d->adj.init[d->adj.off[in.sys] + r]) : R(0);
          }
        }
      } catch (const Fail&) {
        return false;
      }
      return true;
    };
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t nth = std::min<uint32_t>(std::min(hw, 16u), (L->count + 1023) / 1024);
    bool ok = true;
    if (nth > 1) {
      std::vector<std::thread> th;
      std::vector<char> res(nth, 1);
      for (uint32_t t = 0; t < nth; ++t) {
        const uint32_t i0 = uint32_t(uint64_t(L->count) * t / nth);
        const uint32_t i1 = uint32_t(uint64_t(L->count) * (t + 1) / nth);
        th.emplace_back([&, t, i0, i1] { res[t] = fill(i0, i1) ? 1 : 0; });
      }
      for (std::thread& x : th) x.join();
      for (char r : res) ok = ok && r;
    } else {
      ok = fill(0, L->count);
    }
    PASIF_REQUIRE(ok, PASIF_ERR_UNSUPPORTED, "lane plan: structure mismatch");
    tmax = std::max(tmax, L->tmax);
    if (adjoint && lay.waves == 1) {
      // wave w runs instances [w*ipw, (w+1)*ipw), the first with the most steps
      const uint32_t ipw = kLaneWave / G;
      uint64_t* off = reinterpret_cast<uint64_t*>(host.data() + f.ck);
      uint64_t tot = 0;
      for (uint32_t w = 0; w < L->waves; ++w) {
        const uint32_t Tw = w * ipw < L->count ? stp[size_t(w) * ipw] : 0u;
        off[w] = tot;
        tot += uint64_t(std::max<uint32_t>(1, (Tw + L->seg - 1) / L->seg)) * L->nf * kLaneWave;
      }
      L->ck_elems = tot;
    }
  }
  const auto u0 = std::chrono::steady_clock::now();
  upload_bytes(d, P->prog, host.data(), top);
  for (auto& v : views)
    *v.first = P->prog.p ? static_cast<const unsigned char*>(P->prog.p) + v.second : nullptr;
  if (d->dev >= 0) PASIF_HIP(hipStreamSynchronize(d->stream));
  if (d->pool && d->pool->timing)
    d->pool->tm[4] +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - u0).count();
  P->blocks = block;
  P->tmax = tmax;
  P->module = part;
  pk.lane = true;
  pk.ltmax = std::max(pk.ltmax, tmax);
  pk.out_len_f = out_len_f;
  pk.out_len_a = out_len_a;
  pk.lparts.push_back(std::move(P));
  return pk.lparts.back().get();
}

void require_device(pasif_driver* d) {
  PASIF_REQUIRE(d->dev >= 0, PASIF_ERR_STATE,
                "this handle was created without a GPU (gpu_id < 0): host-side checks only");
}

template <typename R>
void pool_lane_parts(pasif_driver* d, bool adjoint);

void ensure_packed(pasif_driver* d, bool adjoint) {
  PASIF_REQUIRE(d->sys_ready, PASIF_ERR_STATE, "the system must be allocated before solving");
  if (adjoint)
    PASIF_REQUIRE(d->adj_ready, PASIF_ERR_STATE, "the adjoint must be allocated before solving");
  const uint64_t key = pack_key(d, adjoint);
  if (d->pk.valid && d->pk.key == key) return;
  d->pk.valid = false;
  d->pk.release();
  if (!d->pool) {
    d->lane_cls.clear();
    d->big_sys.clear();
  }
  // the exact discrete adjoint runs on single-wave lane kernels, tiles and
  // the row-parallel path (not on pool jobs' lane modules)
  const bool exact = adjoint && d->adj_mode != 0;
  PASIF_REQUIRE(!(d->pool && d->kernel_path == PASIF_KERNEL_BIG), PASIF_ERR_UNSUPPORTED,
                "pool jobs run on the tile and lane kernels");
  const uint32_t S = uint32_t(d->fwd.size.size());
  if (d->pool && exact) d->lane_cls.clear();  // the pool's lane modules run the reference recursion
  if (d->kernel_path == PASIF_KERNEL_BIG) {
    d->big_sys.assign(S, 1);
  } else if (d->kernel_path != PASIF_KERNEL_TILE && !d->pool) {
    if (d->precision == 64)
      (void)lane_plan<double>(d, adjoint);
    else
      (void)lane_plan<float>(d, adjoint);
  } else if (d->pool && !d->lane_cls.empty() && !exact) {
    if (d->precision == 64)
      pool_lane_parts<double>(d, adjoint);
    else
      pool_lane_parts<float>(d, adjoint);
  }
  // under the automatic choice, systems the lane kernel did not take that do
  // not fit a tile run on the row-parallel path (not for pool jobs)
  if (!d->pool && d->kernel_path == PASIF_KERNEL_AUTO) {
    d->big_sys.assign(S, 0);
    bool any = false;
    for (uint32_t s = 0; s < S; ++s)
      if (in_part(d, s, kPartTile) && !fits_tile(d, s, adjoint)) {
        d->big_sys[s] = 1;
        any = true;
      }
    if (!any) d->big_sys.clear();
  }
  if (!d->big_sys.empty()) {
    if (d->precision == 64)
      pack_big<double>(d, adjoint);
    else
      pack_big<float>(d, adjoint);
  }
  // the tile kernel takes every other system
  bool rest = false;
  for (uint32_t s = 0; s < S && !rest; ++s) rest = in_part(d, s, kPartTile);
  if (rest) {
    if (d->precision == 64)
      pack_impl<double>(d, adjoint);
    else
      pack_impl<float>(d, adjoint);
    d->pk.tile = true;
  } else {
    d->pk.ntiles = 0;
    d->pk.tmax = 0;
  }
  Packed& pk = d->pk;
  pk.tmax = std::max(pk.tmax, std::max(pk.ltmax, pk.b_tmax));
  if (pk.lane && !pk.tile) {
    pk.ntiles = 0;  // reported: the lane launch's workgroups
    for (const auto& lp : pk.lparts) pk.ntiles += lp->blocks;
    pk.lanes = kLaneWave;
    pk.slot_cfg = 0;
  }
  pk.adjoint = adjoint;
  pk.precision = d->precision;
  pk.valid = true;
  d->pk.key = key;
}

template <typename R>
void ensure_excitations(pasif_driver* d) {
  if (d->ext_exc) return;  // the pool's signal bank
  if (!d->d_exc_dirty && d->d_exc_prec == d->precision) return;
  std::vector<R> e(d->exc.begin(), d->exc.end());
  upload(d, d->d_exc, e);
  std::vector<R> m(d->interp.begin(), d->interp.end());
  upload(d, d->d_interp, m);
  PASIF_HIP(hipStreamSynchronize(d->stream));
  d->d_exc_dirty = false;
  d->d_exc_prec = d->precision;
}

template <typename R>
void fill_program(Program<R>& p, const DevBuf& info, const DevBuf& out, const DevBuf& src,
                  const DevBuf& steps, const DevBuf& exc, const DevBuf& fslot,
                  const DevBuf& force, const DevBuf& val, const DevBuf& idx, const DevBuf& init) {
  p.info = info.as<uint32_t>();
  p.out = out.as<uint32_t>();
  p.src = src.as<uint32_t>();
  p.steps = steps.as<uint32_t>();
  p.exc = exc.as<int32_t>();
  p.fslot = fslot.as<uint32_t>();
  p.force = force.as<R>();
  p.val = val.as<R>();
  p.idx = idx.as<uint32_t>();
  p.init = init.as<R>();
}

// Run buffers of one solve, without the outputs: the driver's own device
// buffers, allocated on first use and reused by later solves, or -- for
// pasif_solve_noalloc -- consecutive 256-byte aligned pieces of the library's
// static scratch array.  A sizing pass only adds up the bytes.
struct RunMem {
  pasif_driver* d = nullptr;
  unsigned char* base = nullptr;  // scratch region (null: the driver's buffers)
  size_t cap = 0, used = 0;
  size_t ckpt = 0;                // bytes of checkpoint buffers among `used`
  bool sizing = false;
  void* get(DevBuf& own, size_t bytes, bool checkpoint = false) {
    if (checkpoint) ckpt += bytes;
    if (!base && !sizing) {
      dev_reserve(d, own, bytes);
      return own.p;
    }
    used = (used + 255) & ~size_t(255);
    void* p = sizing ? nullptr : base + used;
    used += bytes ? bytes : 16;
    PASIF_REQUIRE(sizing || used <= cap, PASIF_ERR_MEMORY, "run scratch overflow");
    return p;
  }
  void zero(void* p, size_t bytes) {
    if (!sizing) PASIF_HIP(hipMemsetAsync(p, 0, bytes, d->stream));
  }
};

// Arguments of a run of the packed tiles with the run buffers from `mem`
// (progress cleared) and the given outputs.  *work is the number of steps
// of work each tile needs (3T with the adjoint: forward, recompute, reverse).
template <typename R>
TileArgs<R> tile_run_args(pasif_driver* d, bool adjoint, uint32_t save_steps, RunMem& mem,
                          void* out_q, void* out_max, void* traj, void* out_grad,
                          uint64_t* work_out, uint32_t* nslots_out) {
  Packed& pk = d->pk;
  const uint32_t L = pk.lanes, nt = pk.ntiles;
  const size_t rows = size_t(nt) * L;
  TileArgs<R> a{};
  a.hdr = pk.hdr.as<TileHeader>();
  a.lanes = L;
  fill_program(a.fwd, pk.f_info, pk.f_out, pk.f_src, pk.f_steps, pk.f_exc, pk.f_fslot,
               pk.f_force, pk.f_val, pk.f_idx, pk.f_init);
  if (adjoint)
    fill_program(a.adj, pk.a_info, pk.a_out, pk.a_src, pk.a_steps, pk.a_exc, pk.a_fslot,
                 pk.a_force, pk.a_val, pk.a_idx, pk.a_init);
  a.max_forced = pk.max_forced;
  a.excit = d->ext_exc ? static_cast<const R*>(d->ext_exc) : d->d_exc.as<R>();
  a.exc_len = d->exc_len;
  a.interp = d->interp_w ? (d->ext_interp ? static_cast<const R*>(d->ext_interp)
                                          : d->d_interp.as<R>())
                         : nullptr;
  a.interp_w = d->interp_w;
  // h = 1/sampleRate in the working type, h2 and h6 from that h in double
  // (src/__GpuDriver.cu:632-636)
  const R h = R(1.0 / double(d->sample_rate));
  a.h = h;
  a.h2 = R(double(h) / 2.0);
  a.h6 = R(double(h) / 6.0);
  a.qstate = static_cast<R*>(mem.get(d->qstate, rows * sizeof(R)));
  a.progress = static_cast<uint32_t*>(mem.get(d->progress, size_t(nt) * 4 * sizeof(uint32_t)));
  mem.zero(a.progress, size_t(nt) * 4 * sizeof(uint32_t));
  a.window = d->window;
  if (d->window) {
    a.mxstate = static_cast<R*>(mem.get(d->mxstate, rows * sizeof(R)));
    a.out_max = static_cast<R*>(out_max);
  }
  a.out_q = static_cast<R*>(out_q);
  a.out_len = pk.out_len_f;
  uint32_t nslots = 0;
  if (save_steps) {
    nslots = pk.tmax / save_steps;
    a.traj = static_cast<R*>(traj);
    a.save_every = save_steps;
    a.nslots = nslots;
  }
  uint64_t work = pk.tmax;
  if (adjoint) {
    const uint32_t K = pk.ck_every;
    a.lstate = static_cast<R*>(mem.get(d->lstate, rows * sizeof(R)));
    a.ckpt = static_cast<R*>(mem.get(d->ckpt, size_t(pk.ck_slots) * L * sizeof(R), true));
    a.ck_tile = pk.ck_tile.as<uint32_t>();
    a.exact = d->adj_mode != 0 ? 1u : 0u;
    a.segbuf = static_cast<R*>(mem.get(d->segbuf, rows * (K + 1) * (a.exact ? 4 : 1) * sizeof(R)));
    a.out_grad = static_cast<R*>(out_grad);
    a.ck_every = K;
    a.adjoint = 1;
    work = 3ull * pk.tmax;
  }
  *work_out = work;
  *nslots_out = nslots;
  return a;
}

// Arguments of a run of the row-parallel instances with the run buffers
// from `mem` (progress cleared) and the given outputs; *work as for tiles.
template <typename R>
BigArgs<R> big_run_args(pasif_driver* d, bool adjoint, RunMem& mem, void* out_q,
                        void* out_grad, uint64_t* work_out) {
  Packed& pk = d->pk;
  BigArgs<R> a{};
  a.inst = pk.b_inst.as<BigInst>();
  a.fwd = {pk.b_rf.as<uint32_t>(), pk.b_vf.as<R>(), pk.b_xf.as<uint32_t>(), pk.b_ff.as<R>(),
           pk.b_if.as<R>(), pk.b_sgf.as<uint32_t>(), pk.b_spf.as<uint32_t>()};
  if (adjoint)
    a.adj = {pk.b_ra.as<uint32_t>(), pk.b_va.as<R>(), pk.b_xa.as<uint32_t>(), nullptr,
             pk.b_ia.as<R>(), pk.b_sga.as<uint32_t>(), pk.b_spa.as<uint32_t>()};
  a.excit = d->ext_exc ? static_cast<const R*>(d->ext_exc) : d->d_exc.as<R>();
  a.exc_len = d->exc_len;
  a.interp = d->interp_w ? (d->ext_interp ? static_cast<const R*>(d->ext_interp)
                                          : d->d_interp.as<R>())
                         : nullptr;
  a.interp_w = d->interp_w;
  const R h = R(1.0 / double(d->sample_rate));
  a.h = h;
  a.h2 = R(double(h) / 2.0);
  a.h6 = R(double(h) / 6.0);
  const size_t el = sizeof(R);
  a.vec_f = static_cast<R*>(mem.get(d->bvec_f, size_t(pk.b_vec_f) * el));
  // the partial sums of one side at a time live in LDS when the largest fits
  a.part_lds = size_t(pk.b_part_max) * el <= kBigPartLds ? 1u : 0u;
  a.part = a.part_lds ? nullptr
                      : static_cast<R*>(mem.get(d->bpart, size_t(std::max<uint64_t>(pk.b_part, 1)) * el));
  a.progress = static_cast<uint32_t*>(mem.get(d->bprog, size_t(pk.nbig) * 4 * sizeof(uint32_t)));
  mem.zero(a.progress, size_t(pk.nbig) * 4 * sizeof(uint32_t));
  a.out_q = static_cast<R*>(out_q);
  a.ck_every = pk.b_every;
  uint64_t work = pk.b_tmax;
  if (adjoint) {
    a.vec_a = static_cast<R*>(mem.get(d->bvec_a, size_t(pk.b_vec_a) * el));
    a.setp = static_cast<R*>(mem.get(d->bsetp, size_t(pk.b_vec_f) * el));
    a.ckpt = static_cast<R*>(mem.get(d->bckpt, size_t(pk.b_ck) * el, true));
    a.segbuf = static_cast<R*>(mem.get(d->bseg, size_t(pk.b_seg) * el));
    a.out_grad = static_cast<R*>(out_grad);
    a.adjoint = 1;
    a.exact = d->adj_mode != 0 ? 1u : 0u;
    work = 3ull * pk.b_tmax;
  }
  *work_out = work;
  return a;
}

template <typename R>
uint32_t run_outputs(pasif_driver* d, bool adjoint, uint32_t save_steps);

// Pack (if needed), allocate the outputs and run buffers and clear the tile
// progress: the arguments of a run of the packed tiles.
template <typename R>
TileArgs<R> prepare_run(pasif_driver* d, bool adjoint, uint32_t save_steps, uint64_t* work_out,
                        uint32_t* nslots_out) {
  ensure_packed(d, adjoint);
  (void)run_outputs<R>(d, adjoint, save_steps);
  RunMem mem;
  mem.d = d;
  return tile_run_args<R>(d, adjoint, save_steps, mem, d->out_q.p, d->out_max.p, d->traj.p,
                          adjoint ? d->out_grad.p : nullptr, work_out, nslots_out);
}

// Copy the results of a finished run back to the host.
template <typename R>
void finish_run(pasif_driver* d, bool adjoint, uint32_t nslots, std::vector<double>* traj_out,
                std::vector<double>* grad_out) {
  Packed& pk = d->pk;
  std::vector<R> hq(pk.out_len_f);
  PASIF_HIP(hipMemcpy(hq.data(), d->out_q.p, hq.size() * sizeof(R), hipMemcpyDeviceToHost));
  d->res_q.assign(hq.begin(), hq.end());
  d->have_max = d->window != 0;
  if (d->window) {
    std::vector<R> hm(pk.out_len_f);
    PASIF_HIP(hipMemcpy(hm.data(), d->out_max.p, hm.size() * sizeof(R), hipMemcpyDeviceToHost));
    d->res_max.assign(hm.begin(), hm.end());
  }
  if (traj_out) {
    std::vector<R> ht(size_t(nslots) * pk.out_len_f);
    if (!ht.empty())
      PASIF_HIP(hipMemcpy(ht.data(), d->traj.p, ht.size() * sizeof(R), hipMemcpyDeviceToHost));
    traj_out->assign(ht.begin(), ht.end());
    d->traj.release();

  }
  if (grad_out) {
    std::vector<R> hg(pk.out_len_a);
    PASIF_HIP(hipMemcpy(hg.data(), d->out_grad.p, hg.size() * sizeof(R), hipMemcpyDeviceToHost));
    grad_out->assign(hg.begin(), hg.end());

  }
}


// Steps of work per launch for the solver mode, and the launches that cover
// `work` steps (a launch on finished work returns at once).
uint32_t launch_budget(const pasif_driver* d, uint64_t work, uint64_t* launches) {
  if (d->solver == PASIF_SOLVER_PLAIN) {
    *launches = std::max<uint64_t>(work, 1);
    return 1u;
  }
  if (d->solver == PASIF_SOLVER_GRAPH) {
    const uint32_t b = std::max<uint32_t>(1, d->graph_slice);
    *launches = std::max<uint64_t>(1, (work + b - 1) / b);
    return b;
  }
  *launches = 1;
  return 0xFFFFFFFFu;
}

// Issue `launches` calls of launch() on the driver's stream -- directly, or
// for the graph solver as replays of a captured graph of up to
// graph_launches calls -- and time them with HIP events.  No memset is
// captured: run state is cleared before, outside the graph.
template <class F>
void drive(pasif_driver* d, uint64_t launches, F&& launch) {
  const bool graph = d->solver == PASIF_SOLVER_GRAPH;
  hipGraphExec_t gexec = nullptr;
  uint32_t per_graph = 1;
  if (graph) {
    per_graph = uint32_t(std::min<uint64_t>(launches, d->graph_launches));
    hipGraph_t g = nullptr;
    PASIF_HIP(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t i = 0; i < per_graph; ++i) {
      hipError_t e = launch();
      if (e != hipSuccess) {
        (void)hipStreamEndCapture(d->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        set_err(d, "kernel capture failed: %s", hipGetErrorName(e));
        throw Fail{PASIF_ERR_DEVICE};
      }
    }
    PASIF_HIP(hipStreamEndCapture(d->stream, &g));
    hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      set_err(d, "graph instantiation failed: %s", hipGetErrorName(e));
      throw Fail{PASIF_ERR_DEVICE};
    }
  }
  PASIF_HIP(hipEventRecord(d->ev0, d->stream));
  if (graph) {
    const uint64_t replays = (launches + per_graph - 1) / per_graph;
    for (uint64_t i = 0; i < replays; ++i) {
      hipError_t e = hipGraphLaunch(gexec, d->stream);
      if (e != hipSuccess) {
        (void)hipGraphExecDestroy(gexec);
        set_err(d, "graph launch failed: %s", hipGetErrorName(e));
        throw Fail{PASIF_ERR_DEVICE};
      }
    }
    launches = replays * per_graph;
  } else {
    for (uint64_t i = 0; i < launches; ++i) {
      hipError_t e = launch();
      if (e != hipSuccess) {
        set_err(d, "kernel launch failed: %s", hipGetErrorName(e));
        throw Fail{PASIF_ERR_DEVICE};
      }
    }
  }
  PASIF_HIP(hipEventRecord(d->ev1, d->stream));
  PASIF_HIP(hipEventSynchronize(d->ev1));
  if (gexec) (void)hipGraphExecDestroy(gexec);
  float ms = 0.f;
  PASIF_HIP(hipEventElapsedTime(&ms, d->ev0, d->ev1));
  d->last_ms = ms;
  d->last_launches = uint32_t(launches);
}

// Output buffers of a solve (written by the lane and the tile parts); the
// device excitations in the working precision.
template <typename R>
uint32_t run_outputs(pasif_driver* d, bool adjoint, uint32_t save_steps) {
  ensure_excitations<R>(d);
  Packed& pk = d->pk;
  dev_reserve(d, d->out_q, size_t(pk.out_len_f) * sizeof(R));
  if (d->window) {
    dev_reserve(d, d->out_max, size_t(pk.out_len_f) * sizeof(R));
    PASIF_HIP(hipMemsetAsync(d->out_max.p, 0, size_t(pk.out_len_f) * sizeof(R), d->stream));
  }
  uint32_t nslots = 0;
  if (save_steps) {
    nslots = pk.tmax / save_steps;
    dev_reserve(d, d->traj, size_t(nslots) * pk.out_len_f * sizeof(R));
    PASIF_HIP(hipMemsetAsync(d->traj.p, 0, size_t(nslots) * pk.out_len_f * sizeof(R), d->stream));
  }
  if (adjoint) dev_reserve(d, d->out_grad, size_t(pk.out_len_a) * sizeof(R));
  return nslots;
}

// Run buffers of one lane part and its per-class kernel arguments (budget
// filled in by the caller).  K is the part's checkpoint interval.  Buffers
// come from `mem` (null: the driver's own), outputs are the driver's unless
// given.
template <typename R>
std::vector<LaneArgs> lane_part_args(pasif_driver* d, LanePart& P, bool adjoint,
                                     uint32_t save_steps, uint32_t nslots, const void* excit,
                                     const void* interp, RunMem* mem_in = nullptr,
                                     void* out_q = nullptr, void* out_grad = nullptr) {
  Packed& pk = d->pk;
  RunMem own;
  own.d = d;
  RunMem& mem = mem_in ? *mem_in : own;
  if (!out_q) out_q = d->out_q.p;
  if (!out_grad && adjoint) out_grad = d->out_grad.p;
  const R h = R(1.0 / double(d->sample_rate));
  std::vector<LaneArgs> args(P.cls.size());
  for (size_t c = 0; c < P.cls.size(); ++c) {
    LaneCls& L = *P.cls[c];
    const uint32_t K = adjoint ? L.seg : 1u;
    const size_t st = L.stride;
    LaneArgs& a = args[c];
    std::memset(&a, 0, sizeof a);
    a.qstate = mem.get(L.qstate, size_t(L.nf) * st * sizeof(R));
    a.progress = mem.get(L.progress, size_t(L.waves) * 4 * sizeof(uint32_t));
    mem.zero(a.progress, size_t(L.waves) * 4 * sizeof(uint32_t));
    if (d->window) a.mxstate = mem.get(L.mxstate, size_t(L.nf) * st * sizeof(R));
    uint32_t nck = 1;
    if (adjoint) {
      nck = std::max<uint32_t>(1, (L.tmax + K - 1) / K);
      a.lstate = mem.get(L.lstate, size_t(std::max<uint32_t>(L.na, 1)) * st * sizeof(R));
      if (P.lay.waves == 1)
        a.ckpt = mem.get(L.ckpt, size_t(L.ck_elems) * sizeof(R), true);
      else
        a.ckpt = mem.get(L.ckpt, size_t(nck) * L.nf * st * sizeof(R), true);
      a.segbuf = mem.get(L.segbuf, lane_seg_values(K, P.lay.exact) * L.nf * st * sizeof(R));
    }
    a.coef_f = L.coef_f;
    a.coef_a = L.coef_a;
    a.init_f = L.init_f;
    a.init_a = L.init_a;
    a.steps = L.steps;
    a.exc = L.exc;
    a.out_f = L.out_f;
    a.out_a = L.out_a;
    a.na_out = L.na_out;
    a.ck_off = L.ckoff;
    a.out_q = out_q;
    a.out_max = d->window ? d->out_max.p : nullptr;
    a.traj = save_steps ? d->traj.p : nullptr;
    a.out_grad = adjoint ? out_grad : nullptr;
    a.excit = excit;
    a.interp = d->interp_w ? interp : nullptr;
    a.stride = st;
    a.count = L.count;
    a.block_begin = L.block_begin;
    a.klass = L.klass;
    a.exc_len = d->exc_len;
    a.interp_w = d->interp_w;
    a.window = d->window;
    a.out_len = pk.out_len_f;
    a.save_every = save_steps;
    a.nslots = nslots;
    a.ck_every = K;
    a.nck = nck;
    a.adjoint = adjoint ? 1 : 0;
    // h = 1/sampleRate in the working type, h2 and h6 from that h in double
    // (src/__GpuDriver.cu:632-636), carried exactly in a double
    a.h = double(h);
    a.h2 = double(R(double(h) / 2.0));
    a.h6 = double(R(double(h) / 6.0));
  }
  return args;
}

hipError_t lane_launch(const LanePart& P, const void* dargs, uint32_t n, hipStream_t st) {
  // multi-wave kernels declare their LDS statically
  const uint32_t lds = P.lay.waves > 1 ? 0u : uint32_t(P.lds);
  void* dp = const_cast<void*>(dargs);
  void* params[] = {&dp, &n};
  return hipModuleLaunchKernel(P.kern.fn, P.blocks, 1, 1, kLaneWave * P.lay.waves, 1, 1, lds,
                               st, params, nullptr);
}

void ensure_aux(pasif_driver* d, int n) {
  if (!d->fork) PASIF_HIP(hipEventCreateWithFlags(&d->fork, hipEventDisableTiming));
  for (int i = 0; i < n; ++i) {
    if (d->aux[i]) continue;
    PASIF_HIP(hipStreamCreateWithFlags(&d->aux[i], hipStreamNonBlocking));
    PASIF_HIP(hipEventCreateWithFlags(&d->join[i], hipEventDisableTiming));
  }
}

// Device bytes the run buffers of a solve of the packed problem take
// (outputs included; the packed program itself is not), and of those the
// checkpoints: a sizing pass over the same requests a solve makes.
template <typename R>
void run_footprint(pasif_driver* d, bool adjoint, uint32_t save_steps, size_t* total,
                   size_t* ckpt) {
  Packed& pk = d->pk;
  RunMem m;
  m.d = d;
  m.sizing = true;
  DevBuf unused;
  uint64_t work = 0;
  uint32_t nslots = save_steps ? pk.tmax / save_steps : 0;
  m.get(unused, size_t(pk.out_len_f) * sizeof(R));
  if (d->window) m.get(unused, size_t(pk.out_len_f) * sizeof(R));
  if (nslots) m.get(unused, size_t(nslots) * pk.out_len_f * sizeof(R));
  if (adjoint) m.get(unused, size_t(pk.out_len_a) * sizeof(R));
  if (pk.tile)
    (void)tile_run_args<R>(d, adjoint, save_steps, m, nullptr, nullptr, nullptr, nullptr, &work,
                           &nslots);
  if (pk.big) (void)big_run_args<R>(d, adjoint, m, nullptr, nullptr, &work);
  for (auto& lp : pk.lparts) {
    auto args = lane_part_args<R>(d, *lp, adjoint, save_steps, nslots, nullptr, nullptr, &m,
                                  nullptr, nullptr);
    m.get(unused, args.size() * sizeof(LaneArgs));
  }
  *total = m.used;
  *ckpt = m.ckpt;
}

// Bytes of run buffers (and outputs) the driver holds now: a solve reuses
// them, so they count as available to it.
size_t held_run_bytes(const pasif_driver* d) {
  size_t b = 0;
  const DevBuf* own[] = {&d->qstate, &d->lstate, &d->mxstate, &d->progress, &d->out_q,
                         &d->out_max, &d->traj,   &d->out_grad, &d->ckpt,  &d->segbuf,
                         &d->bvec_f, &d->bvec_a,  &d->bsetp,    &d->bckpt, &d->bseg,
                         &d->bprog,  &d->bpart};
  for (const DevBuf* x : own) b += x->bytes;
  for (const auto& lp : d->pk.lparts)
    for (const auto& c : lp->cls)
      b += c->qstate.bytes + c->lstate.bytes + c->mxstate.bytes + c->progress.bytes +
           c->ckpt.bytes + c->segbuf.bytes;
  return b;
}

// The memory plan of a solve: its run buffers must fit in 80 % of the device
// memory free now plus what the driver already holds for them (the 0.8 x
// free rule of the reference's planner, src/__GpuDriver.cu:896-938), or the
// solve is refused before anything is allocated.
template <typename R>
void check_run_memory(pasif_driver* d, bool adjoint, uint32_t save_steps) {
  if (d->dev < 0 || d->pool) return;
  size_t need = 0, ck = 0, freeb = 0, totalb = 0;
  run_footprint<R>(d, adjoint, save_steps, &need, &ck);
  PASIF_HIP(hipMemGetInfo(&freeb, &totalb));
  const double avail = 0.8 * double(freeb + held_run_bytes(d));
  PASIF_REQUIRE(double(need) <= avail, PASIF_ERR_MEMORY,
                "the solve needs %zu bytes of run buffers (%zu of them checkpoints), more than "
                "80%% of the free device memory (%.0f bytes available); lengthen the checkpoint "
                "interval or solve fewer systems at once",
                need, ck, avail);
}

// Run the packed problem to completion: the lane part, the tile part, or
// both side by side on two streams.  save_steps > 0 records a trajectory.
template <typename R>
void solve_impl(pasif_driver* d, bool adjoint, uint32_t save_steps, std::vector<double>* traj_out,
                std::vector<double>* grad_out) {
  ensure_packed(d, adjoint);
  Packed& pk = d->pk;
  const bool lane = pk.lane, tile = pk.tile, big = pk.big;
  PASIF_REQUIRE(!big || (save_steps == 0 && d->window == 0), PASIF_ERR_UNSUPPORTED,
                "systems on the row-parallel path (too large for a tile) record neither a "
                "trajectory nor an amplitude window");
  check_run_memory<R>(d, adjoint, save_steps);
  uint64_t twork = 0, bwork = 0;
  uint32_t nslots = 0;
  TileArgs<R> ta{};
  if (tile) ta = prepare_run<R>(d, adjoint, save_steps, &twork, &nslots);
  BigArgs<R> ba{};
  if (big) {
    (void)run_outputs<R>(d, adjoint, save_steps);
    RunMem mem;
    mem.d = d;
    ba = big_run_args<R>(d, adjoint, mem, d->out_q.p, adjoint ? d->out_grad.p : nullptr, &bwork);
  }
  std::vector<std::vector<LaneArgs>> largs;
  if (lane) {
    nslots = run_outputs<R>(d, adjoint, save_steps);
    for (auto& lp : pk.lparts)
      largs.push_back(lane_part_args<R>(d, *lp, adjoint, save_steps, nslots, d->d_exc.p,
                                        d->d_interp.p));
  }
  const uint64_t lwork = lane ? (adjoint ? 3ull * pk.ltmax : pk.ltmax) : 0;
  uint64_t launches = 1;
  const uint32_t budget = launch_budget(d, std::max(std::max(twork, lwork), bwork), &launches);
  ta.budget = budget;
  ba.budget = budget;
  for (size_t i = 0; i < largs.size(); ++i) {
    for (LaneArgs& a : largs[i]) a.budget = budget;
    upload(d, pk.lparts[i]->args, largs[i]);
  }
  // lane modules with work, each on a stream of its own where one is free:
  // the driver's stream when no tile or row-parallel launch runs on it, then
  // up to three side streams (modules beyond that share them in turn)
  std::vector<size_t> lmods;
  for (size_t i = 0; i < largs.size(); ++i)
    if (pk.lparts[i]->blocks) lmods.push_back(i);
  const bool main_busy = tile || big;
  const int nside = int(std::min<size_t>(pasif_driver::kSide,
                                         lmods.size() - (main_busy || lmods.empty() ? 0 : 1)));
  if (nside > 0) ensure_aux(d, nside);
  auto lane_stream = [&](size_t j) -> int {  // -1: the driver's stream
    if (!main_busy && j == 0) return -1;
    return int((j - (main_busy ? 0 : 1)) % size_t(nside));
  };
  const uint32_t nb = pk.nbig;
  const uint32_t nt = pk.ntiles;
  drive(d, launches, [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    if (nside > 0) {
      e = hipEventRecord(d->fork, d->stream);
      for (int k = 0; k < nside && e == hipSuccess; ++k)
        e = hipStreamWaitEvent(d->aux[k], d->fork, 0);
    }
    for (size_t j = 0; j < lmods.size() && e == hipSuccess; ++j) {
      const LanePart& P = *pk.lparts[lmods[j]];
      const int k = lane_stream(j);
      e = lane_launch(P, P.args.p, uint32_t(largs[lmods[j]].size()),
                      k < 0 ? d->stream : d->aux[k]);
    }
    if (tile && e == hipSuccess) e = launch_tiles(ta, nt, int(pk.slot_cfg), d->stream);
    if (big && e == hipSuccess)
      e = launch_big(ba, nb, ba.part_lds ? size_t(pk.b_part_max) * sizeof(R) : 0, d->stream);
    for (int k = 0; k < nside; ++k) {
      if (e == hipSuccess) e = hipEventRecord(d->join[k], d->aux[k]);
      if (e == hipSuccess) e = hipStreamWaitEvent(d->stream, d->join[k], 0);
    }
    return e;
  });
  finish_run<R>(d, adjoint, nslots, traj_out, grad_out);
}

// pasif_solve_noalloc: the packed problem run with every run buffer and the
// lane kernels' argument tables carved from the library's static scratch
// array, results written into the caller's device buffers.  Nothing is
// allocated or freed: the problem must already be packed for this mode and
// its excitations uploaded (the output-length query of pasif_get_amplitudes
// / pasif_get_gradient does both).  Lane and tile parts run one after the
// other on the driver's stream.
template <typename R>
void solve_noalloc_impl(pasif_driver* d, bool adjoint, void* out_q, void* out_grad) {
  Packed& pk = d->pk;
  PASIF_REQUIRE(pk.valid && pk.key == pack_key(d, adjoint), PASIF_ERR_STATE,
                "pasif_solve_noalloc runs a problem already packed for this mode and these "
                "settings: query the output length first (pasif_get_%s with out == NULL)",
                adjoint ? "gradient" : "amplitudes");
  PASIF_REQUIRE(d->ext_exc || (!d->d_exc_dirty && d->d_exc_prec == d->precision),
                PASIF_ERR_STATE,
                "the excitations are not on the device in this precision: query the output "
                "length first (pasif_get_%s with out == NULL)", adjoint ? "gradient" : "amplitudes");
  PASIF_REQUIRE(d->window == 0, PASIF_ERR_UNSUPPORTED,
                "pasif_solve_noalloc does not keep the amplitude window (set it to 0)");
  PASIF_REQUIRE(d->solver != PASIF_SOLVER_GRAPH, PASIF_ERR_UNSUPPORTED,
                "pasif_solve_noalloc runs the persistent or the plain solver, not the graph one");
  PASIF_REQUIRE(!d->pool, PASIF_ERR_UNSUPPORTED, "pool jobs run through pasif_pool_tick");
  PASIF_REQUIRE(!pk.big, PASIF_ERR_UNSUPPORTED,
                "pasif_solve_noalloc runs the tile and lane kernels, not the row-parallel path");
  PASIF_REQUIRE(out_q != nullptr, PASIF_ERR_ARGUMENT, "out_q (device, %u values) is required",
                pk.out_len_f);
  PASIF_REQUIRE(!adjoint || out_grad != nullptr, PASIF_ERR_ARGUMENT,
                "out_grad (device, %u values) is required for the adjoint", pk.out_len_a);
  void* base = nullptr;
  size_t cap = 0;
  PASIF_HIP(scratch_region(d->dev, &base, &cap));
  const bool lane = pk.lane, tile = pk.tile;
  DevBuf unused;
  TileArgs<R> ta{};
  uint64_t twork = 0;
  std::vector<std::vector<LaneArgs>> largs;
  std::vector<void*> tables;
  auto plan = [&](RunMem& mem) {
    largs.clear();
    tables.clear();
    uint32_t nslots = 0;
    if (tile)
      ta = tile_run_args<R>(d, adjoint, 0, mem, out_q, nullptr, nullptr, out_grad, &twork,
                            &nslots);
    if (lane)
      for (auto& lp : pk.lparts) {
        largs.push_back(lane_part_args<R>(d, *lp, adjoint, 0, 0, d->d_exc.p, d->d_interp.p,
                                          &mem, out_q, out_grad));
        tables.push_back(mem.get(unused, largs.back().size() * sizeof(LaneArgs)));
      }
  };
  RunMem sizing;
  sizing.d = d;
  sizing.sizing = true;
  plan(sizing);
  PASIF_REQUIRE(sizing.used <= cap, PASIF_ERR_MEMORY,
                "this problem needs %zu bytes of run buffers; pasif_solve_noalloc holds %zu "
                "bytes per GPU (solve it with pasif_get_%s instead)",
                sizing.used, cap, adjoint ? "gradient" : "amplitudes");
  RunMem mem;
  mem.d = d;
  mem.base = static_cast<unsigned char*>(base);
  mem.cap = cap;
  plan(mem);
  const uint64_t lwork = lane ? (adjoint ? 3ull * pk.ltmax : pk.ltmax) : 0;
  uint64_t launches = 1;
  const uint32_t budget = launch_budget(d, std::max(twork, lwork), &launches);
  ta.budget = budget;
  for (size_t i = 0; i < largs.size(); ++i) {
    for (LaneArgs& a : largs[i]) a.budget = budget;
    PASIF_HIP(hipMemcpyAsync(tables[i], largs[i].data(), largs[i].size() * sizeof(LaneArgs),
                             hipMemcpyHostToDevice, d->stream));
  }
  const uint32_t nt = pk.ntiles;
  drive(d, launches, [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < largs.size() && e == hipSuccess; ++i)
      if (pk.lparts[i]->blocks)
        e = lane_launch(*pk.lparts[i], tables[i], uint32_t(largs[i].size()), d->stream);
    if (tile && e == hipSuccess) e = launch_tiles(ta, nt, int(pk.slot_cfg), d->stream);
    return e;
  });
}

void solve(pasif_driver* d, bool adjoint, uint32_t save_steps, std::vector<double>* traj,
           std::vector<double>* grad) {
  require_device(d);
  PASIF_HIP(hipSetDevice(d->dev));
  if (d->precision == 64)
    solve_impl<double>(d, adjoint, save_steps, traj, grad);
  else
    solve_impl<float>(d, adjoint, save_steps, traj, grad);
}

void copy_out(pasif_driver* d, const std::vector<double>& v, double* out, size_t cap, size_t* len) {
  if (len) *len = v.size();
  if (!out) return;
  PASIF_REQUIRE(cap >= v.size(), PASIF_ERR_ARGUMENT, "output buffer holds %zu, need %zu", cap,
                v.size());
  std::copy(v.begin(), v.end(), out);
}

void load_exc(pasif_driver* d, const double* e, uint32_t n_exc, uint32_t len, uint32_t sr) {
  PASIF_REQUIRE(sr > 0, PASIF_ERR_ARGUMENT, "sample rate must be positive");
  PASIF_REQUIRE(n_exc > 0 && len > 0, PASIF_ERR_ARGUMENT, "empty excitation set");
  PASIF_REQUIRE(e != nullptr, PASIF_ERR_ARGUMENT, "null excitation data");
  d->exc.assign(e, e + size_t(n_exc) * len);
  d->n_exc = n_exc;
  d->exc_len = len;
  d->sample_rate = sr;
  d->d_exc_dirty = true;
  d->pk.valid = false;
}

void set_csr(pasif_driver* d, CsrIn& K, const uint32_t n[2], const double* v, const uint32_t* ix,
             const uint32_t* ptr, size_t nnz) {
  PASIF_REQUIRE(n != nullptr && ptr != nullptr, PASIF_ERR_ARGUMENT, "null CSR argument");
  PASIF_REQUIRE(nnz == 0 || (v && ix), PASIF_ERR_ARGUMENT, "null CSR values");
  K.n0 = n[0];
  K.n1 = n[1];
  K.val.assign(v, v + nnz);
  K.idx.assign(ix, ix + nnz);
  K.ptr.assign(ptr, ptr + size_t(n[0]) + 1);
  K.set = true;
}

void set_coo(pasif_driver* d, CooIn& t, int order, const uint32_t* n, const double* v,
             const uint32_t* ix, size_t nnz) {
  PASIF_REQUIRE(n != nullptr, PASIF_ERR_ARGUMENT, "null tensor dims");
  PASIF_REQUIRE(nnz == 0 || (v && ix), PASIF_ERR_ARGUMENT, "null tensor data");
  t.order = order;
  for (int a = 0; a < order; ++a) t.dims[a] = n[a];
  t.val.assign(v, v + nnz);
  t.idx.assign(ix, ix + nnz * order);
  t.set = true;
}

// pool job handles (defined with the pool)
void pool_remove_job(pasif_pool* P, size_t k);
void pool_detach(pasif_driver* d);
}  // namespace

// =================================================================== C ABI
extern "C" {

const char* pasif_version(void) { return "pasif_hip 0.1 gfx950"; }

int pasif_create(pasif_driver** out, const double* excitations, uint32_t n_exc, uint32_t exc_len,
                 uint32_t sample_rate, uint32_t numsteps, int d_compute, int d_system,
                 int d_solver, int gpu_id) {
  if (!out) return PASIF_ERR_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<pasif_driver> holder(new (std::nothrow) pasif_driver());
  if (!holder) return PASIF_ERR_MEMORY;
  pasif_driver* d = holder.get();
  const int rc = guarded(d, [&] {
    d->dev = gpu_id;
    if (gpu_id >= 0) {
      int ndev = 0;
      PASIF_HIP(hipGetDeviceCount(&ndev));
      PASIF_REQUIRE(gpu_id < ndev, PASIF_ERR_ARGUMENT, "GPU %d of %d", gpu_id, ndev);
      PASIF_HIP(hipSetDevice(gpu_id));
      PASIF_HIP(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
      PASIF_HIP(hipEventCreate(&d->ev0));
      PASIF_HIP(hipEventCreate(&d->ev1));
    }
    d->d_compute = d_compute != 0;
    d->d_system = d_system != 0;
    d->d_solver = d_solver != 0;
    load_exc(d, excitations, n_exc, exc_len, sample_rate);
    d->numsteps = numsteps ? numsteps : exc_len;
  });
  // on failure the handle is still returned so the caller can read the message
  *out = holder.release();
  return rc;
}

int pasif_destroy(pasif_driver* d) {
  if (!d) return PASIF_OK;
  if (d->pool) {
    // a pool job: leaves the pool (a resident one that a tick in flight
    // still runs is deleted by the pool once the tick is over)
    pasif_pool* P = d->pool;
    for (size_t k = 0; k < P->jobs.size(); ++k) {
      if (P->jobs[k].d != d) continue;
      if (P->in_flight && !P->jobs[k].finished) {
        d->orphan = true;
        return PASIF_OK;
      }
      (void)hipSetDevice(P->dev);
      pool_remove_job(P, k);
      break;
    }
    P->handles.erase(d);
    pool_detach(d);
  }
  if (d->dev >= 0) (void)hipSetDevice(d->dev);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  d->pk.release();
  for (int k = 0; k < pasif_driver::kSide; ++k) {
    if (!d->aux[k]) continue;
    (void)hipStreamSynchronize(d->aux[k]);
    (void)hipStreamDestroy(d->aux[k]);
    if (d->join[k]) (void)hipEventDestroy(d->join[k]);
  }
  if (d->fork) (void)hipEventDestroy(d->fork);
  if (!d->pool) {
    if (d->ev0) (void)hipEventDestroy(d->ev0);
    if (d->ev1) (void)hipEventDestroy(d->ev1);
    if (d->stream) (void)hipStreamDestroy(d->stream);
  }
  delete d;
  return PASIF_OK;
}

const char* pasif_last_error(const pasif_driver* d) { return d ? d->err.c_str() : "null handle"; }

int pasif_set_fwd_K(pasif_driver* d, const uint32_t n[2], const double* values,
                    const uint32_t* indices, const uint32_t* indptr, size_t nnz) {
  return guarded(d, [&] {
    set_csr(d, d->fin.K, n, values, indices, indptr, nnz);
    d->sys_ready = false;
  });
}
int pasif_set_bwd_K(pasif_driver* d, const uint32_t n[2], const double* values,
                    const uint32_t* indices, const uint32_t* indptr, size_t nnz) {
  return guarded(d, [&] {
    set_csr(d, d->bin.K, n, values, indices, indptr, nnz);
    d->adj_ready = false;
  });
}
int pasif_set_fwd_gamma(pasif_driver* d, const uint32_t n[3], const double* v, const uint32_t* ix,
                        size_t nnz) {
  return guarded(d, [&] { set_coo(d, d->fin.G, 3, n, v, ix, nnz); d->sys_ready = false; });
}
int pasif_set_fwd_lambda(pasif_driver* d, const uint32_t n[4], const double* v, const uint32_t* ix,
                         size_t nnz) {
  return guarded(d, [&] { set_coo(d, d->fin.L, 4, n, v, ix, nnz); d->sys_ready = false; });
}
int pasif_set_fwd_psi(pasif_driver* d, const uint32_t n[5], const double* v, const uint32_t* ix,
                      size_t nnz) {
  return guarded(d, [&] { set_coo(d, d->fin.P, 5, n, v, ix, nnz); d->sys_ready = false; });
}
int pasif_set_bwd_gamma(pasif_driver* d, const uint32_t n[3], const double* v, const uint32_t* ix,
                        size_t nnz) {
  return guarded(d, [&] { set_coo(d, d->bin.G, 3, n, v, ix, nnz); d->adj_ready = false; });
}
int pasif_set_bwd_lambda(pasif_driver* d, const uint32_t n[4], const double* v, const uint32_t* ix,
                         size_t nnz) {
  return guarded(d, [&] { set_coo(d, d->bin.L, 4, n, v, ix, nnz); d->adj_ready = false; });
}
int pasif_set_bwd_psi(pasif_driver* d, const uint32_t n[5], const double* v, const uint32_t* ix,
                      size_t nnz) {
  return guarded(d, [&] { set_coo(d, d->bin.P, 5, n, v, ix, nnz); d->adj_ready = false; });
}
int pasif_set_fwd_force_pattern(pasif_driver* d, const double* f, size_t n) {
  return guarded(d, [&] {
    PASIF_REQUIRE(n == 0 || f, PASIF_ERR_ARGUMENT, "null force pattern");
    d->fin.force.assign(f, f + n);
    d->fin.force_set = true;
    d->sys_ready = false;
  });
}
int pasif_set_fwd_initial_conditions(pasif_driver* d, const double* q0, size_t n) {
  return guarded(d, [&] {
    PASIF_REQUIRE(n == 0 || q0, PASIF_ERR_ARGUMENT, "null initial conditions");
    d->fin.init.assign(q0, q0 + n);
    d->fin.init_set = true;
    d->sys_ready = false;
  });
}
int pasif_set_bwd_force_pattern(pasif_driver* d, const double* f, size_t n) {
  return guarded(d, [&] {
    PASIF_REQUIRE(n == 0 || f, PASIF_ERR_ARGUMENT, "null force pattern");
    d->bin.force.assign(f, f + n);
    d->bin.force_set = true;
    d->adj_ready = false;
  });
}
int pasif_set_bwd_initial_conditions(pasif_driver* d, const double* l, size_t n) {
  return guarded(d, [&] {
    PASIF_REQUIRE(n == 0 || l, PASIF_ERR_ARGUMENT, "null terminal condition");
    d->bin.init.assign(l, l + n);
    d->bin.init_set = true;
    d->adj_ready = false;
  });
}

int pasif_allocate_system(pasif_driver* d) {
  return guarded(d, [&] {
    d->sys_ready = false;
    d->adj_ready = false;
    d->pk.valid = false;
    build_rows(d, d->fin, d->fin.K.n0, false, d->fwd, "system");
    std::vector<uint32_t> sizes = d->hint_f.empty() ? detect_blocks(d->fwd, true) : d->hint_f;
    apply_partition(d, d->fwd, sizes, "system");
    validate_blocks(d, d->fwd, nullptr, "system");
    if (d->mapped || !d->map_steps.empty()) {
      const size_t S = d->fwd.size.size();
      PASIF_REQUIRE(d->layout_systems == S, PASIF_ERR_ARGUMENT,
                    "the layout describes %u systems, the system has %zu", d->layout_systems, S);
    }
    d->sys_ready = true;
    if (d->d_system)
      printf("[pasif] system: %u DOFs in %zu independent systems\n", d->fwd.n,
             d->fwd.size.size());
  });
}

int pasif_allocate_adjoint(pasif_driver* d) {
  return guarded(d, [&] {
    d->adj_ready = false;
    d->pk.valid = false;
    PASIF_REQUIRE(d->sys_ready, PASIF_ERR_STATE, "the system must be set before the jacobian");
    build_rows(d, d->bin, d->fwd.n, true, d->adj, "adjoint");
    std::vector<uint32_t> sizes = d->hint_a.empty() ? detect_blocks(d->adj, false) : d->hint_a;
    PASIF_REQUIRE(sizes.size() == d->fwd.size.size(), PASIF_ERR_ARGUMENT,
                  "the adjoint splits into %zu blocks but the system into %zu; pass the "
                  "per-system adjoint sizes",
                  sizes.size(), d->fwd.size.size());
    apply_partition(d, d->adj, sizes, "adjoint");
    d->adj.xoff = d->fwd.off;
    RowTerms xs;
    xs.off = d->fwd.off;
    xs.size = d->fwd.size;
    validate_blocks(d, d->adj, &xs, "adjoint");
    d->adj_ready = true;
  });
}

int pasif_load_excitations(pasif_driver* d, const double* e, uint32_t n_exc, uint32_t len,
                           uint32_t sr) {
  return guarded(d, [&] {
    PASIF_REQUIRE(!d->ext_exc, PASIF_ERR_STATE, "this job uses its pool's signal bank");
    load_exc(d, e, n_exc, len, sr);
  });
}

int pasif_set_interpolation_matrix(pasif_driver* d, const double* m, size_t size, uint32_t w) {
  return guarded(d, [&] {
    PASIF_REQUIRE(!d->ext_interp && !d->ext_exc, PASIF_ERR_STATE,
                  "this job uses its pool's signal bank and interpolation");
    if (size == 0) {
      d->interp.clear();
      d->interp_w = 0;
    } else {
      PASIF_REQUIRE(m != nullptr, PASIF_ERR_ARGUMENT, "null interpolation matrix");
      PASIF_REQUIRE(w >= 2, PASIF_ERR_ARGUMENT, "interpolation window must be >= 2 (got %u)", w);
      PASIF_REQUIRE(w <= kMaxInterpWindow, PASIF_ERR_UNSUPPORTED,
                    "interpolation window %u exceeds %u", w, kMaxInterpWindow);
      PASIF_REQUIRE(size % w == 0, PASIF_ERR_ARGUMENT,
                    "interpolation matrix size %zu is not a multiple of the window %u", size, w);
      PASIF_REQUIRE(size / w >= 2, PASIF_ERR_ARGUMENT,
                    "interpolation matrix needs at least 2 rows (rows 0 and 1 are read)");
      d->interp.assign(m, m + 2 * size_t(w));
      d->interp_w = w;
    }
    d->d_exc_dirty = true;
  });
}

int pasif_set_modulation_buffer(pasif_driver* d, const double* b, size_t n) {
  return guarded(d, [&] {
    PASIF_REQUIRE(n > 0 && b, PASIF_ERR_ARGUMENT, "The modulation buffer must be non-empty.");
    PASIF_REQUIRE(n <= 32000, PASIF_ERR_ARGUMENT,
                  "The modulation buffer don't fit in the GPU cst memory.");
    d->modulation.assign(b, b + n);
  });
}

void ensure_excitations_any(pasif_driver* d) {
  if (d->dev < 0) return;  // host-only handle: nothing to upload
  if (d->precision == 64)
    ensure_excitations<double>(d);
  else
    ensure_excitations<float>(d);
}

int pasif_get_amplitudes(pasif_driver* d, double* out, size_t cap, size_t* len) {
  return guarded(d, [&] {
    if (!out) {  // length query without solving; packs and uploads the problem
      ensure_packed(d, false);
      ensure_excitations_any(d);
      if (len) *len = d->pk.out_len_f;
      return;
    }
    solve(d, false, 0, nullptr, nullptr);
    copy_out(d, d->res_q, out, cap, len);
  });
}

int pasif_get_trajectory(pasif_driver* d, uint32_t save_steps, double* out, size_t cap,
                         size_t* len) {
  return guarded(d, [&] {
    PASIF_REQUIRE(save_steps > 0, PASIF_ERR_ARGUMENT, "saveSteps must be positive");
    if (!out) {
      ensure_packed(d, false);
      if (len) *len = size_t(d->pk.tmax / save_steps) * d->pk.out_len_f;
      return;
    }
    std::vector<double> traj;
    solve(d, false, save_steps, &traj, nullptr);
    copy_out(d, traj, out, cap, len);
  });
}

int pasif_get_gradient(pasif_driver* d, uint32_t save, double* out, size_t cap, size_t* len) {
  return guarded(d, [&] {
    PASIF_REQUIRE(save == 0, PASIF_ERR_UNSUPPORTED,
                  "partial gradients (save != 0) are not supported; the full schedule runs");
    if (!out) {  // length query without solving; packs and uploads the problem
      ensure_packed(d, true);
      ensure_excitations_any(d);
      if (len) *len = d->pk.out_len_a;
      return;
    }
    std::vector<double> g;
    solve(d, true, 0, nullptr, &g);
    copy_out(d, g, out, cap, len);
  });
}

int pasif_set_adjoint_mode(pasif_driver* d, int mode) {
  return guarded(d, [&] {
    PASIF_REQUIRE(mode == 0 || mode == 1, PASIF_ERR_ARGUMENT,
                  "adjoint mode %d (0 reference recursion, 1 exact discrete adjoint)", mode);
    d->adj_mode = mode;
  });
}

int pasif_run_footprint(pasif_driver* d, int adjoint, size_t* run_bytes, size_t* ckpt_bytes,
                        size_t* free_bytes) {
  return guarded(d, [&] {
    ensure_packed(d, adjoint != 0);
    size_t total = 0, ck = 0;
    if (d->precision == 64)
      run_footprint<double>(d, adjoint != 0, 0, &total, &ck);
    else
      run_footprint<float>(d, adjoint != 0, 0, &total, &ck);
    if (run_bytes) *run_bytes = total;
    if (ckpt_bytes) *ckpt_bytes = ck;
    if (free_bytes) {
      *free_bytes = 0;
      if (d->dev >= 0) {
        size_t f = 0, t = 0;
        PASIF_HIP(hipSetDevice(d->dev));
        PASIF_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = f + held_run_bytes(d);
      }
    }
  });
}

int pasif_solve_noalloc(pasif_driver* d, int adjoint, void* out_q, void* out_grad) {
  return guarded(d, [&] {
    require_device(d);
    PASIF_HIP(hipSetDevice(d->dev));
    struct Scope {
      Scope() { t_no_alloc = true; }
      ~Scope() { t_no_alloc = false; }
    } scope;
    if (d->precision == 64)
      solve_noalloc_impl<double>(d, adjoint != 0, out_q, out_grad);
    else
      solve_noalloc_impl<float>(d, adjoint != 0, out_q, out_grad);
  });
}

int pasif_set_systems_layout(pasif_driver* d, uint32_t n_systems, const uint32_t* sizes_fwd,
                             const uint32_t* sizes_adj, const int32_t* exc_of_system,
                             const uint32_t* steps_of_system) {
  return guarded(d, [&] {
    d->hint_f.clear();
    d->hint_a.clear();
    d->map_exc.clear();
    d->map_steps.clear();
    d->mapped = false;
    d->layout_systems = n_systems;
    if (sizes_fwd) d->hint_f.assign(sizes_fwd, sizes_fwd + n_systems);
    if (sizes_adj) d->hint_a.assign(sizes_adj, sizes_adj + n_systems);
    if (exc_of_system) {
      d->map_exc.assign(exc_of_system, exc_of_system + n_systems);
      d->mapped = true;
    }
    if (steps_of_system) d->map_steps.assign(steps_of_system, steps_of_system + n_systems);
    d->pk.valid = false;
    if (d->sys_ready && !d->hint_f.empty()) {
      apply_partition(d, d->fwd, d->hint_f, "system");
      validate_blocks(d, d->fwd, nullptr, "system");
    }
  });
}

int pasif_set_precision(pasif_driver* d, int bits) {
  return guarded(d, [&] {
    PASIF_REQUIRE(bits == 64 || bits == 32, PASIF_ERR_ARGUMENT, "precision must be 32 or 64");
    d->precision = bits;
  });
}

int pasif_set_solver(pasif_driver* d, int solver) {
  return guarded(d, [&] {
    PASIF_REQUIRE(solver == PASIF_SOLVER_PERSISTENT || solver == PASIF_SOLVER_PLAIN ||
                      solver == PASIF_SOLVER_GRAPH,
                  PASIF_ERR_ARGUMENT, "unknown solver %d", solver);
    d->solver = solver;
  });
}

int pasif_set_tuning(pasif_driver* d, uint32_t tile_lanes, uint32_t slot_config,
                     uint32_t checkpoint_every) {
  return guarded(d, [&] {
    PASIF_REQUIRE(tile_lanes == 0 || (tile_lanes % 64 == 0 && tile_lanes <= kMaxLanes),
                  PASIF_ERR_ARGUMENT, "tile lanes must be a multiple of 64 up to %u", kMaxLanes);
    PASIF_REQUIRE(slot_config <= 3, PASIF_ERR_ARGUMENT, "slot configuration 0 .. 3");
    d->tile_lanes = tile_lanes;
    d->slot_cfg = slot_config;
    d->ck_every = checkpoint_every;
  });
}

int pasif_set_graph_slicing(pasif_driver* d, uint32_t steps_per_launch,
                            uint32_t launches_per_graph) {
  return guarded(d, [&] {
    PASIF_REQUIRE(steps_per_launch > 0 && launches_per_graph > 0 && launches_per_graph <= 4096,
                  PASIF_ERR_ARGUMENT, "graph slicing: steps per launch and launches per graph "
                                      "must be positive (at most 4096 launches)");
    d->graph_slice = steps_per_launch;
    d->graph_launches = launches_per_graph;
  });
}

int pasif_set_amplitude_window(pasif_driver* d, uint32_t window) {
  return guarded(d, [&] { d->window = window; });
}

int pasif_get_max_amplitudes(pasif_driver* d, double* out, size_t cap, size_t* len) {
  return guarded(d, [&] {
    PASIF_REQUIRE(d->have_max, PASIF_ERR_STATE, "no windowed solve has run");
    copy_out(d, d->res_max, out, cap, len);
  });
}

int pasif_set_kernel(pasif_driver* d, int kernel) {
  return guarded(d, [&] {
    PASIF_REQUIRE(kernel == PASIF_KERNEL_AUTO || kernel == PASIF_KERNEL_TILE ||
                      kernel == PASIF_KERNEL_LANE || kernel == PASIF_KERNEL_BIG,
                  PASIF_ERR_ARGUMENT, "unknown kernel %d", kernel);
    d->kernel_path = kernel;
  });
}

int pasif_kernel_info(pasif_driver* d, int adjoint, int* kernel, uint32_t* classes,
                      uint32_t* checkpoint_every) {
  return guarded(d, [&] {
    ensure_packed(d, adjoint != 0);
    const Packed& pk = d->pk;
    if (kernel) {
      const int parts = int(pk.lane) + int(pk.tile) + int(pk.big);
      *kernel = parts > 1  ? PASIF_KERNEL_MIXED
                : pk.lane  ? PASIF_KERNEL_LANE
                : pk.big   ? PASIF_KERNEL_BIG
                           : PASIF_KERNEL_TILE;
    }
    if (classes) *classes = pk.lane ? uint32_t(pk.lshapes.size()) : 0u;
    if (checkpoint_every)
      *checkpoint_every = !adjoint ? 0u
                          : pk.lane ? pk.lseg
                          : (pk.big && !pk.tile) ? pk.b_every
                                                 : (d->ck_every ? d->ck_every : 256u);
  });
}

int pasif_set_lane_group(pasif_driver* d, uint32_t lanes) {
  return guarded(d, [&] {
    PASIF_REQUIRE(lanes <= kLaneWave && (lanes & (lanes - 1u)) == 0u, PASIF_ERR_ARGUMENT,
                  "lanes per instance must be 0 (automatic) or a power of two up to %u, got %u",
                  kLaneWave, lanes);
    d->lane_group = lanes;
  });
}

int pasif_set_lane_bias(pasif_driver* d, int bias) {
  return guarded(d, [&] {
    PASIF_REQUIRE(bias >= -6 && bias <= 6, PASIF_ERR_ARGUMENT, "lane bias %d outside [-6, 6]", bias);
    d->lane_bias = bias;
  });
}

int pasif_lane_groups(pasif_driver* d, int adjoint, uint32_t* out, size_t cap, size_t* len) {
  return guarded(d, [&] {
    ensure_packed(d, adjoint != 0);
    std::vector<uint32_t> g;
    if (d->pk.lane)
      g = d->pk.lgroups;
    if (len) *len = g.size();
    if (out) {
      PASIF_REQUIRE(cap >= g.size(), PASIF_ERR_ARGUMENT, "buffer holds %zu, need %zu", cap, g.size());
      std::copy(g.begin(), g.end(), out);
    }
  });
}

int pasif_lane_source(pasif_driver* d, int adjoint, char* out, size_t cap, size_t* len,
                      int compile) {
  return guarded(d, [&] {
    // the lane modules of the automatic plan (every class on lanes when the
    // lane kernel is forced; with tiles forced, the plan the lane kernel
    // would take)
    const int saved = d->kernel_path;
    if (saved != PASIF_KERNEL_AUTO) d->kernel_path = PASIF_KERNEL_LANE;
    const int dev = d->dev;
    d->dev = -1;  // plan on the host only; the compilation below needs no GPU
    d->pk.release();
    const int rc = guarded(d, [&] {
      if (d->precision == 64)
        (void)lane_plan<double>(d, adjoint != 0);
      else
        (void)lane_plan<float>(d, adjoint != 0);
    });
    d->dev = dev;
    d->kernel_path = saved;
    const auto mods = d->pk.lmods;
    d->pk.release();
    if (rc != PASIF_OK) throw Fail{rc};
    // one kernel per module, in module order
    std::string src;
    std::vector<std::string> srcs;
    std::vector<bool> ilp;
    for (size_t m = 0; m < mods.size(); ++m) {
      srcs.push_back(lane_source(mods[m].first, d->precision, mods[m].second));
      ilp.push_back(mods[m].second.ilp);
      if (mods.size() > 1) src += "// ---- lane module " + std::to_string(m) + "\n";
      src += srcs.back();
    }
    if (len) *len = src.size() + 1;
    if (out) {
      PASIF_REQUIRE(cap > src.size(), PASIF_ERR_ARGUMENT, "source buffer holds %zu, need %zu",
                    cap, src.size() + 1);
      std::memcpy(out, src.c_str(), src.size() + 1);
    }
    if (compile) {
      std::string err;
      if (!lane_jit_flags_ok(&err)) {
        set_err(d, "%s", err.c_str());
        throw Fail{PASIF_ERR_ARGUMENT};
      }
      if (!lane_compile_many(srcs, ilp, -1, std::vector<LaneKernel*>(srcs.size(), nullptr),
                             &err)) {
        set_err(d, "%s", err.c_str());
        throw Fail{PASIF_ERR_DEVICE};
      }
    }
  });
}

int pasif_last_timing(const pasif_driver* d, double* kernel_ms, uint32_t* launches,
                      uint32_t* tiles, uint32_t* lanes) {
  if (!d) return PASIF_ERR_ARGUMENT;
  if (kernel_ms) *kernel_ms = d->last_ms;
  if (launches) *launches = d->last_launches;
  if (tiles) *tiles = d->pk.ntiles;
  if (lanes) *lanes = d->pk.lanes;
  return PASIF_OK;
}

}  // extern "C"

// ================================================================= pool
namespace {

template <class F>
int pool_guarded(pasif_pool* P, F&& f) {
  if (!P) return PASIF_ERR_ARGUMENT;
  try {
    P->err.clear();
    f();
    return PASIF_OK;
  } catch (const Fail& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    P->err = "host memory exhausted";
    return PASIF_ERR_MEMORY;
  } catch (const std::exception& e) {
    P->err = e.what();
    return PASIF_ERR_ARGUMENT;
  }
}

#define POOL_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      P->err = std::string("HIP error ") + hipGetErrorName(e_) + " in " + #call;         \
      throw Fail{PASIF_ERR_DEVICE};                                                      \
    }                                                                                    \
  } while (0)

#define POOL_REQUIRE(cond, code, msg) \
  do {                                \
    if (!(cond)) {                    \
      P->err = (msg);                 \
      throw Fail{code};               \
    }                                 \
  } while (0)

double job_bytes(const pasif_driver* d) {
  const DevBuf* run[] = {&d->qstate, &d->lstate, &d->mxstate, &d->progress, &d->out_q,
                         &d->out_max, &d->out_grad, &d->ckpt, &d->segbuf, &d->d_exc,
                         &d->d_interp};
  double b = double(d->pk.program_bytes) + double(d->pk.hdr.bytes);
  for (const DevBuf* x : run) b += double(x->bytes);
  for (const auto& lp : d->pk.lparts) {
    b += double(lp->prog.bytes);
    for (const auto& L : lp->cls) {
      const DevBuf* lb[] = {&L->qstate, &L->lstate, &L->mxstate, &L->progress, &L->ckpt,
                            &L->segbuf};
      for (const DevBuf* x : lb) b += double(x->bytes);
    }
  }
  return b;
}

void release_run(pasif_driver* d) {
  DevBuf* run[] = {&d->qstate, &d->lstate, &d->mxstate, &d->progress, &d->out_q,
                   &d->out_max, &d->out_grad, &d->ckpt, &d->segbuf};
  for (DevBuf* x : run) x->release();
  d->pk.release();
  d->pool_largs.clear();
}

// One system's block of a side, re-indexed from 0: factor 0 relative to the
// block's own rows, factors 1..3 relative to its set-point block (the forward
// rows of the same system for the adjoint, its own rows for the forward).
RowTerms slice_block(const RowTerms& rt, size_t s, bool adjoint) {
  RowTerms o;
  const uint32_t off = rt.off[s], n = rt.size[s];
  const uint32_t xo = adjoint ? rt.xoff[s] : off;
  o.n = n;
  o.nset = n;
  for (int k = 0; k < 4; ++k) {
    const int nf = k + 1;
    const uint32_t b = rt.ptr[k][off], e = rt.ptr[k][off + n];
    o.ptr[k].resize(n + 1);
    for (uint32_t r = 0; r <= n; ++r) o.ptr[k][r] = rt.ptr[k][off + r] - b;
    o.val[k].assign(rt.val[k].begin() + b, rt.val[k].begin() + e);
    o.idx[k].assign(rt.idx[k].begin() + size_t(b) * nf, rt.idx[k].begin() + size_t(e) * nf);
    for (size_t t = 0; t < size_t(e - b); ++t) {
      o.idx[k][t * nf] -= off;
      for (int a = 1; a < nf; ++a) o.idx[k][t * nf + a] -= xo;
    }
  }
  o.force.assign(rt.force.begin() + off, rt.force.begin() + off + n);
  o.init.assign(rt.init.begin() + off, rt.init.begin() + off + n);
  o.size = {n};
  o.off = {0};
  if (adjoint) o.xoff = {0};
  return o;
}

// Concatenate re-indexed blocks into one side with the given partition; for
// the adjoint, factors 1..3 of block i are shifted by xoff[i] (the forward
// offset of system i).
void concat_blocks(const std::vector<const RowTerms*>& parts, bool adjoint,
                   const std::vector<uint32_t>& xoff, RowTerms& o) {
  o = RowTerms{};
  size_t n = 0, terms[4] = {0, 0, 0, 0};
  for (const RowTerms* p : parts) {
    n += p->n;
    for (int k = 0; k < 4; ++k) terms[k] += p->val[k].size();
  }
  o.n = uint32_t(n);
  o.size.reserve(parts.size());
  o.off.reserve(parts.size());
  o.force.reserve(n);
  o.init.reserve(n);
  for (int k = 0; k < 4; ++k) {
    o.ptr[k].reserve(n + 1);
    o.ptr[k].push_back(0);
    o.val[k].reserve(terms[k]);
    o.idx[k].reserve(terms[k] * (k + 1));
  }
  uint32_t off = 0;
  for (size_t i = 0; i < parts.size(); ++i) {
    const RowTerms& p = *parts[i];
    const uint32_t xo = adjoint ? xoff[i] : off;
    for (int k = 0; k < 4; ++k) {
      const int nf = k + 1;
      const uint32_t base = o.ptr[k].back();
      for (uint32_t r = 1; r <= p.n; ++r) o.ptr[k].push_back(base + p.ptr[k][r]);
      o.val[k].insert(o.val[k].end(), p.val[k].begin(), p.val[k].end());
      const size_t first = o.idx[k].size();
      o.idx[k].insert(o.idx[k].end(), p.idx[k].begin(), p.idx[k].end());
      for (size_t t = first; t < o.idx[k].size(); t += nf) {
        o.idx[k][t] += off;
        for (int a = 1; a < nf; ++a) o.idx[k][t + a] += xo;
      }
    }
    o.force.insert(o.force.end(), p.force.begin(), p.force.end());
    o.init.insert(o.init.end(), p.init.begin(), p.init.end());
    o.size.push_back(p.n);
    o.off.push_back(off);
    off += p.n;
  }
  if (adjoint) o.xoff = xoff;
}

// a job handle reading the pool's signal bank
pasif_driver* new_bank_job(pasif_pool* P, uint32_t numsteps) {
  POOL_REQUIRE(P->bank_n > 0, PASIF_ERR_STATE, "no signal bank: pasif_pool_set_bank first");
  pasif_driver* d = new pasif_driver();
  d->dev = P->dev;
  d->stream = P->upload;
  d->pool = P;
  d->precision = P->precision;
  d->tile_lanes = P->lanes;
  d->slot_cfg = P->slot_cfg;
  d->ck_every = P->ck_every;
  d->ext_exc = P->bank.p;
  d->ext_interp = P->bank_w ? P->bank_interp.p : nullptr;
  d->n_exc = P->bank_n;
  d->exc_len = P->bank_len;
  d->sample_rate = P->bank_rate;
  d->interp_w = P->bank_w;
  d->interp.assign(size_t(2) * P->bank_w, 0.0);
  d->numsteps = numsteps ? numsteps : P->bank_len;
  d->d_exc_dirty = false;
  return d;
}

// Remove job k from the pool and free its device memory (its handle stays
// with the caller).
void pool_remove_job(pasif_pool* P, size_t k) {
  auto& j = P->jobs[k];
  pasif_driver* d = j.d;
  release_run(d);
  d->d_exc.release();
  d->d_interp.release();
  d->d_exc_dirty = true;
  P->used -= j.bytes;
  P->tiles_used -= j.tiles;
  P->jobs.erase(P->jobs.begin() + long(k));
}

// A pool handle's device memory (all of it in the pool's arena) is freed and
// the handle no longer refers to the pool: pasif_destroy deletes it as a
// plain driver.
void pool_detach(pasif_driver* d) {
  release_run(d);
  d->d_exc.release();
  d->d_interp.release();
  d->traj.release();
  d->pool = nullptr;
  d->stream = nullptr;
  d->ext_exc = nullptr;
  d->ext_interp = nullptr;
}

// Delete the jobs their owners destroyed while a tick ran them (those no
// launch in flight uses).
void pool_reap(pasif_pool* P) {
  for (size_t k = P->jobs.size(); k-- > 0;) {
    pasif_driver* d = P->jobs[k].d;
    if (!d->orphan || (P->in_flight && !P->jobs[k].finished)) continue;
    pool_remove_job(P, k);
    P->handles.erase(d);
    pool_detach(d);
    delete d;
  }
}

// Pack a loaded job, allocate its run buffers from the arena and admit it
// if the slot cap and the HBM budget allow (PASIF_ERR_MEMORY otherwise,
// leaving the job unadmitted with no device memory).
void admit_job(pasif_pool* P, pasif_driver* d, int adjoint) {
  for (const auto& j : P->jobs)
    POOL_REQUIRE(j.d != d, PASIF_ERR_STATE, "job already submitted");
  // full already: refuse before packing
  POOL_REQUIRE(P->jobs.empty() || P->max_tiles == 0 || P->tiles_used < P->max_tiles,
               PASIF_ERR_MEMORY, "no free slots: every tile slot of the pool is in use");
  POOL_REQUIRE(P->jobs.empty() || P->used < P->hbm_budget, PASIF_ERR_MEMORY,
               "HBM budget exhausted");
  POOL_HIP(hipSetDevice(P->dev));
  d->precision = P->precision;
  d->tile_lanes = P->lanes;
  d->slot_cfg = P->slot_cfg;
  d->ck_every = P->ck_every;
  d->window = 0;
  uint64_t work = 0;
  uint32_t nslots = 0;
  using clk = std::chrono::steady_clock;
  const auto c0 = clk::now();
  auto c1 = c0;
  int rc = guarded(d, [&] {
    ensure_packed(d, adjoint != 0);
    c1 = clk::now();
    if (d->pk.tile) {
      if (P->precision == 64) {
        TileArgs<double> a = prepare_run<double>(d, adjoint != 0, 0, &work, &nslots);
        d->pool_ck = a.ck_every;
      } else {
        TileArgs<float> a = prepare_run<float>(d, adjoint != 0, 0, &work, &nslots);
        d->pool_ck = a.ck_every;
      }
      PASIF_REQUIRE(d->pk.lanes == P->lanes, PASIF_ERR_UNSUPPORTED,
                    "a system of this request needs %u lanes, the pool runs %u", d->pk.lanes,
                    P->lanes);
      PASIF_REQUIRE(d->pk.slot_cfg == P->slot_cfg, PASIF_ERR_UNSUPPORTED,
                    "slot configuration mismatch");
    }
    d->pool_largs.clear();
    if (d->pk.lane) {
      if (P->precision == 64) {
        (void)run_outputs<double>(d, adjoint != 0, 0);
        for (auto& lp : d->pk.lparts)
          d->pool_largs.push_back(lane_part_args<double>(d, *lp, adjoint != 0, 0, 0,
                                                         P->bank.p, P->bank_interp.p));
      } else {
        (void)run_outputs<float>(d, adjoint != 0, 0);
        for (auto& lp : d->pk.lparts)
          d->pool_largs.push_back(lane_part_args<float>(d, *lp, adjoint != 0, 0, 0,
                                                        P->bank.p, P->bank_interp.p));
      }
      work = std::max<uint64_t>(work, adjoint ? 3ull * d->pk.ltmax : d->pk.ltmax);
    }
  });
  if (rc != PASIF_OK) {
    release_run(d);
    P->err = d->err;
    throw Fail{rc};
  }
  const double bytes = job_bytes(d);
  uint32_t tiles = d->pk.tile ? d->pk.ntiles : 0;
  for (const auto& lp : d->pk.lparts) tiles += lp->blocks;
  if (!P->jobs.empty() && P->max_tiles && P->tiles_used + tiles > P->max_tiles) {
    release_run(d);
    P->err = "no free slots: the request needs " + std::to_string(tiles) + " tiles, " +
             std::to_string(P->max_tiles - P->tiles_used) + " are free";
    throw Fail{PASIF_ERR_MEMORY};
  }
  if (P->used + bytes > P->hbm_budget) {
    release_run(d);
    P->err = "HBM budget exhausted: the request needs " + std::to_string(bytes) +
             " bytes, " + std::to_string(P->hbm_budget - P->used) + " are free";
    throw Fail{PASIF_ERR_MEMORY};
  }
  const auto c2 = clk::now();
  POOL_HIP(hipEventRecord(P->uploaded, P->upload));
  P->used += bytes;
  P->tiles_used += tiles;
  if (P->timing) {
    P->tm[1] += std::chrono::duration<double, std::milli>(c1 - c0).count();
    P->tm[2] += std::chrono::duration<double, std::milli>(c2 - c1).count();
  }
  P->jobs.push_back({d, adjoint != 0, work, 0, bytes, false, tiles});
}

// The forcing rule of the pool's signal bank as the lane kernels compile it
// (1 interpolation, 2 zero-order hold; 0 no bank yet: both).
uint32_t pool_forcing(const pasif_pool* P) {
  return P->bank_n ? (P->bank_w ? 1u : 2u) : 0u;
}

// Pool modules: only the bank's forcing rule, no per-step outputs (pool jobs
// ask for final states and gradients only).
void pool_specialise(const pasif_pool* P, LaneLayout* lay) {
  lay->forcing = pool_forcing(P);
  lay->full = false;
}

// How a class runs on the pool's lane modules: 1 = in one wave (its group
// size chosen here: the shortest step among the sizes within `slack` x the
// least work, default 2, since a tick lasts `budget` steps of the slowest
// class and a request waits for its slowest system; PASIF_POOL_WORK_SLACK=1
// takes the least work, throughput over latency), w >= 2 = over the w waves
// of a workgroup, 0 = on tiles.
uint32_t pool_class_fit(LaneShape& sh, int bits) {
  const auto opt = group_options(sh, bits);
  if (!opt.empty()) {
    double slack = 2.0;
    if (const char* e = std::getenv("PASIF_POOL_WORK_SLACK")) {
      const double v = std::strtod(e, nullptr);
      if (v >= 1.0 && v <= 16.0) slack = v;
    }
    double wmin = 0;
    for (size_t k = 0; k < opt.size(); ++k) {
      const double wk = opt[k].second * opt[k].first;
      if (k == 0 || wk < wmin) wmin = wk;
    }
    size_t b = opt.size();
    for (size_t k = 0; k < opt.size(); ++k)
      if (opt[k].second * opt[k].first <= slack * wmin &&
          (b == opt.size() || opt[k].second < opt[b].second))
        b = k;
    (void)lane_group(sh, opt[b].first);
    return 1;
  }
  for (uint32_t w = 2; w <= kLaneMaxWaves; ++w)
    if (lane_partition(sh, w) <= kLaneMaxWaveLive) return w;
  return 0;
}

// Layout of a single-wave pool module: a fixed LDS share per wave (eight
// waves per CU resident) sets each class's checkpoint interval; registers
// for at most four waves per SIMD.
LaneLayout pool_single_layout(std::vector<LaneShape>& shapes, uint32_t Kreq, size_t el) {
  LaneLayout lay;
  uint32_t live = 0;
  bool adj = false;
  lay.seg = Kreq;
  for (LaneShape& sh : shapes) {
    live = std::max(live, lane_live(sh));
    if (!sh.adjoint) continue;
    adj = true;
    const uint32_t G = sh.group > 1 ? sh.group : 1u;
    const size_t stage = size_t(kLaneStage + (G > 1 ? 3 * kLaneChunk : 0)) * (kLaneWave / G) * el;
    const size_t step = size_t(lane_seg_rows(sh)) * kLaneWave * el;
    const size_t b = std::min<size_t>(kLaneMaxLds, std::max<size_t>(kMaxLdsBytes / 8, stage + 6 * step));
    const size_t fit = b > stage ? (b - stage) / step : 0;
    sh.seg = uint32_t(std::max<size_t>(1, std::min<size_t>(Kreq, fit > 0 ? fit - 1 : 0)));
    lay.seg = std::min(lay.seg, sh.seg);
  }
  if (!adj) lay.seg = 1;
  lay.occ = std::min<uint32_t>(4, std::max<uint32_t>(1, 512 / (live * uint32_t(el * 8 / 32) + 48)));
  return lay;
}

// Layout of a multi-wave pool module (16 instances per workgroup over
// `waves` waves): the longest checkpoint interval up to Kreq whose segment
// fits the workgroup's LDS; false when none does.
bool pool_multi_layout(std::vector<LaneShape>& shapes, uint32_t waves, uint32_t Kreq, size_t el,
                       LaneLayout* out) {
  LaneLayout lay;
  bool adj = false;
  for (LaneShape& sh : shapes) {
    (void)lane_partition(sh, waves);
    adj = adj || sh.adjoint;
  }
  lay.waves = waves;
  lay.lw = 16;
  lay.seg = 1;
  if (adj) {
    uint32_t lo = 1, hi = Kreq;
    lay.seg = 0;
    while (lo <= hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      LaneLayout t = lay;
      t.seg = mid;
      if (lane_lds_bytes(shapes, int(el * 8), t) <= kMaxLdsBytes) {
        lay.seg = mid;
        lo = mid + 1;
      } else {
        hi = mid - 1;
      }
    }
  }
  if (lay.seg == 0 || lane_lds_bytes(shapes, int(el * 8), lay) > kMaxLdsBytes) return false;
  *out = lay;
  return true;
}

// Lane modules over the registered structures: module 0 runs the classes
// that fit one lane or a group of lanes (each at its cheapest group size:
// requests arrive continuously, so throughput, not one request's chain of
// steps, sets the layout), module 1 (multi-wave, 16 instances per
// workgroup) the classes that need a workgroup; classes fitting neither
// stay on tiles.
template <typename R>
void pool_build_modules(pasif_pool* P) {
  const double eps = std::numeric_limits<R>::epsilon();
  std::map<std::vector<uint32_t>, uint32_t> cls_of_key;
  std::vector<LaneShape> all;
  std::vector<std::vector<uint32_t>> keys, fkeys;
  std::vector<uint32_t> count;
  std::vector<uint32_t> key, fkey;
  for (const auto& e : P->reg) {
    LaneShape sh;
    sh.f = lane_side(e.f, 0, false, eps);
    sh.adjoint = e.has_adj;
    if (e.has_adj) sh.a = lane_side(e.a, 0, true, eps);
    key.clear();
    shape_key(sh.f, key);
    const size_t nf = key.size();
    key.push_back(e.has_adj ? 1u : 0u);
    if (e.has_adj) shape_key(sh.a, key);
    auto it = cls_of_key.find(key);
    if (it == cls_of_key.end()) {
      it = cls_of_key.emplace(key, uint32_t(all.size())).first;
      all.push_back(std::move(sh));
      keys.push_back(key);
      fkeys.emplace_back(key.begin(), key.begin() + long(nf));
      count.push_back(0);
    }
    count[it->second] += 1;
  }
  std::vector<uint32_t> need(all.size(), 0);
  const int bits = int(sizeof(R) * 8);
  for (size_t c = 0; c < all.size(); ++c) need[c] = pool_class_fit(all[c], bits);
  P->mods.clear();
  P->mods.resize(2);
  P->mods_forcing = pool_forcing(P);
  const uint32_t Kreq = P->ck_every ? P->ck_every : 256;
  const size_t el = sizeof(R);
  for (int m = 0; m < 2; ++m) {
    std::vector<uint32_t> order;
    uint32_t waves = 1;
    for (uint32_t c = 0; c < all.size(); ++c)
      if (m == 0 ? need[c] == 1 : need[c] >= 2) order.push_back(c);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return count[x] > count[y]; });
    if (order.size() > kLaneMaxClasses) order.resize(kLaneMaxClasses);
    for (uint32_t c : order) waves = std::max(waves, need[c]);
    if (order.empty()) continue;
    auto& M = P->mods[size_t(m)];
    for (uint32_t c : order) M.shapes.push_back(all[c]);
    LaneLayout lay;
    if (m == 0) {
      lay = pool_single_layout(M.shapes, Kreq, el);
    } else if (!pool_multi_layout(M.shapes, waves, Kreq, el, &lay)) {
      M.shapes.clear();  // does not fit: these classes stay on tiles
      continue;
    }
    pool_specialise(P, &lay);
    M.lay = lay;
    M.lds = lane_lds_bytes(M.shapes, int(el * 8), lay);
    POOL_REQUIRE(M.lds <= (lay.waves > 1 ? kMaxLdsBytes : kLaneMaxLds), PASIF_ERR_UNSUPPORTED,
                 "pool lane module LDS");
    const std::string src = lane_source(M.shapes, int(el * 8), lay);
    std::string err;
    POOL_HIP(hipSetDevice(P->dev));
    if (!lane_jit_flags_ok(&err)) {
      P->err = err;
      throw Fail{PASIF_ERR_ARGUMENT};
    }
    if (!lane_compile(src, P->dev, &M.kern, &err)) {
      P->err = err;
      throw Fail{PASIF_ERR_DEVICE};
    }
    for (size_t k = 0; k < order.size(); ++k) {
      P->key_lane[keys[order[k]]] = lane_code(m, uint32_t(k));
      P->fkey_lane.emplace(fkeys[order[k]], lane_code(m, uint32_t(k)));
    }
  }
  P->reg_lane.assign(P->reg.size(), -1);
  size_t id = 0;
  for (const auto& e : P->reg) {
    LaneShape sh;
    sh.f = lane_side(e.f, 0, false, eps);
    if (e.has_adj) sh.a = lane_side(e.a, 0, true, eps);
    key.clear();
    shape_key(sh.f, key);
    key.push_back(e.has_adj ? 1u : 0u);
    if (e.has_adj) shape_key(sh.a, key);
    auto it = P->key_lane.find(key);
    P->reg_lane[id++] = it == P->key_lane.end() ? -1 : it->second;
  }
  P->mods_built = true;
}

// A pool job's lane parts: its systems of each module, laid out as the module.
template <typename R>
void pool_lane_parts(pasif_driver* d, bool adjoint) {
  pasif_pool* P = d->pool;
  for (size_t m = 0; m < P->mods.size(); ++m) {
    bool any = false;
    for (int32_t c : d->lane_cls) any = any || (c >= 0 && lane_part_of(c) == int(m));
    if (!any) continue;
    const auto& M = P->mods[m];
    std::vector<uint32_t> klass(M.shapes.size());
    for (size_t c = 0; c < klass.size(); ++c) klass[c] = uint32_t(c);
    LanePart* lp = lane_build<R>(d, adjoint, int(m), M.shapes, klass, M.lay);
    lp->kern = M.kern;
    lp->lds = M.lds;
  }
}

template <typename R>
void pool_tick_impl(pasif_pool* P) {
  std::vector<TileArgs<R>> table;
  P->map_host.clear();
  size_t lds = 0;
  for (auto& M : P->mods) {
    M.table.clear();
    M.blocks = 0;
  }
  // jobs whose work the completed ticks covered leave the launches: their
  // buffers are then untouched by the GPU and may be collected while a
  // later tick runs
  // (finished jobs hold no slots: the slot cap counts the workgroups ticks
  // still launch)
  for (auto& j : P->jobs)
    if (j.done >= j.work && !j.finished) {
      j.finished = true;
      P->tiles_used -= j.tiles;
      j.tiles = 0;
    }
  for (auto& j : P->jobs) {
    if (j.finished) continue;
    pasif_driver* d = j.d;
    // a lane workgroup of longest instance T has finished once the completed
    // ticks cover its T (3 T with the adjoint: forward, recompute, reverse)
    // steps of work; only the unfinished prefix of each class is launched
    const uint64_t mult = j.adjoint ? 3u : 1u;
    uint32_t held = d->pk.tile ? d->pk.ntiles : 0u;
    for (size_t i = 0; i < d->pk.lparts.size(); ++i) {
      LanePart& lp = *d->pk.lparts[i];
      auto& M = P->mods[size_t(lp.module)];
      for (size_t c = 0; c < lp.cls.size(); ++c) {
        LaneCls& L = *lp.cls[c];
        while (L.live > 0 && mult * L.wave_t[L.live - 1] <= j.done) --L.live;
        held += L.live;
        if (!L.live) continue;
        LaneArgs a = d->pool_largs[i][c];
        a.block_begin = M.blocks;
        a.budget = P->budget;
        M.table.push_back(a);
        M.blocks += L.live;
      }
    }
    P->tiles_used -= j.tiles - held;
    j.tiles = held;
    if (!d->pk.tile) continue;
    // the job's buffers already exist: re-derive its arguments without
    // clearing the progress (prepare_run only clears at admission)
    TileArgs<R> a{};
    {
      Packed& pk = d->pk;
      a.hdr = pk.hdr.as<TileHeader>();
      a.lanes = pk.lanes;
      fill_program(a.fwd, pk.f_info, pk.f_out, pk.f_src, pk.f_steps, pk.f_exc, pk.f_fslot,
                   pk.f_force, pk.f_val, pk.f_idx, pk.f_init);
      if (j.adjoint)
        fill_program(a.adj, pk.a_info, pk.a_out, pk.a_src, pk.a_steps, pk.a_exc, pk.a_fslot,
                     pk.a_force, pk.a_val, pk.a_idx, pk.a_init);
      a.max_forced = pk.max_forced;
      a.excit = d->ext_exc ? static_cast<const R*>(d->ext_exc) : d->d_exc.as<R>();
      a.exc_len = d->exc_len;
      a.interp = d->interp_w ? (d->ext_interp ? static_cast<const R*>(d->ext_interp)
                                              : d->d_interp.as<R>())
                             : nullptr;
      a.interp_w = d->interp_w;
      const R h = R(1.0 / double(d->sample_rate));
      a.h = h;
      a.h2 = R(double(h) / 2.0);
      a.h6 = R(double(h) / 6.0);
      a.qstate = d->qstate.as<R>();
      a.progress = d->progress.as<uint32_t>();
      a.out_q = d->out_q.as<R>();
      a.out_len = pk.out_len_f;
      if (j.adjoint) {
        a.lstate = d->lstate.as<R>();
        a.ckpt = d->ckpt.as<R>();
        a.segbuf = d->segbuf.as<R>();
        a.out_grad = d->out_grad.as<R>();
        a.ck_every = d->pool_ck;
        a.ck_tile = d->pk.ck_tile.as<uint32_t>();
        a.adjoint = 1;
      }
      a.budget = P->budget;
    }
    const uint32_t id = uint32_t(table.size());
    table.push_back(a);
    for (uint32_t t = 0; t < d->pk.ntiles; ++t) {
      P->map_host.push_back(id);
      P->map_host.push_back(t);
    }
    lds = std::max(lds, tile_lds_bytes(sizeof(R), P->lanes, true, a.interp_w, a.max_forced));
  }
  const uint32_t nblocks = uint32_t(P->map_host.size() / 2);
  uint32_t lblocks = 0;
  for (const auto& M : P->mods) lblocks += M.blocks;
  P->last_blocks = nblocks + lblocks;
  if (nblocks == 0 && lblocks == 0) return;
  POOL_HIP(hipStreamWaitEvent(P->stream, P->uploaded, 0));
  if (nblocks) {
    const size_t tb = table.size() * sizeof(TileArgs<R>);
    P->table_host.assign(reinterpret_cast<unsigned char*>(table.data()),
                         reinterpret_cast<unsigned char*>(table.data()) + tb);
    if (P->table_bytes < tb) {
      if (P->table_dev) (void)hipFree(P->table_dev);
      P->table_dev = nullptr;
      POOL_HIP(hipMalloc(&P->table_dev, tb));
      P->table_bytes = tb;
    }
    const size_t mb = P->map_host.size() * sizeof(uint32_t);
    if (P->map_bytes < mb) {
      if (P->map_dev) (void)hipFree(P->map_dev);
      P->map_dev = nullptr;
      POOL_HIP(hipMalloc(&P->map_dev, mb));
      P->map_bytes = mb;
    }
    POOL_HIP(hipMemcpyAsync(P->table_dev, P->table_host.data(), tb, hipMemcpyHostToDevice,
                            P->stream));
    POOL_HIP(hipMemcpyAsync(P->map_dev, P->map_host.data(), mb, hipMemcpyHostToDevice,
                            P->stream));
  }
  for (auto& M : P->mods) {
    if (!M.blocks) continue;
    const size_t b = M.table.size() * sizeof(LaneArgs);
    if (M.table_bytes < b) {
      if (M.table_dev) (void)hipFree(M.table_dev);
      M.table_dev = nullptr;
      POOL_HIP(hipMalloc(&M.table_dev, b));
      M.table_bytes = b;
    }
    POOL_HIP(hipMemcpyAsync(M.table_dev, M.table.data(), b, hipMemcpyHostToDevice, P->stream));
  }
  POOL_HIP(hipEventRecord(P->ev0, P->stream));
  // lane modules on their own streams beside the tile launch
  if (lblocks) POOL_HIP(hipEventRecord(P->fork, P->stream));
  for (size_t m = 0; m < P->mods.size(); ++m) {
    auto& M = P->mods[m];
    if (!M.blocks) continue;
    POOL_HIP(hipStreamWaitEvent(P->aux[m], P->fork, 0));
    const uint32_t lds_m = M.lay.waves > 1 ? 0u : uint32_t(M.lds);
    void* dp = M.table_dev;
    uint32_t n = uint32_t(M.table.size());
    void* params[] = {&dp, &n};
    hipError_t e = hipModuleLaunchKernel(M.kern.fn, M.blocks, 1, 1, kLaneWave * M.lay.waves, 1,
                                         1, lds_m, P->aux[m], params, nullptr);
    if (e != hipSuccess) {
      P->err = std::string("pool lane kernel launch failed: ") + hipGetErrorName(e);
      throw Fail{PASIF_ERR_DEVICE};
    }
    POOL_HIP(hipEventRecord(P->join[m], P->aux[m]));
  }
  if (nblocks) {
    hipError_t e = launch_pool(static_cast<const TileArgs<R>*>(P->table_dev),
                               static_cast<const uint32_t*>(P->map_dev), nblocks, P->lanes,
                               int(P->slot_cfg), lds, P->stream);
    if (e != hipSuccess) {
      P->err = std::string("pool kernel launch failed: ") + hipGetErrorName(e);
      throw Fail{PASIF_ERR_DEVICE};
    }
  }
  for (size_t m = 0; m < P->mods.size(); ++m)
    if (P->mods[m].blocks) POOL_HIP(hipStreamWaitEvent(P->stream, P->join[m], 0));
  POOL_HIP(hipEventRecord(P->ev1, P->stream));
  for (auto& j : P->jobs)
    if (!j.finished) j.done += P->budget;
  P->in_flight = true;
}

}  // namespace

namespace {
// Forget every lane class of the pool (the modules are about to be built
// again): a background build in flight is waited for and dropped, queued
// structures are forgotten (they are queued again when seen).
void pool_forget_modules(pasif_pool* P) {
  if (P->build) {
    if (P->build->th.joinable()) P->build->th.join();
    P->build.reset();
  }
  P->key_lane.clear();
  P->fkey_lane.clear();
  P->queued.clear();
  P->pending.clear();
  P->loose.clear();
  P->loose_mod.clear();
}
void pool_match_lanes(pasif_pool* P, pasif_driver* d, bool adjoint);
void pool_install_build(pasif_pool* P, bool wait);
template <typename R>
void pool_start_build(pasif_pool* P);
void pool_compile_step(pasif_pool* P);
}  // namespace

extern "C" {

int pasif_pool_create(pasif_pool** out, int gpu_id, int precision, uint32_t tile_lanes,
                      uint32_t slot_config, uint32_t checkpoint_every, uint32_t steps_per_tick,
                      double hbm_budget_bytes) {
  if (!out) return PASIF_ERR_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<pasif_pool> holder(new (std::nothrow) pasif_pool());
  if (!holder) return PASIF_ERR_MEMORY;
  pasif_pool* P = holder.get();
  const int rc = pool_guarded(P, [&] {
    POOL_REQUIRE(precision == 64 || precision == 32, PASIF_ERR_ARGUMENT, "precision 32 or 64");
    POOL_REQUIRE(tile_lanes >= 64 && tile_lanes <= kMaxLanes && tile_lanes % 64 == 0,
                 PASIF_ERR_ARGUMENT, "tile lanes must be a multiple of 64 up to 1024");
    POOL_REQUIRE(slot_config >= 1 && slot_config <= 3 && (tile_lanes <= 256 || slot_config == 1),
                 PASIF_ERR_ARGUMENT, "slot configuration 1..3 (1 above 256 lanes)");
    POOL_REQUIRE(steps_per_tick > 0, PASIF_ERR_ARGUMENT, "steps per tick must be positive");
    POOL_REQUIRE(hbm_budget_bytes > 0, PASIF_ERR_ARGUMENT, "HBM budget must be positive");
    int ndev = 0;
    POOL_HIP(hipGetDeviceCount(&ndev));
    POOL_REQUIRE(gpu_id >= 0 && gpu_id < ndev, PASIF_ERR_ARGUMENT, "no such GPU");
    P->dev = gpu_id;
    POOL_HIP(hipSetDevice(gpu_id));
    POOL_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    POOL_HIP(hipStreamCreateWithFlags(&P->upload, hipStreamNonBlocking));
    POOL_HIP(hipEventCreate(&P->ev0));
    POOL_HIP(hipEventCreate(&P->ev1));
    POOL_HIP(hipEventCreateWithFlags(&P->uploaded, hipEventDisableTiming));
    for (int m = 0; m < 2; ++m) {
      P->aux.push_back(nullptr);
      P->join.push_back(nullptr);
      POOL_HIP(hipStreamCreateWithFlags(&P->aux.back(), hipStreamNonBlocking));
      POOL_HIP(hipEventCreateWithFlags(&P->join.back(), hipEventDisableTiming));
    }
    POOL_HIP(hipEventCreateWithFlags(&P->fork, hipEventDisableTiming));
    P->precision = precision;
    P->lanes = tile_lanes;
    P->slot_cfg = slot_config;
    P->ck_every = checkpoint_every ? checkpoint_every : 256;
    P->budget = steps_per_tick;
    P->hbm_budget = hbm_budget_bytes;
    P->arena.reset(new Arena());
    P->arena->chunk_bytes = size_t(std::min(hbm_budget_bytes, double(size_t(1) << 30)));
  });
  *out = holder.release();
  return rc;
}

const char* pasif_pool_last_error(const pasif_pool* P) { return P ? P->err.c_str() : "null pool"; }

int pasif_pool_job_create(pasif_pool* P, pasif_driver** job, const double* excitations,
                          uint32_t n_exc, uint32_t exc_len, uint32_t sample_rate,
                          uint32_t numsteps) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(job != nullptr, PASIF_ERR_ARGUMENT, "null job pointer");
    if (excitations == nullptr) {  // the pool's signal bank: excitation ids index it
      *job = new_bank_job(P, numsteps);
      P->handles.insert(*job);
      return;
    }
    std::unique_ptr<pasif_driver> d(new pasif_driver());
    d->dev = P->dev;
    d->stream = P->upload;
    d->pool = P;
    d->precision = P->precision;
    d->tile_lanes = P->lanes;
    d->slot_cfg = P->slot_cfg;
    d->ck_every = P->ck_every;
    {
      const int rc = guarded(d.get(), [&] {
        pasif_driver* dd = d.get();
        load_exc(dd, excitations, n_exc, exc_len, sample_rate);
        dd->numsteps = numsteps ? numsteps : exc_len;
      });
      POOL_REQUIRE(rc == PASIF_OK, rc, d->err);
    }
    *job = d.release();
    P->handles.insert(*job);
  });
}

int pasif_pool_set_bank(pasif_pool* P, const double* excitations, uint32_t n_exc,
                        uint32_t exc_len, uint32_t sample_rate, const double* interp,
                        uint32_t window) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(P->jobs.empty(), PASIF_ERR_STATE, "the bank can only change on an empty pool");
    POOL_REQUIRE(excitations && n_exc > 0 && exc_len > 0 && sample_rate > 0, PASIF_ERR_ARGUMENT,
                 "empty signal bank");
    POOL_REQUIRE(!interp || (window >= 2 && window <= kMaxInterpWindow), PASIF_ERR_ARGUMENT,
                 "interpolation window must be in [2, 2048]");
    POOL_HIP(hipSetDevice(P->dev));
    const size_t n = size_t(n_exc) * exc_len;
    const size_t el = P->precision == 64 ? 8 : 4;
    P->bank.release();
    P->bank_interp.release();
    POOL_HIP(hipMalloc(&P->bank.p, n * el));
    P->bank.bytes = n * el;
    if (P->precision == 64) {
      POOL_HIP(hipMemcpy(P->bank.p, excitations, n * el, hipMemcpyHostToDevice));
    } else {
      std::vector<float> f(excitations, excitations + n);
      POOL_HIP(hipMemcpy(P->bank.p, f.data(), n * el, hipMemcpyHostToDevice));
    }
    if (interp) {
      POOL_HIP(hipMalloc(&P->bank_interp.p, 2 * size_t(window) * el));
      P->bank_interp.bytes = 2 * size_t(window) * el;
      if (P->precision == 64) {
        POOL_HIP(hipMemcpy(P->bank_interp.p, interp, 2 * size_t(window) * el,
                           hipMemcpyHostToDevice));
      } else {
        std::vector<float> f(interp, interp + 2 * size_t(window));
        POOL_HIP(hipMemcpy(P->bank_interp.p, f.data(), 2 * size_t(window) * el,
                           hipMemcpyHostToDevice));
      }
    }
    P->bank_n = n_exc;
    P->bank_len = exc_len;
    P->bank_rate = sample_rate;
    P->bank_w = interp ? window : 0;
    if (P->mods_built && P->mods_forcing != 0 && P->mods_forcing != pool_forcing(P)) {
      // modules compiled for the other forcing rule: built again at the next
      // submit
      pool_forget_modules(P);
      P->mods_built = false;
    }
  });
}

int pasif_pool_submit(pasif_pool* P, pasif_driver* d, int adjoint) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(d && d->pool == P, PASIF_ERR_ARGUMENT, "job does not belong to this pool");
    pool_install_build(P, false);
    pool_match_lanes(P, d, adjoint != 0);
    admit_job(P, d, adjoint);
    if (P->precision == 64)
      pool_start_build<double>(P);
    else
      pool_start_build<float>(P);
  });
}

int pasif_pool_set_slots(pasif_pool* P, uint32_t max_tiles) {
  return pool_guarded(P, [&] { P->max_tiles = max_tiles; });
}

int pasif_pool_set_compile(pasif_pool* P, int on) {
  return pool_guarded(P, [&] { P->compile_new = on != 0; });
}

int pasif_pool_compile_wait(pasif_pool* P, uint32_t max_builds, uint32_t* modules,
                            uint32_t* failed, uint32_t* queued) {
  return pool_guarded(P, [&] {
    pool_compile_step(P);  // starts a build if structures wait and none runs
    for (uint32_t k = 0; P->build && (max_builds == 0 || k < max_builds); ++k)
      pool_install_build(P, true);
    if (modules) *modules = P->built_modules;
    if (failed) *failed = P->failed_builds;
    if (queued) {
      size_t n = P->pending.size();
      if (P->build) n += P->build->cls.size();
      *queued = uint32_t(n);
    }
  });
}

int pasif_pool_register(pasif_pool* P, const pasif_driver* src, uint32_t* first_id,
                        uint32_t* count) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(src && first_id && count, PASIF_ERR_ARGUMENT, "null argument");
    POOL_REQUIRE(src->sys_ready, PASIF_ERR_STATE,
                 "the source handle's system must be allocated (pasif_allocate_system)");
    const size_t S = src->fwd.size.size();
    POOL_REQUIRE(P->reg.size() + S < (size_t(1) << 31), PASIF_ERR_ARGUMENT, "registry full");
    const size_t first = P->reg.size();
    P->reg.reserve(first + S);
    for (size_t s = 0; s < S; ++s) {
      pasif_pool::Entry e;
      e.f = slice_block(src->fwd, s, false);
      if (src->adj_ready) {
        e.a = slice_block(src->adj, s, true);
        e.a.nset = e.f.n;
        e.has_adj = true;
      }
      P->reg.push_back(std::move(e));
    }
    if (P->jobs.empty()) {
      // no job runs on the lane modules yet: (re)build them over every
      // registered structure now, so that no request waits for a compilation
      pool_forget_modules(P);
      P->mods_built = false;
      if (P->precision == 64)
        pool_build_modules<double>(P);
      else
        pool_build_modules<float>(P);
    } else if (P->mods_built) {
      // structures the lane modules already run join them; new ones use tiles
      const double eps = P->precision == 64 ? std::numeric_limits<double>::epsilon()
                                            : double(std::numeric_limits<float>::epsilon());
      std::vector<uint32_t> key;
      for (size_t id = first; id < P->reg.size(); ++id) {
        const auto& e = P->reg[id];
        key.clear();
        shape_key(lane_side(e.f, 0, false, eps), key);
        key.push_back(e.has_adj ? 1u : 0u);
        if (e.has_adj) shape_key(lane_side(e.a, 0, true, eps), key);
        auto it = P->key_lane.find(key);
        P->reg_lane.push_back(it == P->key_lane.end() ? -1 : it->second);
      }
    }
    *first_id = uint32_t(first);
    *count = uint32_t(S);
  });
}

}  // extern "C"

namespace {
// One request of registered systems: validated, concatenated, packed and
// admitted (throws Fail; PASIF_ERR_MEMORY when slots or budget refuse it).
// The raw structure of system s of a side: per term kind, each row's term
// count, each term's factor indices relative to the system (and its
// set-point block) and whether its value survives the eps drop, then the
// forced rows.  Equal raw structures give equal lane classes.
void raw_structure(const RowTerms& rt, uint32_t s, bool adjoint, double eps,
                   std::vector<uint32_t>& fp) {
  const uint32_t off = rt.off[s], n = rt.size[s];
  const uint32_t xo = adjoint ? rt.xoff[s] : off;
  fp.push_back(n);
  for (int k = 0; k < 4; ++k) {
    const int nfac = k + 1;
    const uint32_t b = rt.ptr[k][off], e = rt.ptr[k][off + n];
    for (uint32_t r = 0; r < n; ++r) fp.push_back(rt.ptr[k][off + r + 1] - rt.ptr[k][off + r]);
    for (uint32_t t = b; t < e; ++t) {
      const uint32_t* ix = rt.idx[k].data() + size_t(t) * nfac;
      fp.push_back(ix[0] - off);
      for (int a = 1; a < nfac; ++a) fp.push_back(ix[a] - xo);
      fp.push_back(std::fabs(rt.val[k][t]) > eps ? 1u : 0u);
    }
  }
  if (!adjoint)
    for (uint32_t r = 0; r < n; ++r)
      if (std::fabs(rt.force[off + r]) > eps) fp.push_back(0x80000000u | r);
}

// raw structures remembered (the map starts over beyond this; an entry of a
// 64-DOF system holds a few KB)
constexpr size_t kPoolRawKeys = 1u << 13;

// A job carrying its own arrays: each system whose structure a pool lane
// module already runs (the same sparsity, any coefficient values) joins that
// module; the others, and every system of a job with its own excitations
// (the modules read the pool's bank), run on tiles.  With compile_new set,
// the structures of the others that fit one wave are queued for a module
// of their own.
void pool_match_lanes(pasif_pool* P, pasif_driver* d, bool adjoint) {
  d->lane_cls.clear();
  if (!d->sys_ready || (adjoint && !d->adj_ready) || d->ext_exc != P->bank.p)
    return;
  // the registered modules come first (indices 0 and 1): modules compiled
  // for new structures are appended after them
  if (!P->mods_built) {
    if (P->precision == 64)
      pool_build_modules<double>(P);
    else
      pool_build_modules<float>(P);
  }
  if (P->key_lane.empty() && !P->compile_new) return;
  const int bits = P->precision == 64 ? 64 : 32;
  const double eps = P->precision == 64 ? std::numeric_limits<double>::epsilon()
                                        : double(std::numeric_limits<float>::epsilon());
  const uint32_t S = uint32_t(d->fwd.size.size());
  std::vector<int32_t> cls(S, -1);
  std::vector<uint32_t> key;
  bool any = false;
  std::vector<uint32_t> fp;
  for (uint32_t s = 0; s < S; ++s) {
    // the raw structure first: a structure seen before skips the lane-side
    // analysis (its class key is cached)
    fp.clear();
    raw_structure(d->fwd, s, false, eps, fp);
    fp.push_back(adjoint ? 1u : 0u);
    if (adjoint) raw_structure(d->adj, s, true, eps, fp);
    auto hit = P->raw_key.find(fp);
    LaneShape sh;
    bool have_shape = false;
    if (hit == P->raw_key.end()) {
      sh.f = lane_side(d->fwd, s, false, eps);
      key.clear();
      shape_key(sh.f, key);
      const uint32_t nf = uint32_t(key.size());
      key.push_back(adjoint ? 1u : 0u);
      if (adjoint) {
        sh.adjoint = true;
        sh.a = lane_side(d->adj, s, true, eps);
        shape_key(sh.a, key);
      }
      have_shape = true;
      if (P->raw_key.size() >= kPoolRawKeys) P->raw_key.clear();
      hit = P->raw_key.emplace(fp, std::make_pair(key, nf)).first;
    }
    const std::vector<uint32_t>& k = hit->second.first;
    const uint32_t nf = hit->second.second;
    const std::map<std::vector<uint32_t>, int32_t>::const_iterator it =
        adjoint ? P->key_lane.find(k)
                : P->fkey_lane.find(std::vector<uint32_t>(k.begin(), k.begin() + long(nf)));
    if (it != (adjoint ? P->key_lane.end() : P->fkey_lane.end())) {
      cls[s] = it->second;
      any = true;
    } else if (P->compile_new && !P->queued.count(k)) {
      P->queued.insert(k);
      if (!have_shape) {
        sh.f = lane_side(d->fwd, s, false, eps);
        if (adjoint) {
          sh.adjoint = true;
          sh.a = lane_side(d->adj, s, true, eps);
        }
      }
      const uint32_t waves = pool_class_fit(sh, bits);
      if (waves >= 1) {
        pasif_pool::NewClass c;
        c.sh = std::move(sh);
        c.fkey.assign(k.begin(), k.begin() + long(nf));
        c.key = k;
        c.waves = waves;
        P->pending.push_back(std::move(c));
      }
    }
  }
  if (any) d->lane_cls = std::move(cls);
}

// Classes per background build: a module's compile time grows with its
// classes (about 2 s each) and a structure waits for its whole batch, while
// every module is one more launch per tick (launches beyond the device's
// hardware queues run one after another, each lasting the tick's steps): a
// build takes half the queue, between 4 and 16 classes.
constexpr size_t kPoolBuildMin = 4, kPoolBuildMax = 16;
constexpr size_t kPoolMergeModules = 3;

// Start compiling the next batch of queued structures (none running).
template <typename R>
void pool_start_build(pasif_pool* P) {
  if (P->build) return;
  const uint32_t Kreq = P->ck_every ? P->ck_every : 256;
  if (P->pending.empty()) {
    // idle: merge the single-wave classes spread over several background
    // modules into one (at most kLaneMaxClasses of them)
    const std::set<int> mods(P->loose_mod.begin(), P->loose_mod.end());
    if (mods.size() < kPoolMergeModules) return;
    std::string err;
    if (!lane_jit_flags_ok(&err)) return;
    std::unique_ptr<pasif_pool::Build> B(new pasif_pool::Build());
    B->merge = true;
    const size_t n = std::min<size_t>(P->loose.size(), kLaneMaxClasses);
    B->cls.assign(P->loose.begin(), P->loose.begin() + long(n));
    for (const auto& c : B->cls) B->M.shapes.push_back(c.sh);
    B->M.lay = pool_single_layout(B->M.shapes, Kreq, sizeof(R));
    pool_specialise(P, &B->M.lay);
    B->M.lds = lane_lds_bytes(B->M.shapes, int(sizeof(R) * 8), B->M.lay);
    if (B->M.lds > kLaneMaxLds) {  // not expected: each class fitted alone
      P->loose.clear();
      P->loose_mod.clear();
      return;
    }
    B->src = lane_source(B->M.shapes, int(sizeof(R) * 8), B->M.lay);
    pasif_pool::Build* b = B.get();
    B->th = std::thread([b] {
      b->ok = lane_compile_code(b->src, false, &b->code, &b->err);
      b->done.store(true, std::memory_order_release);
    });
    P->build = std::move(B);
    return;
  }
  std::string err;
  if (!lane_jit_flags_ok(&err)) {
    P->pending.clear();
    ++P->failed_builds;
    return;
  }
  std::unique_ptr<pasif_pool::Build> B;
  while (!B && !P->pending.empty()) {
    B.reset(new pasif_pool::Build());
    // single-wave classes a few to a module (queue order); a multi-wave
    // class alone (its compile time grows with its waves)
    const uint32_t waves = P->pending.front().waves;
    const size_t batch =
        std::min(kPoolBuildMax, std::max(kPoolBuildMin, P->pending.size() / 2));
    std::vector<size_t> pick;
    for (size_t i = 0; i < P->pending.size() && pick.size() < batch; ++i)
      if ((P->pending[i].waves == 1) == (waves == 1) && (waves == 1 || pick.empty()))
        pick.push_back(i);
    for (size_t i : pick) B->cls.push_back(std::move(P->pending[i]));
    for (size_t j = pick.size(); j-- > 0;) P->pending.erase(P->pending.begin() + long(pick[j]));
    for (const auto& c : B->cls) B->M.shapes.push_back(c.sh);
    bool fits = true;
    if (waves == 1) {
      B->M.lay = pool_single_layout(B->M.shapes, Kreq, sizeof(R));
    } else {
      fits = pool_multi_layout(B->M.shapes, waves, Kreq, sizeof(R), &B->M.lay);
    }
    if (fits) {
      pool_specialise(P, &B->M.lay);
      B->M.lds = lane_lds_bytes(B->M.shapes, int(sizeof(R) * 8), B->M.lay);
      fits = B->M.lds <= (waves > 1 ? kMaxLdsBytes : kLaneMaxLds);
    }
    if (!fits) {
      ++P->failed_builds;  // these structures stay on tiles
      B.reset();
    }
  }
  if (!B) return;
  B->src = lane_source(B->M.shapes, int(sizeof(R) * 8), B->M.lay);
  pasif_pool::Build* b = B.get();
  B->th = std::thread([b] {
    b->ok = lane_compile_code(b->src, false, &b->code, &b->err);
    b->done.store(true, std::memory_order_release);
  });
  P->build = std::move(B);
}

// Install a finished build as a further module (wait: join a running one
// first), then start the next batch.
void pool_install_build(pasif_pool* P, bool wait) {
  if (!P->build) return;
  if (!wait && !P->build->done.load(std::memory_order_acquire)) return;
  P->build->th.join();
  std::unique_ptr<pasif_pool::Build> B = std::move(P->build);
  bool ok = B->ok;
  if (ok) {
    POOL_HIP(hipSetDevice(P->dev));
    std::string err;
    ok = lane_load_code(B->src, false, P->dev, B->code, &B->M.kern, &err);
    if (ok) {
      P->aux.push_back(nullptr);
      P->join.push_back(nullptr);
      POOL_HIP(hipStreamCreateWithFlags(&P->aux.back(), hipStreamNonBlocking));
      POOL_HIP(hipEventCreateWithFlags(&P->join.back(), hipEventDisableTiming));
      const int m = int(P->mods.size());
      const bool single = B->M.lay.waves == 1;
      P->mods.push_back(std::move(B->M));
      for (size_t k = 0; k < B->cls.size(); ++k) {
        const int32_t code = lane_code(m, uint32_t(k));
        P->key_lane[B->cls[k].key] = code;
        P->fkey_lane[B->cls[k].fkey] = code;
      }
      if (B->merge) {
        // later jobs of these classes run on the merged module; the modules
        // they ran on keep only the jobs already admitted (the merge took
        // the first classes of `loose`, which only a merge removes)
        const size_t n = std::min(B->cls.size(), P->loose.size());
        P->loose.erase(P->loose.begin(), P->loose.begin() + long(n));
        P->loose_mod.erase(P->loose_mod.begin(), P->loose_mod.begin() + long(n));
      } else if (single) {
        for (auto& c : B->cls) {
          P->loose.push_back(std::move(c));
          P->loose_mod.push_back(m);
        }
      }
      ++P->built_modules;
    }
  }
  if (!ok) ++P->failed_builds;
  if (P->precision == 64)
    pool_start_build<double>(P);
  else
    pool_start_build<float>(P);
}

void pool_compile_step(pasif_pool* P) {
  pool_install_build(P, false);
  if (P->precision == 64)
    pool_start_build<double>(P);
  else
    pool_start_build<float>(P);
}

pasif_driver* pool_submit_one(pasif_pool* P, uint32_t n, const uint32_t* ids,
                              const int32_t* exc_of_system, const uint32_t* steps_of_system,
                              int adjoint) {
  POOL_REQUIRE(n > 0 && ids && exc_of_system && steps_of_system, PASIF_ERR_ARGUMENT,
               "a request needs at least one system with its excitation and step count");
  POOL_REQUIRE(P->bank_n > 0, PASIF_ERR_STATE, "no signal bank: pasif_pool_set_bank first");
  for (uint32_t i = 0; i < n; ++i) {
    POOL_REQUIRE(ids[i] < P->reg.size(), PASIF_ERR_ARGUMENT,
                 "system id " + std::to_string(ids[i]) + " is not registered");
    POOL_REQUIRE(!adjoint || P->reg[ids[i]].has_adj, PASIF_ERR_ARGUMENT,
                 "system id " + std::to_string(ids[i]) + " has no registered adjoint");
    POOL_REQUIRE(exc_of_system[i] >= -1 && exc_of_system[i] < int32_t(P->bank_n),
                 PASIF_ERR_ARGUMENT,
                 "excitation " + std::to_string(exc_of_system[i]) + " not in the bank");
    POOL_REQUIRE(steps_of_system[i] > 0, PASIF_ERR_ARGUMENT, "step counts must be positive");
  }
  if (!P->mods_built) {
    if (P->precision == 64)
      pool_build_modules<double>(P);
    else
      pool_build_modules<float>(P);
  }
  const auto c0 = std::chrono::steady_clock::now();
  std::unique_ptr<pasif_driver> d(new_bank_job(P, 0));
  d->lane_cls.assign(n, -1);
  for (uint32_t i = 0; i < n; ++i) d->lane_cls[i] = P->reg_lane[ids[i]];
  std::vector<const RowTerms*> parts(n);
  for (uint32_t i = 0; i < n; ++i) parts[i] = &P->reg[ids[i]].f;
  concat_blocks(parts, false, {}, d->fwd);
  d->fwd.nset = d->fwd.n;
  if (adjoint) {
    for (uint32_t i = 0; i < n; ++i) parts[i] = &P->reg[ids[i]].a;
    concat_blocks(parts, true, d->fwd.off, d->adj);
    d->adj.nset = d->fwd.n;
  }
  d->mapped = true;
  d->map_exc.assign(exc_of_system, exc_of_system + n);
  d->map_steps.assign(steps_of_system, steps_of_system + n);
  d->layout_systems = n;
  d->sys_ready = true;
  d->adj_ready = adjoint != 0;
  const auto c1 = std::chrono::steady_clock::now();
  admit_job(P, d.get(), adjoint);
  pasif_driver* job = d.release();
  P->handles.insert(job);
  if (P->timing) {
    const auto c2 = std::chrono::steady_clock::now();
    P->tm[0] += std::chrono::duration<double, std::milli>(c1 - c0).count();
    P->tm[3] += std::chrono::duration<double, std::milli>(c2 - c1).count();
    ++P->tm_n;
  }
  return job;
}
}  // namespace

extern "C" {

int pasif_pool_submit_systems(pasif_pool* P, uint32_t n, const uint32_t* ids,
                              const int32_t* exc_of_system, const uint32_t* steps_of_system,
                              int adjoint, pasif_driver** job) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(job != nullptr, PASIF_ERR_ARGUMENT, "null job pointer");
    *job = nullptr;
    *job = pool_submit_one(P, n, ids, exc_of_system, steps_of_system, adjoint);
  });
}

int pasif_pool_submit_batch(pasif_pool* P, uint32_t n_req, const uint32_t* sys_off,
                            const uint32_t* ids, const int32_t* exc_of_system,
                            const uint32_t* steps_of_system, const int32_t* adjoint,
                            pasif_driver** jobs, uint32_t* admitted) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(admitted != nullptr && jobs != nullptr && (n_req == 0 || (sys_off && adjoint)),
                 PASIF_ERR_ARGUMENT, "null argument");
    *admitted = 0;
    for (uint32_t r = 0; r < n_req; ++r) jobs[r] = nullptr;
    for (uint32_t r = 0; r < n_req; ++r) {
      POOL_REQUIRE(sys_off[r + 1] > sys_off[r], PASIF_ERR_ARGUMENT,
                   "request " + std::to_string(r) + " has no system");
      const uint32_t b = sys_off[r];
      try {
        jobs[r] = pool_submit_one(P, sys_off[r + 1] - b, ids + b, exc_of_system + b,
                                  steps_of_system + b, adjoint[r]);
      } catch (const Fail& f) {
        if (f.code == PASIF_ERR_MEMORY) return;   // full: the rest waits
        throw;
      }
      *admitted = r + 1;
    }
  });
}

int pasif_pool_poll(pasif_pool* P, int* running) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(running != nullptr, PASIF_ERR_ARGUMENT, "null output");
    *running = 0;
    if (!P->in_flight) return;
    const hipError_t e = hipEventQuery(P->ev1);
    if (e == hipErrorNotReady) {
      *running = 1;
      return;
    }
    POOL_HIP(e);
  });
}

int pasif_pool_tick(pasif_pool* P, int wait, uint32_t* active_tiles) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(!P->in_flight, PASIF_ERR_STATE, "a tick is already in flight");
    POOL_HIP(hipSetDevice(P->dev));
    pool_reap(P);
    pool_compile_step(P);
    if (P->precision == 64)
      pool_tick_impl<double>(P);
    else
      pool_tick_impl<float>(P);
    if (active_tiles) *active_tiles = P->last_blocks;
    if (wait && P->in_flight) {
      POOL_HIP(hipEventSynchronize(P->ev1));
      float ms = 0.f;
      POOL_HIP(hipEventElapsedTime(&ms, P->ev0, P->ev1));
      P->last_ms = ms;
      P->in_flight = false;
    }
  });
}

int pasif_pool_wait(pasif_pool* P, double* kernel_ms) {
  return pool_guarded(P, [&] {
    if (P->in_flight) {
      POOL_HIP(hipEventSynchronize(P->ev1));
      float ms = 0.f;
      POOL_HIP(hipEventElapsedTime(&ms, P->ev0, P->ev1));
      P->last_ms = ms;
      P->in_flight = false;
    }
    pool_reap(P);
    if (kernel_ms) *kernel_ms = P->last_ms;
  });
}

int pasif_pool_job_finished(pasif_pool* P, const pasif_driver* d, int* finished) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(finished != nullptr, PASIF_ERR_ARGUMENT, "null output");
    *finished = 0;
    for (auto& j : P->jobs)
      if (j.d == d) {
        *finished = (j.finished || (!P->in_flight && j.done >= j.work)) ? 1 : 0;
        return;
      }
    POOL_REQUIRE(false, PASIF_ERR_ARGUMENT, "job not in this pool");
  });
}

int pasif_pool_collect(pasif_pool* P, pasif_driver* d, double* q, size_t capq, size_t* lenq,
                       double* grad, size_t capg, size_t* leng) {
  return pool_guarded(P, [&] {
    size_t k = 0;
    for (; k < P->jobs.size(); ++k)
      if (P->jobs[k].d == d) break;
    POOL_REQUIRE(k < P->jobs.size(), PASIF_ERR_ARGUMENT, "job not in this pool");
    auto& j = P->jobs[k];
    // a job in the running tick waits for it; one the tick left out does not
    POOL_REQUIRE(!P->in_flight || j.finished, PASIF_ERR_STATE, "wait for the tick first");
    POOL_REQUIRE(j.done >= j.work, PASIF_ERR_STATE, "job not finished");
    POOL_HIP(hipSetDevice(P->dev));
    // every tile must have reached its last phase
    if (d->pk.tile) {
      std::vector<uint32_t> prog(size_t(d->pk.ntiles) * 4);
      POOL_HIP(hipMemcpy(prog.data(), d->progress.p, prog.size() * 4, hipMemcpyDeviceToHost));
      for (uint32_t t = 0; t < d->pk.ntiles; ++t) {
        const bool done = j.adjoint ? prog[4 * t + 1] == kPhaseDone
                                    : prog[4 * t] >= d->pk.tmax || prog[4 * t + 1] == kPhaseDone;
        POOL_REQUIRE(done, PASIF_ERR_STATE, "job tiles did not finish");
      }
    }
    for (const auto& lp : d->pk.lparts)
      for (const auto& L : lp->cls) {
        std::vector<uint32_t> prog(size_t(L->waves) * 4);
        POOL_HIP(hipMemcpy(prog.data(), L->progress.p, prog.size() * 4, hipMemcpyDeviceToHost));
        for (uint32_t w = 0; w < L->waves; ++w)
          POOL_REQUIRE(prog[4 * w + 1] == kPhaseDone, PASIF_ERR_STATE,
                       "job lane workgroups did not finish");
      }
    std::vector<double> grad_v;
    const int rc = guarded(d, [&] {
      if (P->precision == 64)
        finish_run<double>(d, j.adjoint, 0, nullptr, j.adjoint ? &grad_v : nullptr);
      else
        finish_run<float>(d, j.adjoint, 0, nullptr, j.adjoint ? &grad_v : nullptr);
      copy_out(d, d->res_q, q, capq, lenq);
      if (j.adjoint) copy_out(d, grad_v, grad, capg, leng);
      else if (leng) *leng = 0;
    });
    if (rc != PASIF_OK) {
      P->err = d->err;
      throw Fail{rc};
    }
    pool_remove_job(P, k);
  });
}

int pasif_pool_collect_batch(pasif_pool* P, uint32_t n, pasif_driver* const* jobs, double* q,
                             const size_t* q_off, double* grad, const size_t* g_off) {
  return pool_guarded(P, [&] {
    POOL_REQUIRE(n == 0 || (jobs && q && q_off && g_off), PASIF_ERR_ARGUMENT, "null argument");
    if (n == 0) return;
    POOL_HIP(hipSetDevice(P->dev));
    const size_t el = P->precision == 64 ? 8 : 4;
    // every job: finished, its slots sized right; then its progress records
    // and outputs copied into one pinned buffer with a single wait
    std::vector<size_t> at(n);
    struct Piece {
      const void* src;
      size_t bytes, off;
    };
    std::vector<Piece> pieces;
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes) {
      pieces.push_back({src, bytes, total});
      total += (bytes + 15) & ~size_t(15);
    };
    for (uint32_t i = 0; i < n; ++i) {
      size_t k = 0;
      for (; k < P->jobs.size(); ++k)
        if (P->jobs[k].d == jobs[i]) break;
      POOL_REQUIRE(k < P->jobs.size(), PASIF_ERR_ARGUMENT,
                   "job " + std::to_string(i) + " is not in this pool");
      const auto& j = P->jobs[k];
      POOL_REQUIRE(!P->in_flight || j.finished, PASIF_ERR_STATE, "wait for the tick first");
      POOL_REQUIRE(j.done >= j.work, PASIF_ERR_STATE,
                   "job " + std::to_string(i) + " not finished");
      const pasif_driver* d = j.d;
      POOL_REQUIRE(q_off[i + 1] - q_off[i] == d->pk.out_len_f, PASIF_ERR_ARGUMENT,
                   "job " + std::to_string(i) + ": " + std::to_string(d->pk.out_len_f) +
                       " final states, slot of " + std::to_string(q_off[i + 1] - q_off[i]));
      if (j.adjoint)
        POOL_REQUIRE(grad && g_off[i + 1] - g_off[i] == d->pk.out_len_a, PASIF_ERR_ARGUMENT,
                     "job " + std::to_string(i) + ": gradient slot size");
      at[i] = pieces.size();
      if (d->pk.tile) add(d->progress.p, size_t(d->pk.ntiles) * 4 * sizeof(uint32_t));
      for (const auto& lp : d->pk.lparts)
        for (const auto& L : lp->cls) add(L->progress.p, size_t(L->waves) * 4 * sizeof(uint32_t));
      add(d->out_q.p, size_t(d->pk.out_len_f) * el);
      if (j.adjoint) add(d->out_grad.p, size_t(d->pk.out_len_a) * el);
    }
    if (P->pinned_bytes < total) {
      if (P->pinned) (void)hipHostFree(P->pinned);
      P->pinned = nullptr;
      P->pinned_bytes = 0;
      const size_t want = std::max(total, size_t(1) << 20) * 2;
      POOL_HIP(hipHostMalloc(reinterpret_cast<void**>(&P->pinned), want, hipHostMallocDefault));
      P->pinned_bytes = want;
    }
    for (const Piece& c : pieces)
      POOL_HIP(hipMemcpyAsync(P->pinned + c.off, c.src, c.bytes, hipMemcpyDeviceToHost,
                              P->upload));
    POOL_HIP(hipStreamSynchronize(P->upload));
    for (uint32_t i = 0; i < n; ++i) {
      size_t k = 0;
      for (; k < P->jobs.size(); ++k)
        if (P->jobs[k].d == jobs[i]) break;
      const bool adj = P->jobs[k].adjoint;
      const pasif_driver* d = P->jobs[k].d;
      size_t x = at[i];
      auto prog = [&](uint32_t recs) {
        return reinterpret_cast<const uint32_t*>(P->pinned + pieces[x++].off);
      };
      if (d->pk.tile) {
        const uint32_t* pr = prog(d->pk.ntiles);
        for (uint32_t t = 0; t < d->pk.ntiles; ++t) {
          const bool done = adj ? pr[4 * t + 1] == kPhaseDone
                                : pr[4 * t] >= d->pk.tmax || pr[4 * t + 1] == kPhaseDone;
          POOL_REQUIRE(done, PASIF_ERR_STATE, "job tiles did not finish");
        }
      }
      for (const auto& lp : d->pk.lparts)
        for (const auto& L : lp->cls) {
          const uint32_t* pr = prog(L->waves);
          for (uint32_t w = 0; w < L->waves; ++w)
            POOL_REQUIRE(pr[4 * w + 1] == kPhaseDone, PASIF_ERR_STATE,
                         "job lane workgroups did not finish");
        }
      auto widen = [&](const Piece& c, double* out, size_t len) {
        if (el == 8) {
          std::memcpy(out, P->pinned + c.off, len * sizeof(double));
        } else {
          const float* f = reinterpret_cast<const float*>(P->pinned + c.off);
          for (size_t e = 0; e < len; ++e) out[e] = double(f[e]);
        }
      };
      widen(pieces[x++], q + q_off[i], d->pk.out_len_f);
      if (adj) widen(pieces[x++], grad + g_off[i], d->pk.out_len_a);
      pool_remove_job(P, k);
      pasif_destroy(jobs[i]);
    }
  });
}

int pasif_pool_stats(const pasif_pool* P, uint32_t* jobs, double* used_bytes,
                     double* last_kernel_ms, uint32_t* last_blocks) {
  if (!P) return PASIF_ERR_ARGUMENT;
  if (jobs) *jobs = uint32_t(P->jobs.size());
  if (used_bytes) *used_bytes = P->used;
  if (last_kernel_ms) *last_kernel_ms = P->last_ms;
  if (last_blocks) *last_blocks = P->last_blocks;
  return PASIF_OK;
}

int pasif_pool_slots(const pasif_pool* P, uint32_t* used, uint32_t* max_tiles) {
  if (!P) return PASIF_ERR_ARGUMENT;
  if (used) *used = P->tiles_used;
  if (max_tiles) *max_tiles = P->max_tiles;
  return PASIF_OK;
}

int pasif_pool_destroy(pasif_pool* P) {
  if (!P) return PASIF_OK;
  if (P->build && P->build->th.joinable()) P->build->th.join();
  if (P->timing && P->tm_n)
    fprintf(stderr,
            "[pasif] admission of %llu requests, ms per request: concatenation %.3f, "
            "packing %.3f (of which lane program uploads %.3f), run buffers and arguments "
            "%.3f, total in admit %.3f\n",
            (unsigned long long)P->tm_n, P->tm[0] / double(P->tm_n), P->tm[1] / double(P->tm_n),
            P->tm[4] / double(P->tm_n), P->tm[2] / double(P->tm_n), P->tm[3] / double(P->tm_n));
  (void)hipSetDevice(P->dev);
  if (P->stream) (void)hipStreamSynchronize(P->stream);
  if (P->upload) (void)hipStreamSynchronize(P->upload);
  // orphaned jobs are deleted; every other handle the pool created stays
  // with its owner, detached (pasif_destroy it as usual)
  for (auto& j : P->jobs)
    if (j.d->orphan) {
      P->handles.erase(j.d);
      pool_detach(j.d);
      delete j.d;
    }
  P->jobs.clear();
  for (pasif_driver* d : P->handles) pool_detach(d);
  P->handles.clear();
  P->arena.reset();
  if (P->pinned) (void)hipHostFree(P->pinned);
  P->pinned = nullptr;
  for (size_t m = 0; m < P->aux.size(); ++m) {
    if (P->aux[m]) {
      (void)hipStreamSynchronize(P->aux[m]);
      (void)hipStreamDestroy(P->aux[m]);
    }
    if (P->join[m]) (void)hipEventDestroy(P->join[m]);
  }
  if (P->fork) (void)hipEventDestroy(P->fork);
  for (auto& M : P->mods)
    if (M.table_dev) (void)hipFree(M.table_dev);
  if (P->table_dev) (void)hipFree(P->table_dev);
  if (P->map_dev) (void)hipFree(P->map_dev);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  if (P->uploaded) (void)hipEventDestroy(P->uploaded);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  if (P->upload) (void)hipStreamDestroy(P->upload);
  delete P;
  return PASIF_OK;
}

}  // extern "C"
