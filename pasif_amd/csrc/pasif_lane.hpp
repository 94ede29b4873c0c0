// Lane kernels: one lane integrates one (system, excitation) instance with
// its whole state and its coefficients in registers.  Instances are grouped
// into classes of identical sparsity structure; for each set of classes a
// kernel is generated with every class's right-hand side unrolled into
// straight-line code (compile-time row indices), compiled with hipRTC for
// gfx950 at pack time and cached (per process and on disk).
//
// Compared with the tile kernel there is no LDS exchange and no barrier per
// RK stage: a step is a dependent chain of FMAs spread over the rows of the
// system, which is what bounds small systems.  Large systems (more state and
// coefficients than the register file holds) stay on the tile kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace pasif {

// Kernel arguments of one class.  Every field is 8 bytes so the host and the
// generated device code (compiled separately) agree on the layout; the field
// list is shared through PASIF_LANE_FIELDS.
#define PASIF_LANE_FIELDS(X)                                                 \
  X(const void*, coef_f)   /* [TF][stride] forward coefficients */          \
  X(const void*, coef_a)   /* [TA][stride] adjoint coefficients */          \
  X(const void*, init_f)   /* [NF][stride] initial state */                 \
  X(const void*, init_a)   /* [NA][stride] terminal adjoint state */        \
  X(const void*, steps)    /* u32 [stride], descending within the class */  \
  X(const void*, exc)      /* i32 [stride], -1 = unforced */                \
  X(const void*, out_f)    /* u32 [stride] output base, forward rows */     \
  X(const void*, out_a)    /* u32 [stride] output base, adjoint rows */     \
  X(const void*, na_out)   /* u32 [stride] adjoint rows the instance owns    \
                              (a merged class may pad its adjoint) */       \
  X(void*, qstate)         /* [NF][stride] state between launches */        \
  X(void*, lstate)         /* [NA][stride] */                               \
  X(void*, mxstate)        /* [NF][stride] running max |q| */               \
  X(void*, progress)       /* u32 [waves][4]: step, phase, segment */       \
  X(void*, ckpt)           /* checkpoints (single-wave: per-wave blocks) */ \
  X(const void*, ck_off)   /* u64 [waves] first element of each wave's      \
                              checkpoint block [nck_w][NF][64] (single-wave) */ \
  X(void*, segbuf)         /* [K+1][NF][stride] recomputed segment */       \
  X(void*, out_q)                                                            \
  X(void*, out_max)                                                          \
  X(void*, traj)                                                             \
  X(void*, out_grad)                                                         \
  X(const void*, excit)    /* [n_exc][exc_len] */                           \
  X(const void*, interp)   /* [2][interp_w] or null */                      \
  X(uint64_t, stride)      /* instance slots (multiple of 64) */            \
  X(uint64_t, count)       /* instances */                                  \
  X(uint64_t, block_begin) /* first workgroup of the class */               \
  X(uint64_t, klass)       /* the class's case in the kernel's dispatch */  \
  X(uint64_t, exc_len)                                                       \
  X(uint64_t, interp_w)                                                      \
  X(uint64_t, budget)                                                        \
  X(uint64_t, window)                                                        \
  X(uint64_t, out_len)                                                       \
  X(uint64_t, save_every)                                                    \
  X(uint64_t, nslots)                                                        \
  X(uint64_t, ck_every)                                                      \
  X(uint64_t, nck)                                                           \
  X(uint64_t, adjoint)                                                       \
  X(double, h)             /* step constants in the working precision */    \
  X(double, h2)                                                              \
  X(double, h6)

struct LaneArgs {
#define PASIF_LANE_DECL(T, name) T name;
  PASIF_LANE_FIELDS(PASIF_LANE_DECL)
#undef PASIF_LANE_DECL
};

// One term of a class's row: kind 0..3 (A, Gamma, Lambda, Psi) and its
// factor rows (f[0] in the side's own state, f[1..kind] in the set-point
// state: the forward state itself for the forward side).
struct LaneTerm {
  uint32_t kind;
  uint32_t f[4];
};

struct LaneSide {
  uint32_t n = 0;
  std::vector<std::vector<LaneTerm>> rows;  // terms of each row, coefficient order
  std::vector<uint32_t> forced;             // rows driven by the excitation
  uint32_t coefs() const;                   // coefficient slots incl. forcing
};

// Lane groups: G lanes (a power of two, aligned within the wave) integrate
// one instance.  A side's rows are laid out in bands of G rows (band slot r
// of position p holds row band[r] + p * stride[r]) and shared rows (every
// lane of the group holds the same value).  The right-hand side of a band
// slot is one program for all positions (the union of the positions' terms,
// each lane with its own coefficients, zero where its row lacks the term);
// a factor owned by another position comes through a lane rotation within
// the group.  A shared row's terms are spread over the positions and summed
// across the group with an xor butterfly (every lane gets the same sum).
struct GSrc {
  bool shared = false;
  uint32_t slot = 0;  // lane slot (band slots first, then shared rows)
  uint32_t d = 0;     // rotation: the value of position (p + d) mod G
  bool operator<(const GSrc& o) const {
    return shared != o.shared ? shared < o.shared : (slot != o.slot ? slot < o.slot : d < o.d);
  }
  bool operator==(const GSrc& o) const { return shared == o.shared && slot == o.slot && d == o.d; }
};
struct GTerm {
  uint32_t kind = 0;
  GSrc f[4];
  std::vector<int32_t> coef;  // [G] system coefficient slot per position, -1 = none
};
struct GRow {
  std::vector<GTerm> terms;
  std::vector<int32_t> force;  // [G] force slot per position, empty if no position is forced
};
struct GSide {
  uint32_t RL = 0, NS = 0;
  // band slot r of position p holds row band[r] + p * stride[r]: stride 1
  // for G consecutive rows, m for the j-th slot of a block of G m rows in
  // which position p holds rows m p .. m p + m - 1
  std::vector<uint32_t> band;    // first row of each band slot
  std::vector<uint32_t> stride;  // row step between positions of each band slot
  std::vector<uint32_t> shared;  // row of each shared slot
  std::vector<GRow> rows;        // [RL + NS]
  // shared slots no term reads (forward: nor the adjoint's set points): the
  // group sums their slopes once per step instead of once per stage
  std::vector<uint8_t> sink;     // [RL + NS]
  uint32_t nl() const { return RL + NS; }
  uint32_t coefs() const;        // per-lane coefficient slots: row by row, terms then force
};

struct LaneShape {
  LaneSide f, a;
  bool adjoint = false;
  // doubles the larger of the forward and adjoint phases keeps live
  uint32_t live_values() const;
  // Multi-wave layout (kernel waves > 1): the wave that owns each forward /
  // adjoint row.  Every wave of a workgroup works on the same instances
  // (one per lane); stage values go through LDS.
  std::vector<uint32_t> fwave, awave;
  // Lane-group layout (group > 1, single-wave kernels): the programs of both
  // sides and the forward values (set points) the adjoint side reads.
  uint32_t group = 1;
  GSide gf, ga;
  std::vector<GSrc> gx;
  // single-wave kernels: this class's checkpoint interval (0: the kernel's)
  uint32_t seg = 0;
};

// Lay `sh` out for groups of G lanes (G = 1 restores one lane per instance).
// False if G does not divide the wave or no band fits.
bool lane_group(LaneShape& sh, uint32_t G);
// values (of the working type) one lane keeps live in the larger phase
uint32_t lane_live(const LaneShape& sh);
// adjoint runs: the forward lane slots the reverse step reads as set points,
// which are all the recomputed segment in LDS holds.  map[r] = the slot's row
// in the segment or -1; returns the segment's rows per step (at least 1).
uint32_t lane_seg_rows(const LaneShape& sh, std::vector<int>* map = nullptr);
// estimated vector instructions per step of one wave: forward step, and
// forward + recompute + reverse step of an adjoint run
double lane_step_cost(const LaneShape& sh, int bits);
// system coefficient slots (lane_coefs order) -> the lane coefficients of
// position p of a group side (0 where the position has no such term)
void lane_group_coefs(const GSide& gs, const std::vector<double>& sys, uint32_t p,
                      std::vector<double>& out);

// Kernel-level layout.  waves == 1: one instance per lane (or per group of
// lanes, LaneShape::group), the whole state in registers, independent
// waves.  waves > 1: a workgroup of `waves` waves
// runs `lw` instances (lanes >= lw mirror lane % lw), rows split across the
// waves; `seg` is the adjoint segment length (checkpoint interval) in LDS.
struct LaneLayout {
  uint32_t waves = 1, lw = 64, seg = 1;
  // single-wave kernels: waves each SIMD should hold at once (the register
  // budget the compiler gets is 512 / occ VGPRs); 0 or 1: no target
  uint32_t occ = 0;
  // single-wave kernels whose waves each have a SIMD to themselves: a lone
  // wave's dependency chains bound them, so they are compiled with the
  // scheduler that maximises instruction-level parallelism
  bool ilp = false;
  // single-wave adjoint kernels: the exact discrete adjoint (the recompute
  // keeps every step's stage inputs y2..y4 beside the post-step states)
  bool exact = false;
  // step loops compiled: forcing 0 = both forcing rules (chosen at run time
  // from the interpolation window), 1 = interpolation only, 2 = zero-order
  // hold only; full = false leaves out the loops with per-step outputs
  // (trajectory, window maximum), which a caller that never asks for them
  // (the pool) does not need; smaller code compiles faster
  uint32_t forcing = 0;
  bool full = true;
};
// Values per lane and set-point row that a checkpoint segment of K steps
// keeps: K + 1 post-step states, and three stage inputs per step for the
// exact discrete adjoint.
inline size_t lane_seg_values(uint32_t K, bool exact) {
  return size_t(K) + 1 + (exact ? 3 * size_t(K) : 0);
}

// Split the rows of `sh` over `waves` waves (largest term cost first onto
// the least loaded wave) and return the largest per-wave live estimate
// (coefficients of the larger phase plus two values per owned row).
uint32_t lane_partition(LaneShape& sh, uint32_t waves);

// samples of the excitation staged per lane, per chunk of steps
constexpr uint32_t kLaneChunk = 16;
constexpr uint32_t kLaneStage = kLaneChunk + 3;
constexpr uint32_t kLaneWave = 64;

// Generated source of one kernel covering `shapes` (class c runs the
// workgroups of LaneArgs c) in precision `bits` with layout `lay`.
std::string lane_source(const std::vector<LaneShape>& shapes, int bits, const LaneLayout& lay);
// LDS bytes of that kernel (dynamic for waves == 1, static otherwise)
size_t lane_lds_bytes(const std::vector<LaneShape>& shapes, int bits, const LaneLayout& lay);

struct LaneKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};
// Compile (or fetch from the caches) the kernel of `src` for device `dev`.
// Returns false with a message in *err on failure.
bool lane_compile(const std::string& src, int dev, LaneKernel* out, std::string* err,
                  bool ilp = false);
// Several kernels at once (one per module of a lane plan): those in neither
// cache are compiled on worker threads, one program each (one after the
// other when the options carry -mllvm flags, which LLVM keeps in
// process-wide state).  dev < 0 compiles only (no cache, no module load;
// outs may hold null).
bool lane_compile_many(const std::vector<std::string>& srcs, const std::vector<bool>& ilp,
                       int dev, const std::vector<LaneKernel*>& outs, std::string* err);
// hipRTC compilation only (no device needed): size of the code object
bool lane_compile_only(const std::string& src, size_t* code_bytes, std::string* err,
                       bool ilp = false);
// checks the caller's PASIF_JIT_FLAGS against the accepted options
bool lane_jit_flags_ok(std::string* err);
// Two halves of lane_compile for a caller that compiles on a thread of its
// own: the code object of `src` (the disk cache, else hipRTC; no device
// needed; serialised with the library's other compilations), then its
// module loaded on `dev` (from the thread that owns the device's work).
bool lane_compile_code(const std::string& src, bool ilp, std::vector<char>* code,
                       std::string* err);
bool lane_load_code(const std::string& src, bool ilp, int dev, const std::vector<char>& code,
                    LaneKernel* out, std::string* err);

}  // namespace pasif
