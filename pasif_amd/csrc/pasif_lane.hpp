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
  X(void*, qstate)         /* [NF][stride] state between launches */        \
  X(void*, lstate)         /* [NA][stride] */                               \
  X(void*, mxstate)        /* [NF][stride] running max |q| */               \
  X(void*, progress)       /* u32 [waves][4]: step, phase, segment */       \
  X(void*, ckpt)           /* [nck][NF][stride] checkpoints */              \
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

struct LaneShape {
  LaneSide f, a;
  bool adjoint = false;
  // doubles the larger of the forward and adjoint phases keeps live
  uint32_t live_values() const;
  // Multi-wave layout (kernel waves > 1): the wave that owns each forward /
  // adjoint row.  Every wave of a workgroup works on the same instances
  // (one per lane); stage values go through LDS.
  std::vector<uint32_t> fwave, awave;
};

// Kernel-level layout.  waves == 1: one instance per lane, the whole state
// in registers, independent waves.  waves > 1: a workgroup of `waves` waves
// runs `lw` instances (lanes >= lw mirror lane % lw), rows split across the
// waves; `seg` is the adjoint segment length (checkpoint interval) in LDS.
struct LaneLayout {
  uint32_t waves = 1, lw = 64, seg = 1;
};

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
bool lane_compile(const std::string& src, int dev, LaneKernel* out, std::string* err);
// hipRTC compilation only (no device needed): size of the code object
bool lane_compile_only(const std::string& src, size_t* code_bytes, std::string* err);

}  // namespace pasif
