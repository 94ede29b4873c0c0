// Tile program format shared by the host runtime (pasif_driver.cpp) and the
// gfx950 kernels (pasif_kernels.hip).
//
// A tile is one workgroup's worth of independent systems ("instances": a
// system paired with the excitation that drives it).  Every row (DOF) of an
// instance is owned by one lane; a row with more terms than the lane's slot
// budget is split over several consecutive lanes ("segments") whose partial
// sums the owner adds in lane order.  Each lane holds the coefficients and
// packed LDS indices of its segment's terms in registers for the whole solve;
// the stage state y lives in LDS, one double per row, double-buffered across
// the four RK4 stages.
//
// Term kinds by number of state factors:
//   kind 1: A[o,j] S_j                       (CSR B+K)
//   kind 2: G[o,j,k] S_j X_k                 (Gamma)
//   kind 3: L[o,j,k,l] S_j X_k X_l           (Lambda)
//   kind 4: P[o,j,k,l,m] S_j X_k X_l X_m     (Psi)
// S is the integrated state, X the set point (S itself in the forward pass,
// the stored forward trajectory in the adjoint pass).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pasif {

constexpr uint32_t kRowIdle = 0xFFFFu;   // lane without a segment
constexpr uint32_t kMaxLanes = 1024;
constexpr size_t kMaxLdsBytes = 160 * 1024;   // LDS per CU on gfx950
constexpr uint32_t kMaxInterpWindow = 2048;
// excitation samples staged in LDS per forced row and per chunk of steps
constexpr uint32_t kForceChunk = 64;
constexpr uint32_t kForceStage = kForceChunk + 3;

// per-lane word: [0,16) tile-local row, [16,26) segment count (owner lanes
// only), bit 31 owner flag
__host__ __device__ inline uint32_t lane_row(uint32_t w) { return w & 0xFFFFu; }
__host__ __device__ inline uint32_t lane_nseg(uint32_t w) { return (w >> 16) & 0x3FFu; }
__host__ __device__ inline bool lane_owner(uint32_t w) { return (w >> 31) != 0u; }
__host__ __device__ inline uint32_t make_lane(uint32_t row, uint32_t nseg, bool owner) {
  return (row & 0xFFFFu) | ((nseg & 0x3FFu) << 16) | (owner ? 0x80000000u : 0u);
}

// slot-width configurations (slots per lane for kinds 1..4)
struct SlotConfig { int c1, c2, c3, c4; };
// slot configuration ids: 1 narrow, 2 wide, 3 small
constexpr SlotConfig kSlotNarrow{4, 4, 2, 2};
constexpr SlotConfig kSlotWide{8, 8, 4, 2};
constexpr SlotConfig kSlotSmall{4, 2, 2, 1};

__host__ __device__ constexpr int slot_values(SlotConfig c) { return c.c1 + c.c2 + c.c3 + c.c4; }
// index words hold LDS byte offsets: kind 1 -> 1 (S), kind 2 -> 1
// (S | X<<16), kind 3 -> 2 (S | X1<<16, X2), kind 4 -> 2 (S | X1<<16,
// X2 | X3<<16).  S offsets address the stage double buffer [row][2]; X
// offsets address it too in the forward program and the set-point buffer
// [row][4] in the adjoint program.
__host__ __device__ constexpr int slot_words(SlotConfig c) { return c.c1 + c.c2 + 2 * c.c3 + 2 * c.c4; }

struct alignas(16) TileHeader {
  uint32_t rows_f;     // forward rows in the tile
  uint32_t rows_a;     // adjoint rows
  uint32_t cnt_f[4];   // slot trip counts of the forward program
  uint32_t cnt_a[4];   // slot trip counts of the adjoint program
  uint32_t split_f;    // some forward row spans several lanes
  uint32_t split_a;
  uint32_t steps;      // largest step count among the tile's instances
  uint32_t nforced;    // forced rows in the tile (LDS sample slots)
};
static_assert(sizeof(TileHeader) == 64, "tile header layout");

// Per-tile program of one pass (forward or adjoint).  Arrays are
// [tile][...][lanes]; `lanes` is the launch's workgroup size.
template <typename R>
struct Program {
  const uint32_t* info;   // [tile][lanes] lane word
  const uint32_t* out;    // [tile][lanes] owner: output index
  const uint32_t* src;    // [tile][lanes] owner: initial-condition index
  const uint32_t* steps;  // [tile][lanes] owner: step count of the instance
  const int32_t* exc;     // [tile][lanes] owner: excitation id or -1
  const uint32_t* fslot;  // [tile][lanes] forced owner: LDS sample slot
  const R* force;         // [tile][lanes] owner: force pattern value
  const R* val;           // [tile][slot_values][lanes]
  const uint32_t* idx;    // [tile][slot_words][lanes]
  const R* init;          // initial state (forward q0 / adjoint lambda_T)
};

// adjoint-pass phases of a tile (stored per tile in `phase`)
enum : uint32_t { kPhaseForward = 0, kPhaseRecompute = 1, kPhaseBackward = 2, kPhaseDone = 3 };

template <typename R>
struct TileArgs {
  const TileHeader* hdr;
  uint32_t lanes;          // workgroup size == program stride
  Program<R> fwd;
  Program<R> adj;
  // forcing
  const R* excit;          // [n_exc][exc_len]
  uint32_t exc_len;
  const R* interp;         // [2][interp_w] rows 0 and 1, or null (zero-order hold)
  uint32_t interp_w;
  uint32_t max_forced;     // largest nforced over the tiles (LDS sizing)
  R h, h2, h6;
  // resumable state, [tile][lanes] by row
  R* qstate;
  R* lstate;               // adjoint lambda
  R* mxstate;
  uint32_t* progress;      // [tile][4]: step, phase, segment, spare
  uint32_t budget;         // steps of work per launch
  // outputs
  R* out_q;
  R* out_max;
  uint32_t window;
  R* traj;
  uint32_t out_len;
  uint32_t save_every;
  uint32_t nslots;
  R* out_grad;
  // adjoint scratch, [tile][...][lanes] by forward row
  R* ckpt;                 // per tile ceil(steps/ck_every) slots of [lanes], from ck_tile
  const uint32_t* ck_tile; // [tile] first checkpoint slot of the tile (sized by its steps)
  R* segbuf;               // [tile][ck_every+1][lanes]; exact: [tile][4 (ck_every+1)][lanes]
  uint32_t ck_every;
  uint32_t exact;          // adjoint: 0 the reference's recursion, 1 the exact discrete adjoint
  uint32_t adjoint;        // run the adjoint program after the forward pass
};

// ---------------------------------------------------------------------------
// Row-parallel path for systems larger than a workgroup's lanes: one
// workgroup per instance, rows strided over its threads, every vector of the
// instance in global memory (L2-resident), a barrier between stages.
//
// Terms of a side, per instance, row-major and kind-ordered: the terms of
// row r and kind k are [rptr[4r+k], rptr[4r+k+1]) of val / idx, with idx[4e]
// the S factor row and idx[4e+1..k] the set-point factor rows (local rows).
template <typename R>
struct BigSide {
  const uint32_t* rptr;   // [rows * 4 + 1] first term of each (row, kind)
  const R* val;           // [terms]
  const uint32_t* idx;    // [terms][4]
  const R* force;         // [rows] (forward side; 0 where unforced)
  const R* init;          // [rows] initial state / terminal adjoint state
  // segments of at most kBigSegTerms terms of one (row, kind): [segs][3]
  // (first term, end, kind) and the first segment of each (row, kind)
  const uint32_t* seg;
  const uint32_t* sptr;   // [rows * 4 + 1]
};

constexpr uint32_t kBigSegTerms = 8;
// segment partial sums kept in LDS up to this many bytes per workgroup
constexpr size_t kBigPartLds = 64 * 1024;

struct BigInst {
  uint64_t rptr_f, term_f, rptr_a, term_a;  // offsets into the packed sides
  uint64_t row_f, row_a;                     // offsets of force / init rows
  uint64_t scr_f, scr_a, ck, seg;            // offsets into the scratch arrays
  uint64_t sptr_f, sseg_f, sptr_a, sseg_a;   // offsets of the segment lists
  uint64_t part_f, part_a;                   // offsets of the segment partial sums
  uint32_t n_f, n_a, steps, out_f, out_a;
  int32_t exc;
  uint32_t nseg_f, nseg_a;                   // segments of each side
  uint32_t split_f, split_a;                 // some (row, kind) has several segments
};

template <typename R>
struct BigArgs {
  const BigInst* inst;
  BigSide<R> fwd, adj;      // packed sides (base pointers; per-instance offsets)
  const R* excit;           // [n_exc][exc_len]
  uint32_t exc_len;
  const R* interp;          // [2][interp_w] or null (zero-order hold)
  uint32_t interp_w;
  R h, h2, h6;
  R* vec_f;                 // per instance [4][n_f]: q, y0, y1, acc
  R* vec_a;                 // per instance [4][n_a]: lam, y0, y1, acc
  R* setp;                  // per instance [4][n_f]: set points of the reverse step
  R* ckpt;                  // per instance [ceil(steps/K)][n_f]
  R* segbuf;                // per instance [K+1][n_f] or (exact) [K][4][n_f]
  R* part;                  // per instance: segment partial sums of each side
  uint32_t part_lds;        // partial sums in LDS (every side's segments fit)
  uint32_t* progress;       // [inst][4]: step, phase, segment
  R* out_q;
  R* out_grad;
  uint32_t budget;
  uint32_t ck_every;
  uint32_t adjoint;
  uint32_t exact;
};

}  // namespace pasif

namespace pasif {
// kernel launchers (pasif_kernels.hip); slot_cfg 1 = narrow, 2 = wide, 3 = small
hipError_t launch_tiles(const TileArgs<double>& a, uint32_t ntiles, int slot_cfg, hipStream_t s);
hipError_t launch_tiles(const TileArgs<float>& a, uint32_t ntiles, int slot_cfg, hipStream_t s);
// many jobs in one launch: block b runs tile map[2b+1] of job map[2b]
hipError_t launch_pool(const TileArgs<double>* jobs, const uint32_t* map, uint32_t nblocks,
                       uint32_t lanes, int slot_cfg, size_t lds, hipStream_t s);
hipError_t launch_pool(const TileArgs<float>* jobs, const uint32_t* map, uint32_t nblocks,
                       uint32_t lanes, int slot_cfg, size_t lds, hipStream_t s);
size_t tile_lds_bytes(size_t elem, uint32_t lanes, bool adjoint, uint32_t W, uint32_t nforced);
// row-parallel path: one workgroup of kBigThreads per instance
constexpr uint32_t kBigThreads = 1024;
// lds: bytes of dynamic LDS for the partial sums (0 when they stay in device memory)
hipError_t launch_big(const BigArgs<double>& a, uint32_t ninst, size_t lds, hipStream_t s);
hipError_t launch_big(const BigArgs<float>& a, uint32_t ninst, size_t lds, hipStream_t s);
// Size of the static device scratch array that holds the run buffers of
// pasif_solve_noalloc, and the address of `dev`'s copy of it (looked up once
// per device).
constexpr size_t kScratchBytes = size_t(64) << 20;
hipError_t scratch_region(int dev, void** ptr, size_t* bytes);
}  // namespace pasif
