// Host-side structures shared by the tile packer and the lane-kernel
// planner: per-row term lists of one side of the problem and the
// (system, excitation) instances a solve runs.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace pasif {

// Per-row term lists of one side (forward or adjoint) in global indices:
// kind k terms of row r are [ptr[k][r], ptr[k][r+1]) of val[k] / idx[k]
// (idx[k] holds k + 1 index words per term: S, X1, X2, X3), in the order the
// reference applies them (CSR order, then COO order).
struct RowTerms {
  uint32_t n = 0;      // rows
  uint32_t nset = 0;   // size of the set-point index space
  std::array<std::vector<uint32_t>, 4> ptr;
  std::array<std::vector<double>, 4> val;
  std::array<std::vector<uint32_t>, 4> idx;
  std::vector<double> force;  // dense, eps-dropped at pack time
  std::vector<double> init;
  std::vector<uint32_t> off, size;   // system partition (rows)
  std::vector<uint32_t> xoff;        // adjoint: forward offset of each system
};

// one (system, excitation) pair
struct Instance {
  uint32_t sys;
  int32_t exc;         // -1: unforced
  uint32_t steps;
  uint32_t out_f;      // output base of forward rows
  uint32_t out_a;      // output base of adjoint rows
  uint32_t lanes;      // lanes the instance occupies in a tile
};

}  // namespace pasif
