// plan.hpp -- CSR row-block plan shared by the host planner and the HIP kernels.
//
// A block is a run of at most kCsrWG consecutive rows holding at most kCsrNnzCap entries, processed
// by one 256-lane workgroup: the workgroup reads the block's values and column indices with
// contiguous, coalesced loads, stores the products a_ij * x_j in LDS, and then reduces the rows
// from LDS -- a row of at most kShortRow entries by one lane in order (the host's summation order),
// a longer row by a whole wavefront.  Mode kModeMixed marks blocks that hold such longer rows, so
// that blocks of short rows skip that pass.  A row longer than kCsrNnzCap is a block of its own
// (mode kModeLongRow) that the workgroup reduces straight from global memory.
#pragma once

#include <cstdint>

namespace vclhip {

constexpr uint32_t kPlanRowLimit = 1u << 28;  // first row is packed into 28 bits
constexpr uint32_t kCsrWG = 256;              // lanes per workgroup
constexpr uint32_t kCsrNnzCap = 2048;         // LDS staging capacity (entries) per block
constexpr uint32_t kShortRow = 32;            // rows up to this length are summed by one lane
constexpr uint32_t kModeShort = 0;
constexpr uint32_t kModeMixed = 1;
constexpr uint32_t kModeLongRow = 7;

inline uint32_t plan_row(uint32_t packed) { return packed & (kPlanRowLimit - 1); }
inline uint32_t plan_mode(uint32_t packed) { return packed >> 28; }

// Greedy partition; `blocks` holds n_rows + 1 entries, the entry after the last block is n_rows.
inline uint32_t plan_csr_blocks(const uint32_t* rp, uint32_t n, uint32_t* blocks)
{
  uint32_t nb = 0;
  uint32_t r = 0;
  while (r < n) {
    uint32_t len = rp[r + 1] - rp[r];
    if (len > kCsrNnzCap) {
      blocks[nb++] = r | (kModeLongRow << 28);
      ++r;
      continue;
    }
    const uint32_t start = r;
    uint32_t nnz = len;
    bool mixed = len > kShortRow;
    ++r;
    while (r < n && r - start < kCsrWG) {
      len = rp[r + 1] - rp[r];
      if (len > kCsrNnzCap || nnz + len > kCsrNnzCap) break;
      nnz += len;
      mixed = mixed || len > kShortRow;
      ++r;
    }
    blocks[nb++] = start | ((mixed ? kModeMixed : kModeShort) << 28);
  }
  blocks[nb] = n;
  return nb;
}

}  // namespace vclhip
