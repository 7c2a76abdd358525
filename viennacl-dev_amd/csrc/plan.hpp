// plan.hpp -- CSR row-block plan shared by the host planner and the HIP kernels.
//
// A block is a run of consecutive rows processed by one 256-lane workgroup.  Blocks whose
// rows hold at most kCsrNnzCap entries in total are staged through LDS: the workgroup reads
// the block's values and column indices with contiguous, coalesced loads, stores the products
// a_ij * x_j in LDS, and then 2^mode lanes reduce each row.  A row longer than kCsrNnzCap is a
// block of its own (mode kModeLongRow) that the workgroup reduces straight from global memory.
// The per-block lane count follows the longest row of the block so that short rows get one
// lane and long rows get up to a full wavefront.
#pragma once

#include <cstdint>

namespace vclhip {

constexpr uint32_t kPlanRowLimit = 1u << 28;  // first row is packed into 28 bits
constexpr uint32_t kCsrWG = 256;              // lanes per workgroup
constexpr uint32_t kCsrNnzCap = 2048;         // LDS staging capacity (entries) per block
constexpr uint32_t kLaneTarget = 8;           // products each lane sums in the row phase
constexpr uint32_t kModeLongRow = 7;

inline uint32_t plan_row(uint32_t packed) { return packed & (kPlanRowLimit - 1); }
inline uint32_t plan_mode(uint32_t packed) { return packed >> 28; }

// log2 of the lanes per row for a block whose longest row has `maxlen` entries.
inline uint32_t plan_lanes_log2(uint32_t maxlen)
{
  uint32_t want = (maxlen + kLaneTarget - 1) / kLaneTarget;  // lanes to keep ~kLaneTarget products per lane
  uint32_t l2 = 0;
  while ((1u << l2) < want && l2 < 6) ++l2;
  return l2;
}

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
    uint32_t maxlen = len, nnz = len, l2 = plan_lanes_log2(len);
    ++r;
    while (r < n) {
      len = rp[r + 1] - rp[r];
      if (len > kCsrNnzCap) break;
      uint32_t m = maxlen > len ? maxlen : len;
      uint32_t ml2 = plan_lanes_log2(m);
      uint32_t rows = r - start + 1;
      if (nnz + len > kCsrNnzCap || (rows << ml2) > kCsrWG) break;
      maxlen = m;
      nnz += len;
      l2 = ml2;
      ++r;
    }
    blocks[nb++] = start | (l2 << 28);
  }
  blocks[nb] = n;
  return nb;
}

}  // namespace vclhip
