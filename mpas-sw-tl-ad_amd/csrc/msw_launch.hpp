// Block-to-item mapping of the stage launches (grid sizes and the logical
// item each hardware thread handles).  The kernels and the launchers in
// msw_kernels.hip use these functions; a host-side check of the mapping
// (tests/launch_map_check.cpp) compiles the same header with MSW_HD defined
// empty, so what it checks is the code the GPU runs.
//
// Blocks are dealt round-robin to the 8 XCDs (hardware block i runs on XCD
// i % 8).  Logical blocks are grouped in stripes of 2^sl consecutive blocks
// and the stripes are dealt to the XCDs in turn, so each XCD's L2 sees runs
// of consecutive entities along the space-filling curve.  sl == kRangePerXcd
// instead gives each XCD one contiguous range of the segment's blocks.
// Grids are rounded up to a multiple of 8 * 2^sl blocks (8 for
// kRangePerXcd); threads whose logical item lies past the last item return
// at once.  All index arithmetic is 32-bit unsigned (contexts keep every
// launch below 2^31 items).
#pragma once

#ifndef MSW_HD
#define MSW_HD __host__ __device__
#endif

namespace msw {
namespace launch {

constexpr int kRangePerXcd = 7;

// logical block of hardware block i in a segment of nseg blocks
MSW_HD inline unsigned stripe_block(unsigned i, int sl, unsigned nseg)
{
    const unsigned x = i & 7u, j = i >> 3;
    if (sl == kRangePerXcd) return x * (nseg >> 3) + j;
    return ((j >> sl) << (3 + sl)) + (x << sl) + (j & ((1u << sl) - 1u));
}

// nb blocks rounded up to the stripe granule
MSW_HD inline unsigned long round_blocks(unsigned long nb, int sl)
{
    const unsigned long q = sl == kRangePerXcd ? 8ul : 8ul << sl;
    return ((nb + q - 1) / q) * q;
}

// blocks for `items` threads (at least one granule)
MSW_HD inline int grid_items(long items, int sl, int block)
{
    const long nb = (items + block - 1) / block;
    return static_cast<int>(round_blocks(static_cast<unsigned long>(nb > 0 ? nb : 1), sl));
}

// blocks of a launch whose first segment covers nE << lb edge items and whose
// second covers nC << lb cell items
MSW_HD inline int grid_edge_cell(int nE, int nC, int lb, int sl, int block)
{
    const long be = ((static_cast<long>(nE) << lb) + block - 1) / block;
    const long bc = ((static_cast<long>(nC) << lb) + block - 1) / block;
    return static_cast<int>(round_blocks(static_cast<unsigned long>(be), sl) +
                            round_blocks(static_cast<unsigned long>(bc), sl));
}

// Item of a thread in a launch of grid_items(total, sl, block) blocks:
// live == false for threads past the last item.
struct Item {
    unsigned t;
    bool live;
    bool is_edge;   // edge/cell launches: the thread is in the edge segment
};

MSW_HD inline Item flat_item(unsigned bid, unsigned tid, unsigned grid, int block, int sl, unsigned total)
{
    const unsigned t = stripe_block(bid, sl, grid) * static_cast<unsigned>(block) + tid;
    return {t, t < total, false};
}

// Edge/cell launch of grid_edge_cell(nE, nC, lb, sl, block) blocks: the first
// round_blocks(ceil(nE << lb / block)) blocks take the edge items, the rest
// the cell items; each segment is striped on its own.
MSW_HD inline Item edge_cell_item(unsigned bid, unsigned tid, unsigned grid, int block, int sl, int nE,
                                  int nC, int lb)
{
    const unsigned tE = static_cast<unsigned>(nE) << lb;
    const unsigned tC = static_cast<unsigned>(nC) << lb;
    const unsigned bs = static_cast<unsigned>(block);
    const unsigned gE = static_cast<unsigned>(round_blocks((tE + bs - 1) / bs, sl));
    const bool is_edge = bid < gE;
    const unsigned t = stripe_block(is_edge ? bid : bid - gE, sl, is_edge ? gE : grid - gE) * bs + tid;
    return {t, t < (is_edge ? tE : tC), is_edge};
}

}  // namespace launch
}  // namespace msw
