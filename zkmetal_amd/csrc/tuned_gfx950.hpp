// Window widths by size class ceil(log2 n) on MI355X (gfx950), as measured by
// the autotuner in zkmetal_amd/tune.py:
//     python -m zkmetal_amd.tune --min-logn 8 --max-logn 22 --reps 3
// kTunedPlain: an MSM called alone (lowest wall time per call); kTunedBatch:
// an MSM inside a multi-MSM call (lowest time per MSM in a call holding 2^21
// points of that class).  0 = the cost model (default_window_bits).  Engines
// start from these tables; zkmsm_set_window_table / _batch_ replace them.
//
// kTunedParts / kTunedPer: the launch schedule by call size class ceil(log2
// entries) -- pipeline parts and buckets per bucket-reduction lane -- for
// calls of one MSM ([0]) and of several ([1]), from the same sweep; 0 = the
// built-in rule (default_parts, and 8 to 32 buckets per lane by bucket
// count).  zkmsm_set_schedule_table replaces them.
#pragma once
#include <stdint.h>

namespace zk {

static const uint32_t kTunedPlain[33] = {0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 8, 8, 10, 10, 10, 12, 15,
                                         15, 16, 16, 16, 16, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static const uint32_t kTunedBatch[33] = {0, 0, 0, 0, 0, 0, 0, 0, 6, 7, 8, 8, 9, 10, 11, 12, 13,
                                         13, 16, 16, 16, 16, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// The sweep kept the built-in parts everywhere and the built-in reduction
// tile for batched calls; single MSMs of entry classes 2^13 and 2^17 .. 2^21
// (2^12 .. 2^17 points) ran 4-7% faster with 4 buckets per reduction lane
// (their reduction is latency-bound: shorter chains before the tree).
static const uint8_t kTunedParts[2][33] = {};
static const uint8_t kTunedPer[2][33] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 4, 0, 0, 0, 4, 4, 4, 4, 4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {}};

}  // namespace zk
