// Window widths by size class ceil(log2 n) on MI355X (gfx950), as measured by
// the autotuner in zkmetal_amd/tune.py:
//     python -m zkmetal_amd.tune --min-logn 8 --max-logn 22 --reps 3
// kTunedPlain: an MSM called alone (lowest wall time per call); kTunedBatch:
// an MSM inside a multi-MSM call (lowest time per MSM in a call holding 2^21
// points of that class).  0 = the cost model (default_window_bits).  Engines
// start from these tables; zkmsm_set_window_table / _batch_ replace them.
#pragma once
#include <stdint.h>

namespace zk {

static const uint32_t kTunedPlain[33] = {0, 0, 0, 0, 0, 0, 0, 0, 8, 8, 8, 8, 10, 10, 10, 12, 15,
                                         15, 16, 16, 16, 16, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
static const uint32_t kTunedBatch[33] = {0, 0, 0, 0, 0, 0, 0, 0, 6, 7, 8, 8, 9, 10, 11, 12, 13,
                                         13, 16, 16, 16, 16, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

}  // namespace zk
