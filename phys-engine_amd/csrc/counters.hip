// Folding of per-workgroup partials into the step counters (see PartialSlot):
// one workgroup per entry reduces the slot's partials and writes the
// destination counter; no device atomics.
#include <stddef.h>

#include "fold.hpp"
#include "internal.hpp"

namespace pg {
namespace {

__global__ __launch_bounds__(kFoldThreads) void k_fold(const uint64_t* __restrict__ partials,
                                                       FoldList list, Counters* counters)
{
    if (list.finish_pairs && blockIdx.x == 0 && threadIdx.x == 0) {
        // launch_finish_pairs' work, after the broad phase's last pair.
        const unsigned long long p = counters->pair_count;
        counters->sort_count = uint32_t(p < list.finish_cap ? p : list.finish_cap);
        counters->mid_segments = 0;
        counters->long_segments = 0;
    }
    if (blockIdx.x < list.count)
        fold_entry(partials, list.e[blockIdx.x], counters);
}

}  // namespace

void launch_fold(physgpu_context* ctx, const FoldList& in)
{
    FoldList list = in;
    if (ctx->finish_pending) {
        list.finish_pairs = 1;
        list.finish_cap = ctx->cap_p;
        ctx->finish_pending = false;
    }
    if (list.count || list.finish_pairs)
        hipLaunchKernelGGL(k_fold, dim3(list.count ? list.count : 1), dim3(kFoldThreads), 0, ctx->stream,
                           ctx->partials, list, ctx->counters);
}

}  // namespace pg
