// wave_fetch.h -- wave-wide dynamic work fetch for persistent traversal kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#ifndef RISE_REFILL
#define RISE_REFILL 28
#endif

namespace rise_hip {

__device__ inline unsigned long long shfl_u64(unsigned long long v, int src)
{
    const unsigned lo = (unsigned)__shfl((int)(unsigned)v, src);
    const unsigned hi = (unsigned)__shfl((int)(unsigned)(v >> 32), src);
    return ((unsigned long long)hi << 32) | lo;
}

// Lanes whose item is done take new item indices from a global counter, in
// one atomic per wave, once at least RISE_REFILL lanes are idle or no lane is
// busy.  A wave exits when the counter is past the last item and none of its
// lanes holds one.  Returns the new item index for an idle lane, -1 if none
// was taken, -2 once the items are exhausted for this lane.  Items below
// `first` are handed out statically (first_item) and the counter counts from
// there, so a launch starts without a burst of atomics on one address.
__device__ inline long long refill(bool idle, bool busy, unsigned long long* next, int64_t n, int64_t first = 0)
{
    const unsigned long long im = __ballot(idle);
    const unsigned long long bm = __ballot(busy);
    const int nidle = __popcll(im);
    if (nidle == 0 || (nidle < RISE_REFILL && bm != 0ull)) return -1;
    const unsigned rank = __builtin_amdgcn_mbcnt_hi((unsigned)(im >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)im, 0u));
    const int leader = __ffsll((unsigned long long)im) - 1;
    unsigned long long base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(next, (unsigned long long)nidle);
    base = shfl_u64(base, leader);
    if (!idle) return -1;
    const unsigned long long idx = (unsigned long long)first + base + rank;
    return idx < (unsigned long long)n ? (long long)idx : -2;
}

// The statically assigned first item of this thread (one per thread of the
// grid, `first` = grid size in threads), or -2 when there is none.
__device__ inline long long first_item(int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return i < n ? (long long)i : -2;
}

}  // namespace rise_hip
