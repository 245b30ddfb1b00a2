// wave.h -- wave64 helpers shared by the binning kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ct2 {

// Counting-sort increment with one atomic per run of equal keys: every lane
// whose key is >= 0 adds 1 to counter[key] and receives the counter's value
// before its own increment (its slot in the key's segment).  Consecutive
// lanes with the same key form a run, and the run's first lane adds the run
// length in one atomic.  Inputs in spatial order (lattice or cell order) give
// long runs, so far fewer atomics meet at one counter.  Which slot a lane gets
// depends on the order the runs' atomics land in; the set of slots of a key is
// always 0 .. count-1.  Every lane of the wave must call it (inactive lanes with
// key < 0); the return value of an inactive lane is meaningless.
__device__ __forceinline__ int run_increment(int* counter, int key) {
    const int lane = (int)__lane_id();
    const int prev = __shfl_up(key, 1);
    const bool head = key >= 0 && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head);
    const uint64_t active = __ballot(key >= 0);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    // the run's first lane: the highest head at or below this lane
    const uint64_t below = heads & upto;
    const int start = below ? 63 - __clzll((long long)below) : lane;
    // the run ends at the next head or inactive lane above this one
    const uint64_t stop = (heads | ~active) & ~upto;
    const int end = stop ? __ffsll((long long)stop) - 1 : 64;
    int base = 0;
    if (head) base = atomicAdd(counter + key, end - lane);
    base = __shfl(base, start);
    return base + lane - start;
}

}  // namespace ct2
