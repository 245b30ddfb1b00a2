// rungs.hip -- ct2's rung bookkeeping (particle.cpp:20-58, 320-347) on the
// device container:
//   particle_sort_by_rung(minrung)  partitions [rung_begin, n) so that the
//       particles with rung < minrung come first; rung_begin moves to the first
//       with rung >= minrung (minrung 0 only resets rung_begin).  The reference
//       swaps in place (unstable); here the partition is stable and
//       deterministic: per-block counts, a scan, and a scatter of the seven
//       container arrays into scratch, copied back.
//   particle_push_rungs / particle_pop_rungs  a stack of rung_begin values.
//   particle_current_range  [rung_begin, n): the sinks every kick acts on.
//   particle_rung_counts  histogram of all rungs (uint64 atomics).
//   particle_set_minrung  rung = max(rung, minrung) for every particle.
#include <algorithm>

#include "internal.h"

namespace ct2 {
namespace {

constexpr int RB = 1024;  // elements per partition block

struct RungScratch {
    uint32_t *x = nullptr, *y = nullptr, *z = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int8_t* rung = nullptr;
    size_t cap = 0;
    int* blocks = nullptr;  // per-block low counts, then their exclusive prefix
    size_t nblocks_cap = 0;
    unsigned long long* counts = nullptr;  // 256 bins
} g_rs;

__global__ __launch_bounds__(RB) void count_low(const int8_t* __restrict__ rung, long r0, long m, int minrung,
                                                int* __restrict__ blocks) {
    const long e = (long)blockIdx.x * RB + threadIdx.x;
    const int low = e < m && (int)rung[r0 + e] < minrung;
    const int cnt = __syncthreads_count(low);
    if (threadIdx.x == 0) blocks[blockIdx.x] = cnt;
}

// Exclusive scan of nb ints in place (one workgroup), total into blocks[nb].
__global__ __launch_bounds__(1024) void scan_blocks(int* __restrict__ blocks, int nb) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b = min(nb, t * per), e = min(nb, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += blocks[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; ++i) {
        const int v = blocks[i];
        blocks[i] = run;
        run += v;
    }
    if (t == 1023) blocks[nb] = part[t];
}

// Stable scatter: low particles to r0 + (prefix of lows), the others after all lows.
__global__ __launch_bounds__(RB) void scatter_by_rung(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                      const uint32_t* __restrict__ Z, const float* __restrict__ VX,
                                                      const float* __restrict__ VY, const float* __restrict__ VZ,
                                                      const int8_t* __restrict__ R, long r0, long m, int minrung,
                                                      const int* __restrict__ blocks, int nb, RungScratch s) {
    __shared__ int wsum[RB / 64];
    const long e = (long)blockIdx.x * RB + threadIdx.x;
    const bool valid = e < m;
    const bool low = valid && (int)R[r0 + e] < minrung;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long ball = __ballot(low);
    const int before = __popcll(ball & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(ball);
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wave; ++w) woff += wsum[w];
    if (!valid) return;
    const int lrank = woff + before;  // lows before this element within the block
    const long lows_before = blocks[blockIdx.x];
    const long total_low = blocks[nb];
    long dst;
    if (low)
        dst = lows_before + lrank;
    else
        dst = total_low + ((long)blockIdx.x * RB - lows_before) + (threadIdx.x - lrank);
    const long src = r0 + e;
    s.x[dst] = X[src];
    s.y[dst] = Y[src];
    s.z[dst] = Z[src];
    s.vx[dst] = VX[src];
    s.vy[dst] = VY[src];
    s.vz[dst] = VZ[src];
    s.rung[dst] = R[src];
}

__global__ __launch_bounds__(256) void rung_hist(const int8_t* __restrict__ R, long n,
                                                 unsigned long long* __restrict__ counts) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        atomicAdd(&h[(uint8_t)R[i]], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(counts + threadIdx.x, (unsigned long long)h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void zero_u64(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ __launch_bounds__(256) void raise_rungs(int8_t* __restrict__ R, long n, int minrung) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n && (int)R[i] < minrung) R[i] = (int8_t)minrung;
}

void free_scratch() {
    for (void* p : {(void*)g_rs.x, (void*)g_rs.y, (void*)g_rs.z, (void*)g_rs.vx, (void*)g_rs.vy, (void*)g_rs.vz,
                    (void*)g_rs.rung})
        dev_free(p);
    g_rs.x = g_rs.y = g_rs.z = nullptr;
    g_rs.vx = g_rs.vy = g_rs.vz = nullptr;
    g_rs.rung = nullptr;
    g_rs.cap = 0;
}

int ensure_scratch_arrays(size_t m, size_t nb) {
    if (m > g_rs.cap) {
        free_scratch();
        CT2_TRY(dev_alloc((void**)&g_rs.x, m * 4, "rung sort x"));
        CT2_TRY(dev_alloc((void**)&g_rs.y, m * 4, "rung sort y"));
        CT2_TRY(dev_alloc((void**)&g_rs.z, m * 4, "rung sort z"));
        CT2_TRY(dev_alloc((void**)&g_rs.vx, m * 4, "rung sort vx"));
        CT2_TRY(dev_alloc((void**)&g_rs.vy, m * 4, "rung sort vy"));
        CT2_TRY(dev_alloc((void**)&g_rs.vz, m * 4, "rung sort vz"));
        CT2_TRY(dev_alloc((void**)&g_rs.rung, m, "rung sort rung"));
        g_rs.cap = m;
    }
    if (nb + 1 > g_rs.nblocks_cap) {
        dev_free(g_rs.blocks);
        g_rs.blocks = nullptr;
        g_rs.nblocks_cap = 0;
        CT2_TRY(dev_alloc((void**)&g_rs.blocks, (nb + 1) * sizeof(int), "rung sort blocks"));
        g_rs.nblocks_cap = nb + 1;
    }
    return CT2_OK;
}

}  // namespace

void release_rungs() {
    free_scratch();
    dev_free(g_rs.blocks);
    g_rs.blocks = nullptr;
    g_rs.nblocks_cap = 0;
    dev_free(g_rs.counts);
    g_rs.counts = nullptr;
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_particle_sort_by_rung(int minrung) {
    CT2_TRY(require_particles());
    if (minrung < 0 || minrung > 127) return set_error(CT2_ERR_INVALID, "minrung %d out of [0, 127]", minrung);
    Context& c = ctx();
    if (minrung == 0) {
        c.rung_begin = 0;
        return CT2_OK;
    }
    const long r0 = (long)c.rung_begin, m = (long)c.n - r0;
    if (m <= 0) return CT2_OK;
    const int nb = (int)((m + RB - 1) / RB);
    CT2_TRY(ensure_scratch_arrays((size_t)m, (size_t)nb));
    hipStream_t st = c.stream;
    count_low<<<nb, RB, 0, st>>>(c.rung, r0, m, minrung, g_rs.blocks);
    CT2_LAUNCH_CHECK();
    scan_blocks<<<1, 1024, 0, st>>>(g_rs.blocks, nb);
    CT2_LAUNCH_CHECK();
    scatter_by_rung<<<nb, RB, 0, st>>>(c.x, c.y, c.z, c.vx, c.vy, c.vz, c.rung, r0, m, minrung, g_rs.blocks, nb,
                                       g_rs);
    CT2_LAUNCH_CHECK();
    CT2_HIP(hipMemcpyAsync(c.x + r0, g_rs.x, m * 4, hipMemcpyDeviceToDevice, st));
    CT2_HIP(hipMemcpyAsync(c.y + r0, g_rs.y, m * 4, hipMemcpyDeviceToDevice, st));
    CT2_HIP(hipMemcpyAsync(c.z + r0, g_rs.z, m * 4, hipMemcpyDeviceToDevice, st));
    CT2_HIP(hipMemcpyAsync(c.vx + r0, g_rs.vx, m * 4, hipMemcpyDeviceToDevice, st));
    CT2_HIP(hipMemcpyAsync(c.vy + r0, g_rs.vy, m * 4, hipMemcpyDeviceToDevice, st));
    CT2_HIP(hipMemcpyAsync(c.vz + r0, g_rs.vz, m * 4, hipMemcpyDeviceToDevice, st));
    CT2_HIP(hipMemcpyAsync(c.rung + r0, g_rs.rung, m, hipMemcpyDeviceToDevice, st));
    int total_low = 0;
    CT2_HIP(hipMemcpyAsync(&total_low, g_rs.blocks + nb, sizeof(int), hipMemcpyDeviceToHost, st));
    CT2_HIP(hipStreamSynchronize(st));
    c.rung_begin = (size_t)(r0 + total_low);
    c.forces_valid = false;  // the stored forces no longer match the order
    return CT2_OK;
}

extern "C" int ct2_particle_push_rungs(void) {
    CT2_TRY(require_particles());
    Context& c = ctx();
    c.rung_stack.push_back(c.rung_begin);
    return CT2_OK;
}

extern "C" int ct2_particle_pop_rungs(void) {
    CT2_TRY(require_particles());
    Context& c = ctx();
    if (c.rung_stack.empty()) return set_error(CT2_ERR_STATE, "rung stack is empty");
    c.rung_begin = c.rung_stack.back();
    c.rung_stack.pop_back();
    return CT2_OK;
}

extern "C" int ct2_particle_current_range(size_t* first, size_t* last) {
    CT2_TRY(require_particles());
    if (!first || !last) return set_error(CT2_ERR_INVALID, "null output");
    *first = ctx().rung_begin;
    *last = ctx().n;
    return CT2_OK;
}

extern "C" int ct2_particle_rung_counts(uint64_t* counts, int max_rungs, int* nrungs) {
    CT2_TRY(require_particles());
    if (!counts || !nrungs || max_rungs <= 0) return set_error(CT2_ERR_INVALID, "bad rung count arguments");
    Context& c = ctx();
    if (!g_rs.counts) CT2_TRY(dev_alloc((void**)&g_rs.counts, 256 * sizeof(unsigned long long), "rung counts"));
    zero_u64<<<1, 256, 0, c.stream>>>(g_rs.counts, 256);
    CT2_LAUNCH_CHECK();
    const long n = (long)c.n;
    const int grid = (int)std::min<long>(1024, (n + 255) / 256);
    rung_hist<<<grid, 256, 0, c.stream>>>(c.rung, n, g_rs.counts);
    CT2_LAUNCH_CHECK();
    unsigned long long h[256];
    CT2_HIP(hipMemcpyAsync(h, g_rs.counts, sizeof h, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    int top = 0;
    for (int r = 0; r < 256; ++r)
        if (h[r]) top = r + 1;
    if (top > 128) return set_error(CT2_ERR_STATE, "negative rungs in the container");
    if (top > max_rungs) return set_error(CT2_ERR_INVALID, "rungs up to %d need %d counters", top - 1, top);
    for (int r = 0; r < top; ++r) counts[r] = h[r];
    *nrungs = top;
    return CT2_OK;
}

extern "C" int ct2_particle_set_minrung(int minrung) {
    CT2_TRY(require_particles());
    if (minrung < 0 || minrung > 127) return set_error(CT2_ERR_INVALID, "minrung %d out of [0, 127]", minrung);
    Context& c = ctx();
    raise_rungs<<<(unsigned)((c.n + 255) / 256), 256, 0, c.stream>>>(c.rung, (long)c.n, minrung);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}
