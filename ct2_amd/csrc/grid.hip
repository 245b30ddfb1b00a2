// grid.hip -- ct2's cell binning of the container (particle_to_grid,
// particle.cpp:59-114, 349-362) and its cell accessors (particle.cpp:290-300).
//
// ct2 partitions the active range [rung_begin, n) recursively in place into
// Ncell^3 cells and keeps, per cell, its range of the (now cell-ordered)
// arrays.  Here the range is sorted stably by cell index
// ((ix Ncell) + iy) Ncell + iz, ix = floor(x Ncell / 2^32): a radix sort of
// (cell, position-in-range) pairs, then a gather of the seven container
// arrays into scratch and a copy back.  Inside a cell the particles keep
// their previous relative order, so the result does not depend on scheduling.
// Cell starts come from a binary search of the sorted keys per cell.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace ct2 {
namespace {

struct GridState {
    uint32_t* key = nullptr;   // [m] cell of each particle of the range
    uint32_t* skey = nullptr;  // [m] sorted
    int* idx = nullptr;        // [m] position in the range
    int* sidx = nullptr;       // [m] sorted
    size_t cap = 0;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    // gather scratch (the seven container arrays of the range)
    uint32_t *x = nullptr, *y = nullptr, *z = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int8_t* rung = nullptr;
    unsigned long long* starts = nullptr;  // [ncell^3 + 1], absolute container indices
    size_t starts_cap = 0;
    int ncell = 0;
    size_t r0 = 0, r1 = 0;  // the range the starts describe
} g_grid;

__global__ __launch_bounds__(256) void cell_key(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                const uint32_t* __restrict__ Z, long r0, long m, uint32_t nc,
                                                uint32_t* __restrict__ key, int* __restrict__ idx) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= m) return;
    const uint32_t ix = (uint32_t)(((uint64_t)X[r0 + e] * nc) >> 32);
    const uint32_t iy = (uint32_t)(((uint64_t)Y[r0 + e] * nc) >> 32);
    const uint32_t iz = (uint32_t)(((uint64_t)Z[r0 + e] * nc) >> 32);
    key[e] = (ix * nc + iy) * nc + iz;
    idx[e] = (int)e;
}

__global__ __launch_bounds__(256) void grid_gather(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                   const uint32_t* __restrict__ Z, const float* __restrict__ VX,
                                                   const float* __restrict__ VY, const float* __restrict__ VZ,
                                                   const int8_t* __restrict__ R, long r0, long m,
                                                   const int* __restrict__ sidx, GridState s) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= m) return;
    const long src = r0 + sidx[e];
    s.x[e] = X[src];
    s.y[e] = Y[src];
    s.z[e] = Z[src];
    s.vx[e] = VX[src];
    s.vy[e] = VY[src];
    s.vz[e] = VZ[src];
    s.rung[e] = R[src];
}

// starts[c] = r0 + (number of sorted keys < c), c in [0, ncells]
__global__ __launch_bounds__(256) void cell_starts(const uint32_t* __restrict__ skey, long m, long ncells, long r0,
                                                   unsigned long long* __restrict__ starts) {
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    if (c > ncells) return;
    long lo = 0, hi = m;
    while (lo < hi) {
        const long mid = (lo + hi) >> 1;
        if ((long)skey[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    starts[c] = (unsigned long long)(r0 + lo);
}

void free_arrays() {
    for (void* p : {(void*)g_grid.key, (void*)g_grid.skey, (void*)g_grid.idx, (void*)g_grid.sidx, (void*)g_grid.x,
                    (void*)g_grid.y, (void*)g_grid.z, (void*)g_grid.vx, (void*)g_grid.vy, (void*)g_grid.vz,
                    (void*)g_grid.rung})
        dev_free(p);
    g_grid.key = g_grid.skey = nullptr;
    g_grid.idx = g_grid.sidx = nullptr;
    g_grid.x = g_grid.y = g_grid.z = nullptr;
    g_grid.vx = g_grid.vy = g_grid.vz = nullptr;
    g_grid.rung = nullptr;
    g_grid.cap = 0;
}

int ensure_arrays(size_t m) {
    if (m <= g_grid.cap) return CT2_OK;
    free_arrays();
    CT2_TRY(dev_alloc((void**)&g_grid.key, m * 4, "grid keys"));
    CT2_TRY(dev_alloc((void**)&g_grid.skey, m * 4, "grid sorted keys"));
    CT2_TRY(dev_alloc((void**)&g_grid.idx, m * 4, "grid indices"));
    CT2_TRY(dev_alloc((void**)&g_grid.sidx, m * 4, "grid sorted indices"));
    CT2_TRY(dev_alloc((void**)&g_grid.x, m * 4, "grid x"));
    CT2_TRY(dev_alloc((void**)&g_grid.y, m * 4, "grid y"));
    CT2_TRY(dev_alloc((void**)&g_grid.z, m * 4, "grid z"));
    CT2_TRY(dev_alloc((void**)&g_grid.vx, m * 4, "grid vx"));
    CT2_TRY(dev_alloc((void**)&g_grid.vy, m * 4, "grid vy"));
    CT2_TRY(dev_alloc((void**)&g_grid.vz, m * 4, "grid vz"));
    CT2_TRY(dev_alloc((void**)&g_grid.rung, m, "grid rung"));
    g_grid.cap = m;
    return CT2_OK;
}

}  // namespace

void release_grid() {
    free_arrays();
    dev_free(g_grid.tmp);
    g_grid.tmp = nullptr;
    g_grid.tmp_bytes = 0;
    dev_free(g_grid.starts);
    g_grid.starts = nullptr;
    g_grid.starts_cap = 0;
    g_grid.ncell = 0;
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_particles_to_grid(int ncell, uint64_t* starts) {
    CT2_TRY(require_particles());
    if (ncell < 1 || ncell > CT2_GRID_MAX_NCELL)
        return set_error(CT2_ERR_INVALID, "ncell %d out of [1, %d]", ncell, CT2_GRID_MAX_NCELL);
    Context& c = ctx();
    const long r0 = (long)c.rung_begin, m = (long)c.n - r0;
    const long ncells = (long)ncell * ncell * ncell;
    hipStream_t st = c.stream;
    if (g_grid.starts_cap < (size_t)ncells + 1) {
        dev_free(g_grid.starts);
        g_grid.starts = nullptr;
        g_grid.starts_cap = 0;
        CT2_TRY(dev_alloc((void**)&g_grid.starts, ((size_t)ncells + 1) * 8, "cell starts"));
        g_grid.starts_cap = (size_t)ncells + 1;
    }
    if (m > 0) {
        CT2_TRY(ensure_arrays((size_t)m));
        const unsigned pb = (unsigned)((m + 255) / 256);
        cell_key<<<pb, 256, 0, st>>>(c.x, c.y, c.z, r0, m, (uint32_t)ncell, g_grid.key, g_grid.idx);
        CT2_LAUNCH_CHECK();
        int bits = 1;
        while ((1L << bits) < ncells) ++bits;
        size_t tmp = 0;
        CT2_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, g_grid.key, g_grid.skey, g_grid.idx, g_grid.sidx,
                                                   (int)m, 0, bits, st));
        CT2_TRY(ensure_scratch(&g_grid.tmp, &g_grid.tmp_bytes, std::max<size_t>(tmp, 256), "grid sort"));
        CT2_HIP(hipcub::DeviceRadixSort::SortPairs(g_grid.tmp, tmp, g_grid.key, g_grid.skey, g_grid.idx, g_grid.sidx,
                                                   (int)m, 0, bits, st));
        grid_gather<<<pb, 256, 0, st>>>(c.x, c.y, c.z, c.vx, c.vy, c.vz, c.rung, r0, m, g_grid.sidx, g_grid);
        CT2_LAUNCH_CHECK();
        CT2_HIP(hipMemcpyAsync(c.x + r0, g_grid.x, m * 4, hipMemcpyDeviceToDevice, st));
        CT2_HIP(hipMemcpyAsync(c.y + r0, g_grid.y, m * 4, hipMemcpyDeviceToDevice, st));
        CT2_HIP(hipMemcpyAsync(c.z + r0, g_grid.z, m * 4, hipMemcpyDeviceToDevice, st));
        CT2_HIP(hipMemcpyAsync(c.vx + r0, g_grid.vx, m * 4, hipMemcpyDeviceToDevice, st));
        CT2_HIP(hipMemcpyAsync(c.vy + r0, g_grid.vy, m * 4, hipMemcpyDeviceToDevice, st));
        CT2_HIP(hipMemcpyAsync(c.vz + r0, g_grid.vz, m * 4, hipMemcpyDeviceToDevice, st));
        CT2_HIP(hipMemcpyAsync(c.rung + r0, g_grid.rung, m, hipMemcpyDeviceToDevice, st));
    }
    cell_starts<<<(unsigned)((ncells + 1 + 255) / 256), 256, 0, st>>>(g_grid.skey, m > 0 ? m : 0, ncells, r0,
                                                                        g_grid.starts);
    CT2_LAUNCH_CHECK();
    if (starts)
        CT2_HIP(hipMemcpyAsync(starts, g_grid.starts, ((size_t)ncells + 1) * 8, hipMemcpyDeviceToHost, st));
    CT2_HIP(hipStreamSynchronize(st));
    g_grid.ncell = ncell;
    g_grid.r0 = (size_t)r0;
    g_grid.r1 = c.n;
    c.forces_valid = false;  // the stored forces no longer match the order
    return CT2_OK;
}
