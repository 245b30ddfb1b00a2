// near.hip -- the near field of the PM force on gfx950: a direct pair sum over
// cell neighbourhoods beside the mesh far field, the structure of ct2's kick()
// (kick.cpp:105-125: neighbour list at :117, direct pairs through kick_cell
// :36-67, far field from the mesh expansion at :118 and :29).
//
// The Gaussian-filtered PM force (pm.hip) already carries, for every pair, the
// filtered Newtonian kernel phi_long = erf(u)/r, u = r/(2 rs).  For each
// minimum-image pair closer than rcut the near field adds the softened kernel
// minus that:
//     F += GM dx (k_soft(r) - [erf(u) - (2u/sqrt(pi)) e^{-u^2}] / r^3)
//     pot += GM (p_soft(r) - erf(u)/r)
// and the potential drops the background integral GM N 4 pi rs^2, so that
// PM + near field tends to the periodic direct sum.
//
// Pipeline (all on the library stream, deterministic):
//   near_key      slice key of each particle: (x, y) column, then z slice
//                 (NEAR_SUB slices per cell), and its index
//   radix sort    hipCUB's stable LSD sort of (key, index): the order is
//                 (column, slice, particle index) for any scheduling and any
//                 clustering, so the summation order never depends on atomics
//   near_bounds   start of every slice and cell (a binary search each)
//   near_scan     exclusive scan over the rows -> work-item prefix (chunks of
//                 NEAR_BLOCK sinks per row)
//   near_gather   the sorted (x, y, z, index)
//   near_items    the (row, chunk) of every work item
//   near_force    one wave per work item: 64 sinks in registers, the 9 (or
//                 25) neighbouring rows' slice ranges read 64 sources at a
//                 time; sources farther than rcut from the sinks' bounding
//                 box are dropped and the rest gathered in an LDS ring, 64
//                 per tile; the filtered kernel from cubic pieces in LDS,
//                 FP32 pair terms with tile-level TwoSum; adds into the
//                 stored force of each particle.
// Cells have side >= rcut, so every partner of a sink lies in the 3x3x3 cells
// around it; a work item's sinks span a run of slices along z in one row, and
// the union of their neighbourhoods is 9 contiguous (or wrapped) slice ranges
// of the sorted array, each widened by the slices rcut spans.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace ct2 {
namespace {

#ifndef NEAR_BLOCK_SIZE
#define NEAR_BLOCK_SIZE 64
#endif
// sinks per work item == sources per LDS tile: one wave, so that a work item's
// sinks span little of their row and the wave skips distant sources
constexpr int NEAR_BLOCK = NEAR_BLOCK_SIZE;
constexpr int NEAR_MAX_NC = 64;  // cells per side at most (along z; x and y: NEAR_MAX_NC * F)
// Dense cells (more than NEAR_SPLIT_DENSITY particles on average) are split
// F = 2 ways in x and y, so a work item's 64 sinks span a squarer box (a
// quarter of the cross-section, four times the z extent); partners then lie in
// the 5 x 5 neighbouring columns.  Below that density the 3 x 3 stencil of
// whole cells needs fewer candidate sources.
constexpr int NEAR_SPLIT_DENSITY = 200;
constexpr int NEAR_MAX_ROWS = 25;  // (2F + 1)^2
// Each cell is cut into NEAR_SUB slices along z; a cell's segment is ordered
// by (slice, particle index), so a row's particles are in slice order and a
// work item's neighbour ranges start and end at slice boundaries.
#ifndef NEAR_SUB_N
#define NEAR_SUB_N 16
#endif
constexpr int NEAR_SUB = NEAR_SUB_N;
// The filtered Newtonian kernel is evaluated from NEAR_NT cubic pieces in r^2
// over [0, rcut^2): within 2e-8 (relative) of the closed form, below FP32
// rounding (built on the host in FP64 by near_kernel_table).
constexpr int NEAR_NT = 64;
constexpr int SCAN_THREADS = 1024;

struct NearState {
    double rcut_rs = 0.0;  // 0 = off
    int nc = 0;
    size_t cap = 0;                // particles the buffers hold
    uint32_t* key = nullptr;       // [cap] slice key of each particle
    int* idx = nullptr;            // [cap] particle indices (sort values)
    uint4* P0 = nullptr;           // [cap] sorted keys and indices (8 of its 16 B per particle)
    uint4* P = nullptr;            // [cap] sorted (x, y, z, index)
    void* sort_tmp = nullptr;      // radix sort scratch
    size_t sort_tmp_bytes = 0;
    size_t ncell_cap = 0, nrow_cap = 0;  // cells and (x, y) columns the index buffers hold
    int* start = nullptr;          // [nc^3 + 1]
    int* row_chunk = nullptr;      // [nc^2 + 1]
    int* subst = nullptr;          // [nc^3 NEAR_SUB + 1] sorted index of each slice's first particle
    int2* item = nullptr;          // [cap / NEAR_BLOCK + nc^2] (row, chunk) per work item
    size_t item_cap = 0;
    float4* table = nullptr;       // [2 (NEAR_NT + 1)]: k_l pieces, then p_l pieces
    double table_rsu = 0.0, table_rcu = 0.0;  // what the table was built for
} g_near;

struct NearConst {
    float rc2u;      // rcut^2 in 2^-64 units
    float rcu;       // rcut in 2^-32 units
    float eps2u;     // Plummer eps^2
    float h2u, hinvu, hinv3u;  // ct2 kernel
    float tscale;    // NEAR_NT / rcut^2: r^2 -> table coordinate
};

__device__ __forceinline__ uint32_t cell_of(uint32_t x, uint32_t nc) {
    return (uint32_t)(((uint64_t)x * nc) >> 32);
}

// z slice of a particle along its row: u = floor(z nzs / 2^32), nzs = nc NEAR_SUB
// slices per row; u / NEAR_SUB is the cell along z.
__device__ __forceinline__ uint32_t slice_of(uint32_t z, uint32_t nzs) {
    return (uint32_t)(((uint64_t)z * nzs) >> 32);
}

// Sort key of each particle: its (x, y) column row = cx nxy + cy and its z
// slice u along the row, key = row nzs + u -- the index of its slice in
// subst (cells: (row nc + u / NEAR_SUB) = key / NEAR_SUB).
__global__ __launch_bounds__(256) void near_key(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                const uint32_t* __restrict__ Z, int n, uint32_t nxy, uint32_t nzs,
                                                uint32_t* __restrict__ key, int* __restrict__ idx) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    key[p] = (cell_of(X[p], nxy) * nxy + cell_of(Y[p], nxy)) * nzs + slice_of(Z[p], nzs);
    idx[p] = p;
}

// After the stable sort by key: subst[q] = first sorted position with key >=
// q (q = 0 .. nkeys, subst[nkeys] = n), and the cell starts start[c] =
// subst[c NEAR_SUB].  One binary search per slice, so empty slices cost the
// same as full ones.
__global__ __launch_bounds__(256) void near_bounds(const uint32_t* __restrict__ key, int n, int nkeys,
                                                   int* __restrict__ subst, int* __restrict__ start) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] < (uint32_t)q)
            lo = mid + 1;
        else
            hi = mid;
    }
    subst[q] = lo;
    if (q % NEAR_SUB == 0) start[q / NEAR_SUB] = lo;
}

// The sorted particles: P[i] = (x, y, z, index) of sorted entry i.
__global__ __launch_bounds__(256) void near_gather(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                   const uint32_t* __restrict__ Z, const int* __restrict__ idx, int n,
                                                   uint4* __restrict__ P) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p = idx[i];
    P[i] = make_uint4(X[p], Y[p], Z[p], (uint32_t)p);
}

// Exclusive scan (one workgroup) of the work items per row: out[0..len],
// out[len] = total; row i holds ceil(m / NEAR_BLOCK) items, m = in[(i+1)*nc] -
// in[i*nc] (in = cell starts).
__global__ __launch_bounds__(SCAN_THREADS) void near_scan(const int* __restrict__ in, int len, int nc,
                                                          int* __restrict__ out) {
    __shared__ int part[SCAN_THREADS];
    const int t = threadIdx.x;
    const int per = (len + SCAN_THREADS - 1) / SCAN_THREADS;
    const int b = min(len, t * per), e = min(len, b + per);
    auto value = [&](int i) -> int {
        const int m = in[(i + 1) * nc] - in[i * nc];
        return (m + NEAR_BLOCK - 1) / NEAR_BLOCK;
    };
    int s = 0;
    for (int i = b; i < e; ++i) s += value(i);
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;  // exclusive prefix of this thread's chunk
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += value(i);
    }
    if (t == SCAN_THREADS - 1) out[len] = part[t];
}

__device__ __forceinline__ void two_sum(float& s, float& c, float v) {
#pragma clang fp contract(off)
    const float t = s + v;
    const float bp = t - s;
    const float e = (s - (t - bp)) + (v - bp);
    s = t;
    c += e;
}

// Softened kernel minus the filtered Newtonian kernel for one pair, in
// 2^-32 units: returns the force factor k (F += dx k) and potential p.  The
// filtered part (k_l, p_l) comes from the cubic pieces in LDS (tk, tp), so
// no branch depends on the separation beyond the kernel's own.
template <int KERNEL>
__device__ __forceinline__ void near_pair(float r2, const NearConst& q, const float4* tk, const float4* tp, float& k,
                                          float& p) {
    float ks, ps;
    if constexpr (KERNEL == CT2_KERNEL_PLUMMER) {
        const float w = __builtin_amdgcn_rsqf(r2 + q.eps2u);
        ks = (w * w) * w;
        ps = w;
    } else {
        if (r2 < q.h2u) {
            ks = q.hinv3u;
            ps = fmaf(-0.5f * q.hinv3u, r2, 1.5f * q.hinvu);
        } else {
            const float w = __builtin_amdgcn_rsqf(r2);
            ks = (w * w) * w;
            ps = w;
        }
    }
    // t in [0, NEAR_NT] inside the cutoff (NEAR_NT itself only by rounding: the
    // table's extra piece NEAR_NT is the constant end value of the last one)
    const float t = r2 * q.tscale;
    const int i = (int)t;
    const float f = __builtin_amdgcn_fractf(t);
    const float4 a = tk[i], b = tp[i];
    const float kl = fmaf(fmaf(fmaf(a.w, f, a.z), f, a.y), f, a.x);
    const float pl = fmaf(fmaf(fmaf(b.w, f, b.z), f, b.y), f, b.x);
    k = ks - kl;
    p = ps - pl;
}

// Work items, one per NEAR_BLOCK sinks of a row: item[row_chunk[r] + k] = (r, k).
__global__ __launch_bounds__(256) void near_items(const int* __restrict__ row_chunk, int nrow, int2* __restrict__ item) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nrow) return;
    for (int k = row_chunk[r], c = 0; k < row_chunk[r + 1]; ++k, ++c) item[k] = make_int2(r, c);
}

template <int KERNEL>
__global__ __launch_bounds__(NEAR_BLOCK) void near_force(const uint4* __restrict__ P, const int* __restrict__ start,
                                                         const int* __restrict__ subst,
                                                         const int* __restrict__ row_chunk,
                                                         const int2* __restrict__ item, int nxy, int nc, int F,
                                                         int R, NearConst q,
                                                         const float4* __restrict__ table, double fscale,
                                                         double pscale, double pconst, float* __restrict__ fx,
                                                         float* __restrict__ fy, float* __restrict__ fz,
                                                         float* __restrict__ pot) {
    __shared__ float4 lk[NEAR_NT + 1], lp[NEAR_NT + 1];  // the kernel's cubic pieces
    __shared__ int rng[2 * NEAR_MAX_ROWS][2];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    if (b >= row_chunk[nxy * nxy]) return;  // uniform across the workgroup
    const int2 it = item[b];
    const int row = it.x;
    const int s0 = start[row * nc] + it.y * NEAR_BLOCK;
    const int s1 = min(start[(row + 1) * nc], s0 + NEAR_BLOCK);
    const int cx = row / nxy, cy = row % nxy;
    const int side = 2 * F + 1, nr = side * side;  // neighbouring rows: (cx, cy) +- F
    if (tid < nr) {  // one neighbouring row per lane: at most two index ranges (z wraps)
        // the sinks' slices (a row is in slice order) widened by R slices, the
        // most a partner within rcut can differ
        const int Z = nc * NEAR_SUB;
        const int lo = (int)slice_of(P[s0].z, (uint32_t)Z) - R, hi = (int)slice_of(P[s1 - 1].z, (uint32_t)Z) + R;
        const int di = tid / side - F, dj = tid % side - F;
        const int* sr = subst + (((cx + di + nxy) % nxy) * nxy + (cy + dj + nxy) % nxy) * Z;  // slice starts of the row
        int a0, e0, a1 = 0, e1 = 0;
        if (hi - lo + 1 >= Z) {  // the whole row
            a0 = sr[0];
            e0 = sr[Z];
        } else if (lo < 0) {
            a0 = sr[lo + Z];
            e0 = sr[Z];
            a1 = sr[0];
            e1 = sr[hi + 1];
        } else if (hi >= Z) {
            a0 = sr[lo];
            e0 = sr[Z];
            a1 = sr[0];
            e1 = sr[hi - Z + 1];
        } else {
            a0 = sr[lo];
            e0 = sr[hi + 1];
        }
        rng[2 * tid][0] = a0;
        rng[2 * tid][1] = e0;
        rng[2 * tid + 1][0] = a1;
        rng[2 * tid + 1][1] = e1;
    }
    for (int e = tid; e <= NEAR_NT; e += NEAR_BLOCK) {
        lk[e] = table[e];
        lp[e] = table[NEAR_NT + 1 + e];
    }
    __syncthreads();
    const int i = s0 + tid;
    const uint4 me = P[i < s1 ? i : s1 - 1];
    float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f, cx_ = 0.f, cy_ = 0.f, cz_ = 0.f, cp_ = 0.f;
    // The sinks' bounding box: per axis the offsets [L, H] about a centre cc
    // (exact integers).  A source whose distance to the box is >= rcut has no
    // partner among the sinks and is dropped as its tile is loaded.  The
    // distance is squared and summed in the pair's own FP32 operation order
    // on per-axis gaps e <= |dx|, so a source with any r2 < rc2u always
    // passes.  An axis whose half-extent plus rcut nears half the box is not
    // tested (there the source's offset from cc could wrap).
    const uint4 ref = P[s0];
    int lo[3] = {(int32_t)(me.x - ref.x), (int32_t)(me.y - ref.y), (int32_t)(me.z - ref.z)};
    int hi[3] = {lo[0], lo[1], lo[2]};
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], m));
            hi[a] = max(hi[a], __shfl_xor(hi[a], m));
        }
    uint32_t cc[3];
    long long L[3], H[3];
    bool on[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const long long mid = (long long)lo[a] + (((long long)hi[a] - (long long)lo[a]) >> 1);
        cc[a] = (a == 0 ? ref.x : a == 1 ? ref.y : ref.z) + (uint32_t)mid;
        L[a] = (long long)lo[a] - mid;
        H[a] = (long long)hi[a] - mid;
        on[a] = (double)(H[a] > -L[a] ? H[a] : -L[a]) + (double)q.rcu < 0x1p31 - 0x1p24;
    }
    __shared__ uint4 ring[2 * NEAR_BLOCK];  // survivors, consumed NEAR_BLOCK at a time from rb
    int fill = 0, rb = 0;
    float tx = 0.f, ty = 0.f, tz = 0.f, tp = 0.f;
    auto pair = [&](const uint4 s) {
        const float dx = (float)(int32_t)(s.x - me.x);
        const float dy = (float)(int32_t)(s.y - me.y);
        const float dz = (float)(int32_t)(s.z - me.z);
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (r2 < q.rc2u) {  // a source out of range of the whole wave is skipped
            float k, p;
            near_pair<KERNEL>(r2, q, lk, lp, k, p);
            tx = fmaf(dx, k, tx);
            ty = fmaf(dy, k, ty);
            tz = fmaf(dz, k, tz);
            tp += p;
        }
    };
    for (int r = 0; r < 2 * nr; ++r) {
        const int a = rng[r][0], e = rng[r][1];
        for (int j0 = a; j0 < e; j0 += NEAR_BLOCK) {
            __syncthreads();  // the ring's consumed half has been read
            bool keep = false;
            uint4 s = make_uint4(0u, 0u, 0u, 0u);
            if (j0 + tid < e) {
                s = P[j0 + tid];
                const uint32_t sv[3] = {s.x, s.y, s.z};
                float g[3];
#pragma unroll
                for (int a3 = 0; a3 < 3; ++a3) {
                    const long long d = (int32_t)(sv[a3] - cc[a3]);
                    long long gap = L[a3] - d > d - H[a3] ? L[a3] - d : d - H[a3];
                    gap = gap < 0 ? 0 : (gap > 0x7fffffffLL ? 0x7fffffffLL : gap);
                    g[a3] = on[a3] ? (float)(int32_t)gap : 0.f;
                }
                keep = fmaf(g[2], g[2], fmaf(g[1], g[1], g[0] * g[0])) < q.rc2u;
            }
            const uint64_t m = __ballot(keep);
            const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            if (keep) ring[(rb + fill + below) & (2 * NEAR_BLOCK - 1)] = s;
            fill += (int)__popcll(m);
            if (fill >= NEAR_BLOCK) {
                __syncthreads();
                const uint4* T = ring + rb;
#pragma unroll 2
                for (int j = 0; j < NEAR_BLOCK; j += 4) {
                    // four sources read together: the in-range branch inside
                    // pair() would otherwise keep each read next to its use
                    const uint4 u0 = T[j], u1 = T[j + 1], u2 = T[j + 2], u3 = T[j + 3];
                    pair(u0);
                    pair(u1);
                    pair(u2);
                    pair(u3);
                }
                two_sum(sx, cx_, tx);
                two_sum(sy, cy_, ty);
                two_sum(sz, cz_, tz);
                two_sum(sp, cp_, tp);
                tx = ty = tz = tp = 0.f;
                rb ^= NEAR_BLOCK;
                fill -= NEAR_BLOCK;
            }
        }
    }
    __syncthreads();
    for (int j = 0; j < fill; ++j) pair(ring[rb + j]);
    two_sum(sx, cx_, tx);
    two_sum(sy, cy_, ty);
    two_sum(sz, cz_, tz);
    two_sum(sp, cp_, tp);
    if (i < s1) {
        const uint32_t p = me.w;
        fx[p] += (float)(((double)sx + (double)cx_) * fscale);
        fy[p] += (float)(((double)sy + (double)cy_) * fscale);
        fz[p] += (float)(((double)sz + (double)cz_) * fscale);
        pot[p] += (float)(((double)sp + (double)cp_) * pscale - pconst);
    }
}

void free_near_buffers() {
    dev_free(g_near.key);
    dev_free(g_near.idx);
    dev_free(g_near.P0);
    dev_free(g_near.P);
    dev_free(g_near.sort_tmp);
    g_near.sort_tmp = nullptr;
    g_near.sort_tmp_bytes = 0;
    dev_free(g_near.start);
    dev_free(g_near.row_chunk);
    dev_free(g_near.subst);
    g_near.subst = nullptr;
    dev_free(g_near.item);
    dev_free(g_near.table);
    g_near.table = nullptr;
    g_near.table_rsu = g_near.table_rcu = 0.0;
    g_near.key = nullptr;
    g_near.idx = nullptr;
    g_near.P0 = g_near.P = nullptr;
    g_near.start = g_near.row_chunk = nullptr;
    g_near.item = nullptr;
    g_near.item_cap = 0;
    g_near.cap = 0;
    g_near.ncell_cap = g_near.nrow_cap = 0;
}

int ensure_buffers(size_t n, size_t ncell, size_t nrow) {
    if (n > g_near.cap) {
        dev_free(g_near.key);
        dev_free(g_near.idx);
        dev_free(g_near.P0);
        dev_free(g_near.P);
        g_near.key = nullptr;
        g_near.idx = nullptr;
        g_near.P0 = g_near.P = nullptr;
        g_near.cap = 0;
        CT2_TRY(dev_alloc((void**)&g_near.key, n * sizeof(uint32_t), "near-field sort keys"));
        CT2_TRY(dev_alloc((void**)&g_near.idx, n * sizeof(int), "near-field sort values"));
        CT2_TRY(dev_alloc((void**)&g_near.P0, n * sizeof(uint4), "near-field sorted keys"));
        CT2_TRY(dev_alloc((void**)&g_near.P, n * sizeof(uint4), "near-field sorted particles"));
        g_near.cap = n;
    }
    if (ncell > g_near.ncell_cap || nrow > g_near.nrow_cap) {
        dev_free(g_near.start);
        dev_free(g_near.row_chunk);
        dev_free(g_near.subst);
        g_near.start = g_near.row_chunk = g_near.subst = nullptr;
        g_near.ncell_cap = g_near.nrow_cap = 0;
        CT2_TRY(dev_alloc((void**)&g_near.start, (ncell + 1) * sizeof(int), "near-field cell starts"));
        CT2_TRY(dev_alloc((void**)&g_near.row_chunk, (nrow + 1) * sizeof(int), "near-field work items"));
        CT2_TRY(dev_alloc((void**)&g_near.subst, (ncell * NEAR_SUB + 1) * sizeof(int), "near-field slice starts"));
        g_near.ncell_cap = ncell;
        g_near.nrow_cap = nrow;
    }
    const size_t items = n / NEAR_BLOCK + 1 + nrow;
    if (items > g_near.item_cap) {
        dev_free(g_near.item);
        g_near.item = nullptr;
        g_near.item_cap = 0;
        CT2_TRY(dev_alloc((void**)&g_near.item, items * sizeof(int2), "near-field work items"));
        g_near.item_cap = items;
    }
    return CT2_OK;
}

}  // namespace

// Long-range (filtered) part of the pair kernel at t = u^2, u = r / (2 rs),
// without the 1/(2 rs)^3 and 1/(2 rs) factors:
//   K(t) = [erf(u) - (2u/sqrt(pi)) e^{-u^2}] / u^3,   P(t) = erf(u) / u,
// by their Taylor series in t below u = 0.05, where the closed forms cancel.
static double lr_K(double t) {
    const double c = 2.0 / std::sqrt(M_PI);
    if (t < 0.0025) return c * (2.0 / 3.0 + t * (-2.0 / 5.0 + t * (1.0 / 7.0 + t * (-1.0 / 27.0))));
    const double u = std::sqrt(t);
    return (std::erf(u) - c * u * std::exp(-t)) / (t * u);
}
static double lr_P(double t) {
    const double c = 2.0 / std::sqrt(M_PI);
    if (t < 0.0025) return c * (1.0 + t * (-1.0 / 3.0 + t * (1.0 / 10.0 + t * (-1.0 / 42.0))));
    const double u = std::sqrt(t);
    return std::erf(u) / u;
}

// nt cubic pieces of k_l = K(t)/(2 rs)^3 and p_l = P(t)/(2 rs) over r^2 in
// [0, rcut^2) (2^-32 position units); piece i is a polynomial in
// f = r^2 nt / rcut^2 - i in [0, 1), coefficients (c0, c1, c2, c3), computed in
// FP64 by interpolation at the four Chebyshev nodes of the piece.
void near_kernel_table(double rsu, double rcu, int nt, float* k4, float* p4) {
    constexpr int D = 4;
    double node[D], inv[D][D];
    for (int j = 0; j < D; ++j) node[j] = 0.5 - 0.5 * std::cos(M_PI * (j + 0.5) / D);
    // inverse of the Vandermonde matrix V[j][m] = node[j]^m (Gauss-Jordan)
    double a[D][2 * D];
    for (int j = 0; j < D; ++j)
        for (int m = 0; m < 2 * D; ++m) a[j][m] = m < D ? std::pow(node[j], m) : (m - D == j ? 1.0 : 0.0);
    for (int col = 0; col < D; ++col) {
        int piv = col;
        for (int r = col + 1; r < D; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        for (int m = 0; m < 2 * D; ++m) std::swap(a[col][m], a[piv][m]);
        const double d = a[col][col];
        for (int m = 0; m < 2 * D; ++m) a[col][m] /= d;
        for (int r = 0; r < D; ++r)
            if (r != col) {
                const double f = a[r][col];
                for (int m = 0; m < 2 * D; ++m) a[r][m] -= f * a[col][m];
            }
    }
    for (int m = 0; m < D; ++m)
        for (int j = 0; j < D; ++j) inv[m][j] = a[m][D + j];
    const double s = 0.5 / rsu;  // 1 / (2 rs)
    const double tmax = rcu * rcu * s * s;  // u^2 at the cutoff
    for (int i = 0; i < nt; ++i) {
        double vk[D], vp[D];
        for (int j = 0; j < D; ++j) {
            const double t = (i + node[j]) * tmax / nt;
            vk[j] = lr_K(t) * s * s * s;
            vp[j] = lr_P(t) * s;
        }
        for (int m = 0; m < D; ++m) {
            double ck = 0.0, cp = 0.0;
            for (int j = 0; j < D; ++j) {
                ck += inv[m][j] * vk[j];
                cp += inv[m][j] * vp[j];
            }
            k4[4 * i + m] = (float)ck;
            p4[4 * i + m] = (float)cp;
        }
    }
}

bool near_enabled() { return g_near.rcut_rs > 0.0; }
double near_rcut_rs() { return g_near.rcut_rs; }

// Adds the near field to the stored PM force (rs in box units).
int near_forces(double rs) {
    Context& c = ctx();
    const double rcut = g_near.rcut_rs * rs;
    int nc = (int)std::floor(1.0 / (rcut * (1.0 + 1e-6)));
    nc = std::min(nc, NEAR_MAX_NC);
    if (nc < 3) return set_error(CT2_ERR_INVALID, "near-field cutoff %g exceeds a third of the box", rcut);
    const int n = (int)c.n;
    const int F = (double)n / ((double)nc * nc * nc) > NEAR_SPLIT_DENSITY ? 2 : 1;
    const int nxy = F * nc;
    const int c3 = nxy * nxy * nc, nrow = nxy * nxy;
    CT2_TRY(ensure_buffers(c.n, (size_t)c3, (size_t)nrow));
    g_near.nc = nc;
    const unsigned pb = (unsigned)((n + 255) / 256);
    hipStream_t st = c.stream;
    // (slice key, particle index) pairs sorted stably by key: the order is
    // (row, z slice, particle index) whatever the scheduling
    const uint32_t nzs = (uint32_t)(nc * NEAR_SUB);
    const int nkeys = c3 * NEAR_SUB;  // = nrow nzs < 2^25
    int bits = 1;
    while ((1 << bits) < nkeys) ++bits;
    uint32_t* kin = g_near.key;
    int* vin = g_near.idx;
    uint32_t* kout = (uint32_t*)g_near.P0;  // P0 (16 B per particle) holds the sorted keys and indices
    int* vout = (int*)g_near.P0 + n;
    near_key<<<pb, 256, 0, st>>>(c.x, c.y, c.z, n, (uint32_t)nxy, nzs, kin, vin);
    CT2_LAUNCH_CHECK();
    size_t tmp = 0;
    CT2_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, n, 0, bits, st));
    CT2_TRY(ensure_scratch(&g_near.sort_tmp, &g_near.sort_tmp_bytes, std::max<size_t>(tmp, 256), "near-field sort"));
    CT2_HIP(hipcub::DeviceRadixSort::SortPairs(g_near.sort_tmp, tmp, kin, kout, vin, vout, n, 0, bits, st));
    near_bounds<<<(unsigned)((nkeys + 256) / 256), 256, 0, st>>>(kout, n, nkeys, g_near.subst, g_near.start);
    CT2_LAUNCH_CHECK();
    near_scan<<<1, SCAN_THREADS, 0, st>>>(g_near.start, nrow, nc, g_near.row_chunk);
    CT2_LAUNCH_CHECK();
    near_gather<<<pb, 256, 0, st>>>(c.x, c.y, c.z, vout, n, g_near.P);
    CT2_LAUNCH_CHECK();

    const double u = 0x1p32;
    const double rsu = rs * u, rcu = rcut * u, h = c.opts.hsoft, hu = h * u;
    NearConst q;
    q.rc2u = (float)(rcu * rcu);
    q.rcu = (float)rcu;
    q.eps2u = (float)(hu * hu);
    q.h2u = (float)(hu * hu);
    q.hinvu = (float)(1.0 / hu);
    q.hinv3u = (float)(1.0 / (hu * hu * hu));
    q.tscale = (float)(NEAR_NT / (rcu * rcu));
    if (!g_near.table)
        CT2_TRY(dev_alloc((void**)&g_near.table, 2 * (NEAR_NT + 1) * sizeof(float4), "near-field table"));
    if (g_near.table_rsu != rsu || g_near.table_rcu != rcu) {
        std::vector<float> t(8 * (NEAR_NT + 1), 0.0f);
        float* tk = t.data();
        float* tp = t.data() + 4 * (NEAR_NT + 1);
        near_kernel_table(rsu, rcu, NEAR_NT, tk, tp);
        for (float* c : {tk, tp}) {  // piece NEAR_NT: the last piece's value at f = 1
            const float* last = c + 4 * (NEAR_NT - 1);
            c[4 * NEAR_NT] = ((last[0] + last[1]) + last[2]) + last[3];
        }
        CT2_HIP(hipMemcpyAsync(g_near.table, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, st));
        CT2_HIP(hipStreamSynchronize(st));  // t is released on return
        g_near.table_rsu = rsu;
        g_near.table_rcu = rcu;
    }
    const double gm = c.opts.GM;
    const double fscale = gm * 0x1p64, pscale = gm * 0x1p32;
    const double pconst = gm * (double)n * 4.0 * M_PI * rs * rs;
    near_items<<<(nrow + 255) / 256, 256, 0, st>>>(g_near.row_chunk, nrow, g_near.item);
    CT2_LAUNCH_CHECK();
    // slices a partner within rcut can be away along z (rcut <= 1 / nc: R <= NEAR_SUB)
    const int R = std::min(NEAR_SUB, (int)std::ceil(rcut * nc * NEAR_SUB));
    // work items: at most one partial chunk per row beyond n / NEAR_BLOCK full ones
    const unsigned items = (unsigned)((n + NEAR_BLOCK - 1) / NEAR_BLOCK + nrow);
    if (c.opts.kernel == CT2_KERNEL_PLUMMER)
        near_force<CT2_KERNEL_PLUMMER><<<items, NEAR_BLOCK, 0, st>>>(g_near.P, g_near.start, g_near.subst,
                                                                     g_near.row_chunk, g_near.item, nxy, nc, F, R,
                                                                     q, g_near.table, fscale, pscale, pconst, c.fx,
                                                                     c.fy, c.fz, c.pot);
    else
        near_force<CT2_KERNEL_CT2><<<items, NEAR_BLOCK, 0, st>>>(g_near.P, g_near.start, g_near.subst,
                                                                 g_near.row_chunk, g_near.item, nxy, nc, F, R, q,
                                                                 g_near.table, fscale, pscale, pconst, c.fx, c.fy,
                                                                 c.fz, c.pot);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

int near_set(double rcut_rs) {
    if (!(rcut_rs >= 0.0) || rcut_rs > 64.0) return set_error(CT2_ERR_INVALID, "bad near-field cutoff %g", rcut_rs);
    g_near.rcut_rs = rcut_rs;
    return CT2_OK;
}

int near_cells() { return g_near.nc; }

void release_near() {
    free_near_buffers();
    g_near.rcut_rs = 0.0;
    g_near.nc = 0;
}

}  // namespace ct2

extern "C" int ct2_near_kernel_table(double rs_u, double rcut_u, int nt, float* k4, float* p4) {
    if (!(rs_u > 0.0) || !(rcut_u > 0.0) || nt < 1 || !k4 || !p4)
        return ct2::set_error(CT2_ERR_INVALID, "ct2_near_kernel_table: bad argument");
    ct2::near_kernel_table(rs_u, rcut_u, nt, k4, p4);
    return CT2_OK;
}
