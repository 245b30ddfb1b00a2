// kick.hip -- the direct-sum kick of kick.cpp:3-125 on gfx950.
//
// kick_direct: one sink per lane; sources stream through LDS in tiles of
// KICK_TILE (coalesced SoA loads, double-buffered, one barrier per tile) and
// are read back as 16-byte LDS broadcasts.  Separations are formed in the
// fixed32 integer domain (exact minimum image) and converted to FP32 in
// units of 2^-32, so no per-pair scale multiply is needed.  Each tile is
// summed in plain FP32 and folded into the running total with TwoSum
// (compensated accumulation); CT2_KICK_NAIVE sums straight into the total.
// Partial sums of a source chunk go to HBM as FP64 so that the chunked
// (split) launch used for small N stays deterministic.
//
// kick_finalize: adds the chunks in order, scales by GM, stores the force
// and runs the half-kick/rung epilogue of kick.cpp:75-99.
//
// kick_sample_fp64: the FP64 judge; one workgroup per sampled sink, sources
// strided over the lanes, reduced with wavefront shuffles then LDS.
#include <algorithm>
#include <cmath>

#include <hipcub/hipcub.hpp>

#include "epilogue.h"
#include "internal.h"

namespace ct2 {
namespace {

constexpr int KICK_BLOCK = 256;  // sinks per workgroup
constexpr int KICK_TILE = 256;   // sources per LDS tile (== KICK_BLOCK: one load per lane)

struct KickConst {
    float eps2u;   // Plummer: (h 2^32)^2
    float h2u;     // ct2 kernel: (h 2^32)^2
    float hinvu;   //             1/(h 2^32)
    float hinv3u;  //             1/(h 2^32)^3
};

__device__ __forceinline__ void two_sum(float& s, float& c, float v) {
#pragma clang fp contract(off)
    const float t = s + v;
    const float bp = t - s;
    const float e = (s - (t - bp)) + (v - bp);
    s = t;
    c += e;
}

template <int KERNEL>
__device__ __forceinline__ void interact(const uint4 src, uint32_t xi, uint32_t yi, uint32_t zi,
                                         const KickConst& kc, float& tx, float& ty, float& tz,
                                         float& tp) {
    const float dx = (float)(int32_t)(src.x - xi);
    const float dy = (float)(int32_t)(src.y - yi);
    const float dz = (float)(int32_t)(src.z - zi);
    if constexpr (KERNEL == CT2_KERNEL_PLUMMER) {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, kc.eps2u)));
        const float w = __builtin_amdgcn_rsqf(r2);
        const float w3 = (w * w) * w;
        tx = fmaf(dx, w3, tx);
        ty = fmaf(dy, w3, ty);
        tz = fmaf(dz, w3, tz);
        tp += w;
    } else {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const bool near = r2 < kc.h2u;
        const float w = __builtin_amdgcn_rsqf(near ? kc.h2u : r2);
        const float k = near ? kc.hinv3u : (w * w) * w;
        const float p = near ? fmaf(-0.5f * kc.hinv3u, r2, 1.5f * kc.hinvu) : w;
        tx = fmaf(dx, k, tx);
        ty = fmaf(dy, k, ty);
        tz = fmaf(dz, k, tz);
        tp += p;
    }
}

// grid: (ceil(nsink/KICK_BLOCK), nsplit), sinks i0 .. i1-1; block y sums
// sources [y*jchunk, min(nsrc, (y+1)*jchunk)) of sources 0 .. nsrc-1.
// Output: partial[(a*nsplit + y)*nsink + i] for a in {x,y,z,pot}, in units
// of 2^-64 (force) and 2^-32 (potential); see the scaling in kick_finalize.
template <int KERNEL, bool NAIVE>
__global__ __launch_bounds__(KICK_BLOCK) void kick_direct(const uint32_t* __restrict__ X,
                                                          const uint32_t* __restrict__ Y,
                                                          const uint32_t* __restrict__ Z, int i0, int i1, int nsrc,
                                                          int jchunk, KickConst kc,
                                                          double* __restrict__ partial) {
    __shared__ uint4 tile[2][KICK_TILE];
    const int tid = threadIdx.x;
    const int ns = i1 - i0;  // sinks i0 .. i1-1
    const int i = blockIdx.x * KICK_BLOCK + tid;
    const int ic = i < ns ? i0 + i : i1 - 1;
    const uint32_t xi = X[ic], yi = Y[ic], zi = Z[ic];

    const int j0 = blockIdx.y * jchunk;
    const int j1 = min(nsrc, j0 + jchunk);
    float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f;
    float cx = 0.f, cy = 0.f, cz = 0.f, cp = 0.f;

    uint4 pre = make_uint4(0u, 0u, 0u, 0u);
    if (j0 + tid < j1) pre = make_uint4(X[j0 + tid], Y[j0 + tid], Z[j0 + tid], 0u);
    int buf = 0;
    for (int jt = j0; jt < j1; jt += KICK_TILE) {
        tile[buf][tid] = pre;
        __syncthreads();
        const int jn = jt + KICK_TILE + tid;
        if (jn < j1) pre = make_uint4(X[jn], Y[jn], Z[jn], 0u);
        const int cnt = min(KICK_TILE, j1 - jt);
        float tx = 0.f, ty = 0.f, tz = 0.f, tp = 0.f;
        if (NAIVE) {
            tx = sx;
            ty = sy;
            tz = sz;
            tp = sp;
        }
        const uint4* T = tile[buf];
        if (cnt == KICK_TILE) {
#pragma unroll 16
            for (int j = 0; j < KICK_TILE; ++j) interact<KERNEL>(T[j], xi, yi, zi, kc, tx, ty, tz, tp);
        } else {
            for (int j = 0; j < cnt; ++j) interact<KERNEL>(T[j], xi, yi, zi, kc, tx, ty, tz, tp);
        }
        if (NAIVE) {
            sx = tx;
            sy = ty;
            sz = tz;
            sp = tp;
        } else {
            two_sum(sx, cx, tx);
            two_sum(sy, cy, ty);
            two_sum(sz, cz, tz);
            two_sum(sp, cp, tp);
        }
        buf ^= 1;
    }
    if (i < ns) {
        const long nsp = gridDim.y, y = blockIdx.y;
        partial[(0 * nsp + y) * ns + i] = (double)sx + (double)cx;
        partial[(1 * nsp + y) * ns + i] = (double)sy + (double)cy;
        partial[(2 * nsp + y) * ns + i] = (double)sz + (double)cz;
        partial[(3 * nsp + y) * ns + i] = (double)sp + (double)cp;
    }
}

// Partial sums of the sinks [first, first + count): part[(c*nsplit + y)*count + i - first].
struct PartialSums {
    const double* part;
    int nsplit, first, count;
};

// Sums, for every sink i in [i0, n), the planes of each of the (up to three)
// partial-sum sets that cover i, in the set order given.
__global__ __launch_bounds__(256) void kick_finalize(PartialSums s0, PartialSums s1, PartialSums s2,
                                                     int n, int i0, double fscale, double pscale, EpiConst e,
                                                     float* __restrict__ vx, float* __restrict__ vy,
                                                     float* __restrict__ vz, int8_t* __restrict__ rung,
                                                     float* __restrict__ fx, float* __restrict__ fy,
                                                     float* __restrict__ fz, float* __restrict__ pot,
                                                     int* __restrict__ maxrung) {
    const int ns = n - i0;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i = i0 + t;
    int rmax = 0;
    if (t < ns) {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for (const PartialSums& ps : {s0, s1, s2}) {
            const int k = i - ps.first;
            if (!ps.part || k < 0 || k >= ps.count) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double s = 0.0;
                for (int y = 0; y < ps.nsplit; ++y) s += ps.part[((long)c * ps.nsplit + y) * ps.count + k];
                a[c] += s;
            }
        }
        const float f0 = (float)(a[0] * fscale), f1 = (float)(a[1] * fscale), f2 = (float)(a[2] * fscale);
        fx[i] = f0;
        fy[i] = f1;
        fz[i] = f2;
        pot[i] = (float)(a[3] * pscale);
        if (e.kick_ascending || e.descending) {
            float u = vx[i], v = vy[i], w = vz[i];
            int8_t r = rung[i];
            epilogue(f0, f1, f2, u, v, w, r, e, rmax);
            vx[i] = u;
            vy[i] = v;
            vz[i] = w;
            rung[i] = r;
        }
    }
    // wave-level max, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) rmax = max(rmax, __shfl_xor(rmax, off));
    if ((threadIdx.x & 63) == 0 && rmax > 0) atomicMax(maxrung, rmax);
}

// Epilogue over stored forces (mesh-based kicks).
__global__ __launch_bounds__(256) void epilogue_kernel(const float* __restrict__ fx, const float* __restrict__ fy,
                                                       const float* __restrict__ fz, int n, int i0, EpiConst e,
                                                       float* __restrict__ vx, float* __restrict__ vy,
                                                       float* __restrict__ vz, int8_t* __restrict__ rung,
                                                       int* __restrict__ maxrung) {
    const int i = i0 + blockIdx.x * 256 + threadIdx.x;
    int rmax = 0;
    if (i < n) {
        float u = vx[i], v = vy[i], w = vz[i];
        int8_t r = rung[i];
        epilogue(fx[i], fy[i], fz[i], u, v, w, r, e, rmax);
        vx[i] = u;
        vy[i] = v;
        vz[i] = w;
        rung[i] = r;
    }
    for (int off = 32; off > 0; off >>= 1) rmax = max(rmax, __shfl_xor(rmax, off));
    if ((threadIdx.x & 63) == 0 && rmax > 0) atomicMax(maxrung, rmax);
}

__global__ void zero_int(int* p) { *p = 0; }

// ---- symmetric all-pairs kick (Newton's third law): each pair evaluated once.
//
// Superblocks of SYM_BLOCK = 1024 particles, one per workgroup of 16 waves
// (64 sinks per wave, positions in registers).  A superblock pair (I, J) is a
// grid of 64x64 tiles: the wave's 64 sinks against 64 sources of J in LDS.  In
// a tile, at step t lane l pairs with source (l + t) & 63; the sink-side sums
// stay in the lane, and the source-side sums (minus the same force, plus the
// same potential) rotate one lane per step (DPP wave_rol:1) so that each lane
// always holds the running sums of the source it is paired with.  After the
// tile, lane l holds the sums of source (l + t_last) & 63.  Per chunk of 64
// sources the 16 waves' source-side sums are added in FP64 in a fixed order.
//
// Schedule (deterministic, race free): a diagonal launch computes each
// superblock with itself (tile (w, c) for c > w in full, c == w over steps
// 1..32, the last one on half the lanes, plus the self potential) and
// initialises plane 0 of the partial sums; then launches over rounds r:
// workgroup g pairs superblock g with J = (g + r) mod Gs (r = Gs/2 only for
// g < Gs/2 when Gs is even), keeps its sink-side sums in registers across the
// launch's rounds (one read-modify-write per launch) and adds the source-side
// sums of round r into plane 1 + (r - r0).  Below 1024 superblocks a launch's
// R rounds are dealt round-robin to `groups` workgroups per superblock, group
// q writing its sink-side sums to plane 0 (q = 0) or R + q: within a launch
// every (plane, superblock) slot has exactly one writer.  kick_finalize then
// adds the planes in order.
constexpr int SYM_WAVES = 16;  // chunks of 64 per superblock (and waves of the diagonal kernel)
constexpr int SYM_BLOCK = 64 * SYM_WAVES;
#ifndef SYM_NS
#define SYM_NS 4  // sinks per lane in the off-diagonal rounds
#endif
constexpr int SYM_THREADS = SYM_BLOCK / SYM_NS;
// smallest core the symmetric kernels take by default: 64 superblocks.  A
// round takes ~0.15 ms per workgroup whatever the core size, so cores of
// 4-32 superblocks (full kicks of 4,096-32,768 particles) ran 1.2-4x slower
// than the ordered kernel; 64 superblocks and up ran 3-22% faster.
constexpr int SYM_MIN_CORE = 64 * SYM_BLOCK;
static_assert(SYM_THREADS >= 256, "the chunk reduction uses 256 threads (4 components x 64 sources)");

__device__ __forceinline__ float rol1(float v) {  // lane l <- lane l+1 (mod 64)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x134, 0xF, 0xF, true));
}

// Separation (2^-32 units) and kernel factors of one pair: F = d k, pot p.
template <int KERNEL>
__device__ __forceinline__ void pair_k(const uint4 s, uint32_t xi, uint32_t yi, uint32_t zi, const KickConst& kc,
                                       float& dx, float& dy, float& dz, float& k, float& p) {
    dx = (float)(int32_t)(s.x - xi);
    dy = (float)(int32_t)(s.y - yi);
    dz = (float)(int32_t)(s.z - zi);
    if constexpr (KERNEL == CT2_KERNEL_PLUMMER) {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, kc.eps2u)));
        const float w = __builtin_amdgcn_rsqf(r2);
        k = (w * w) * w;
        p = w;
    } else {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const bool near = r2 < kc.h2u;
        const float w = __builtin_amdgcn_rsqf(near ? kc.h2u : r2);
        k = near ? kc.hinv3u : (w * w) * w;
        p = near ? fmaf(-0.5f * kc.hinv3u, r2, 1.5f * kc.hinvu) : w;
    }
}

template <int KERNEL>
__device__ __forceinline__ void pair_terms(const uint4 s, uint32_t xi, uint32_t yi, uint32_t zi, const KickConst& kc,
                                           float& fx, float& fy, float& fz, float& p) {
    float dx, dy, dz, k;
    pair_k<KERNEL>(s, xi, yi, zi, kc, dx, dy, dz, k, p);
    fx = dx * k;
    fy = dy * k;
    fz = dz * k;
}

// Steps [T0, T1) of a tile; HALF: the last step only on lanes < 32.  The
// chunk is stored twice in a row in LDS (128 entries), so lane l reads
// Sl[t] = chunk[(l + t) & 63] at an immediate offset from Sl = chunk + l.
template <int KERNEL, int T0, int T1, bool HALF>
__device__ __forceinline__ void sym_tile(const uint4* Sl, int lane, uint32_t xi, uint32_t yi, uint32_t zi,
                                         const KickConst& kc, float& tx, float& ty, float& tz, float& tp, float& ax,
                                         float& ay, float& az, float& ap) {
#pragma unroll 8
    for (int t = T0; t < T1; ++t) {
        const uint4 s = Sl[t];
        float fx, fy, fz, p;
        pair_terms<KERNEL>(s, xi, yi, zi, kc, fx, fy, fz, p);
        if (HALF && t == T1 - 1 && lane >= 32) fx = fy = fz = p = 0.0f;
        tx += fx;
        ty += fy;
        tz += fz;
        tp += p;
        ax = rol1(ax) - fx;
        ay = rol1(ay) - fy;
        az = rol1(az) - fz;
        ap = rol1(ap) + p;
    }
}

template <int KERNEL>
__global__ __launch_bounds__(SYM_BLOCK) void kick_sym_diag(const uint4* __restrict__ P, int n, int ny, KickConst kc,
                                                           double* __restrict__ part) {
    __shared__ uint4 S[SYM_WAVES][128];
    __shared__ float4 jr[SYM_WAVES][64];
    __shared__ double jacc[4][SYM_BLOCK];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long i = (long)blockIdx.x * SYM_BLOCK + tid;
    const uint4 me = P[i];
    S[w][lane] = me;
    S[w][lane + 64] = me;
    __syncthreads();
    float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, cp = 0.f;
    for (int c = 0; c < SYM_WAVES; ++c) {
        float tx = 0.f, ty = 0.f, tz = 0.f, tp = 0.f, ax = 0.f, ay = 0.f, az = 0.f, ap = 0.f;
        if (w < c) {  // wave-uniform
            sym_tile<KERNEL, 0, 64, false>(S[c] + lane, lane, me.x, me.y, me.z, kc, tx, ty, tz, tp, ax, ay, az, ap);
            jr[w][(lane + 63) & 63] = make_float4(ax, ay, az, ap);
        } else if (w == c) {
            sym_tile<KERNEL, 1, 33, true>(S[c] + lane, lane, me.x, me.y, me.z, kc, tx, ty, tz, tp, ax, ay, az, ap);
            jr[w][(lane + 32) & 63] = make_float4(ax, ay, az, ap);
        }
        two_sum(sx, cx, tx);
        two_sum(sy, cy, ty);
        two_sum(sz, cz, tz);
        two_sum(sp, cp, tp);
        __syncthreads();
        if (tid < 256) {
            const int comp = tid >> 6, jj = tid & 63;
            double acc = 0.0;
            for (int v = 0; v <= c; ++v) acc += (double)((const float*)&jr[v][jj])[comp];
            jacc[comp][c * 64 + jj] = acc;
        }
        __syncthreads();
    }
    // the self pair: potential only
    float fx, fy, fz, pself;
    pair_terms<KERNEL>(me, me.x, me.y, me.z, kc, fx, fy, fz, pself);
    two_sum(sp, cp, pself);
    part[(0L * ny) * n + i] = (double)sx + (double)cx + jacc[0][tid];
    part[(1L * ny) * n + i] = (double)sy + (double)cy + jacc[1][tid];
    part[(2L * ny) * n + i] = (double)sz + (double)cz + jacc[2][tid];
    part[(3L * ny) * n + i] = (double)sp + (double)cp + jacc[3][tid];
}

// Off-diagonal rounds on spatially ordered superblocks.  Each lane carries NS
// sinks (the wave 64*NS consecutive particles of the superblock's 1024), all
// paired with the same source at every step, so one rotation of the
// source-side sums serves NS pairs.  Sources stream through LDS one chunk of
// 64 at a time (double-buffered: the next chunk is fetched during the current
// one, one barrier per chunk).
//
// Separations in FP32 offsets.  Every chunk has a reference point (the
// centre of its bounding box) and the exact integer range [lo, hi] of its
// sources' offsets from it (ChunkRef).  Per chunk a wave forms its sinks'
// offsets u = (int32)(x_i - ref) and the sources' v = (int32)(x_j - ref)
// are in LDS as floats, so a separation is one FP32 subtraction v - u
// instead of an integer subtraction and a conversion.  That is the exact
// minimum image whenever v - u cannot leave [-2^31, 2^31) for any pair of
// the (wave, chunk) block; the wave decides that per axis from its sinks'
// integer range (computed once) and the chunk's, in scalar arithmetic, and
// takes the exact integer separation on any axis where it could (MASK bit).
// The offsets are rounded to FP32 once each, so a separation carries an
// error of at most 2^-24 (|u| + |v|): relative to the separation, 2^-24
// times (1 + twice the chunk's half extent / separation).
struct ChunkRef {
    uint4 ref;  // x, y, z reference (fixed32), w unused
    int4 lo;    // per-axis minimum of (int32)(x_j - ref) over the chunk
    int4 hi;    // per-axis maximum
};

template <int KERNEL>
__device__ __forceinline__ void pair_kd(float dx, float dy, float dz, const KickConst& kc, float& k, float& p) {
    if constexpr (KERNEL == CT2_KERNEL_PLUMMER) {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, kc.eps2u)));
        const float w = __builtin_amdgcn_rsqf(r2);
        k = (w * w) * w;
        p = w;
    } else {
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const bool near = r2 < kc.h2u;
        const float w = __builtin_amdgcn_rsqf(near ? kc.h2u : r2);
        k = near ? kc.hinv3u : (w * w) * w;
        p = near ? fmaf(-0.5f * kc.hinv3u, r2, 1.5f * kc.hinvu) : w;
    }
}

// Measured on gfx950 at N = 2^21 (ms per kick, one box): as written 915-941;
// unrolling the step loop by 2 or 4, keeping the softening constant in a VGPR
// (v_fma_f32 with three VGPR operands) or rotating the source-side sums with
// ds_bpermute instead of DPP: 977-1,099 in every combination.
#ifndef SYM_STEP_UNROLL
#define SYM_STEP_UNROLL 1
#endif

// The 64 steps of one (wave, chunk) block.  MASK bit a: axis a takes the
// exact integer separation (int sources Si, int sinks xi) instead of v - u.
template <int KERNEL, int MASK>
__device__ __forceinline__ void sym_chunk(const float4* Sf, const uint4* Si, const float (&ux)[SYM_NS],
                                          const float (&uy)[SYM_NS], const float (&uz)[SYM_NS],
                                          const uint32_t (&xi)[SYM_NS], const uint32_t (&yi)[SYM_NS],
                                          const uint32_t (&zi)[SYM_NS], const KickConst& kc, float (&tx)[SYM_NS],
                                          float (&ty)[SYM_NS], float (&tz)[SYM_NS], float (&tp)[SYM_NS], float& ax,
                                          float& ay, float& az, float& ap) {
#pragma unroll SYM_STEP_UNROLL
    for (int t = 0; t < 64; ++t) {
        const float4 sf = Sf[t];
        uint4 si = make_uint4(0u, 0u, 0u, 0u);
        if constexpr (MASK != 0) si = Si[t];
        float gx, gy, gz, gp;  // source-side sum of this step, started by sink 0
#pragma unroll
        for (int q = 0; q < SYM_NS; ++q) {
            const float dx = (MASK & 1) ? (float)(int32_t)(si.x - xi[q]) : sf.x - ux[q];
            const float dy = (MASK & 2) ? (float)(int32_t)(si.y - yi[q]) : sf.y - uy[q];
            const float dz = (MASK & 4) ? (float)(int32_t)(si.z - zi[q]) : sf.z - uz[q];
            float k, p;
            pair_kd<KERNEL>(dx, dy, dz, kc, k, p);
            tx[q] = fmaf(dx, k, tx[q]);
            ty[q] = fmaf(dy, k, ty[q]);
            tz[q] = fmaf(dz, k, tz[q]);
            tp[q] += p;
            if (q == 0) {
                gx = dx * k;
                gy = dy * k;
                gz = dz * k;
                gp = p;
            } else {
                gx = fmaf(dx, k, gx);
                gy = fmaf(dy, k, gy);
                gz = fmaf(dz, k, gz);
                gp += p;
            }
        }
        ax = rol1(ax) - gx;
        ay = rol1(ay) - gy;
        az = rol1(az) - gz;
        ap = rol1(ap) + gp;
    }
}

__device__ __forceinline__ int wave_min_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Whether v - u stays in [-2^31, 2^31) for every sink offset u and source
// offset v of a (wave, chunk) block on one axis: the wave's sinks lie at
// [ulo, uhi] from its own reference rw, the chunk's sources at [vlo, vhi]
// from ref (all exact integers, scalar).
__device__ __forceinline__ bool axis_float_safe(uint32_t rw, int ulo, int uhi, uint32_t ref, int vlo, int vhi) {
    const long d = (long)(int32_t)(ref - rw);
    const long lo = (long)ulo - d, hi = (long)uhi - d;  // the sinks' offsets from ref, unwrapped
    if (lo < -2147483648L || hi > 2147483647L) return false;
    return (long)vhi - lo < 2147483648L && (long)vlo - hi >= -2147483648L;
}

// Stores chunk (J, c) of P into LDS buffer (float offsets from its reference
// and the integers, each twice in a row): lane l of one wave, source l.
__device__ __forceinline__ void sym_store_chunk(const uint4 src, const ChunkRef& cr, float4* Sf, uint4* Si,
                                                int lane) {
    const float4 f = make_float4((float)(int32_t)(src.x - cr.ref.x), (float)(int32_t)(src.y - cr.ref.y),
                                 (float)(int32_t)(src.z - cr.ref.z), 0.0f);
    Sf[lane] = f;
    Sf[lane + 64] = f;
    Si[lane] = src;
    Si[lane + 64] = src;
}

template <int KERNEL>
// blockIdx.y = q of gridDim.y round groups: workgroup (g, q) takes rounds
// r0 + q, r0 + q + gridDim.y, ... < r1, so that small cores still fill the
// GPU; group q adds its sink-side sums to plane q == 0 ? 0 : R + q.
__global__ __launch_bounds__(SYM_THREADS) void kick_sym_pairs(const uint4* __restrict__ P,
                                                              const ChunkRef* __restrict__ refs, int n, int ny,
                                                              int gs, int r0, int r1, int R, KickConst kc,
                                                              double* __restrict__ part) {
    constexpr int NWAVE = SYM_THREADS / 64;
    __shared__ float4 Sf[2][128];
    __shared__ uint4 Si[2][128];
    __shared__ float4 jr[2][NWAVE][64];  // double-buffered: one barrier per chunk
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int g = blockIdx.x;
    uint32_t xi[SYM_NS], yi[SYM_NS], zi[SYM_NS];
    float sx[SYM_NS], sy[SYM_NS], sz[SYM_NS], sp[SYM_NS], cx[SYM_NS], cy[SYM_NS], cz[SYM_NS], cp[SYM_NS];
#pragma unroll
    for (int q = 0; q < SYM_NS; ++q) {
        const uint4 me = P[(long)g * SYM_BLOCK + (w * SYM_NS + q) * 64 + lane];
        xi[q] = me.x;
        yi[q] = me.y;
        zi[q] = me.z;
        sx[q] = sy[q] = sz[q] = sp[q] = cx[q] = cy[q] = cz[q] = cp[q] = 0.f;
    }
    // the wave's sinks: integer offsets from its first sink, range per axis
    const uint32_t rwx = (uint32_t)__builtin_amdgcn_readfirstlane((int)xi[0]);
    const uint32_t rwy = (uint32_t)__builtin_amdgcn_readfirstlane((int)yi[0]);
    const uint32_t rwz = (uint32_t)__builtin_amdgcn_readfirstlane((int)zi[0]);
    int ulx = INT32_MAX, uhx = INT32_MIN, uly = INT32_MAX, uhy = INT32_MIN, ulz = INT32_MAX, uhz = INT32_MIN;
#pragma unroll
    for (int q = 0; q < SYM_NS; ++q) {
        const int ux = (int32_t)(xi[q] - rwx), uy = (int32_t)(yi[q] - rwy), uz = (int32_t)(zi[q] - rwz);
        ulx = min(ulx, ux);
        uhx = max(uhx, ux);
        uly = min(uly, uy);
        uhy = max(uhy, uy);
        ulz = min(ulz, uz);
        uhz = max(uhz, uz);
    }
    ulx = wave_min_i32(ulx);
    uhx = wave_max_i32(uhx);
    uly = wave_min_i32(uly);
    uhy = wave_max_i32(uhy);
    ulz = wave_min_i32(ulz);
    uhz = wave_max_i32(uhz);

    const bool even = (gs & 1) == 0;
    auto valid = [&](int r) { return !(even && r == gs / 2 && g >= gs / 2); };  // that pair belongs to g - gs/2
    auto next_round = [&](int r) {
        r += (int)gridDim.y;
        while (r < r1 && !valid(r)) r += (int)gridDim.y;
        return r;
    };
    int r = r0 + (int)blockIdx.y;
    while (r < r1 && !valid(r)) r += (int)gridDim.y;
    if (r < r1 && w == 0) {
        const long J = (g + r) % gs;
        sym_store_chunk(P[J * SYM_BLOCK + lane], refs[J * SYM_WAVES], Sf[0], Si[0], lane);
    }
    __syncthreads();
    int b = 0;
    for (; r < r1; r = next_round(r)) {
        const long J = (g + r) % gs;
        for (int c = 0; c < SYM_WAVES; ++c) {
            // the next chunk, fetched now and stored after this one's steps
            const int rn = c + 1 < SYM_WAVES ? r : next_round(r);
            const int cn = c + 1 < SYM_WAVES ? c + 1 : 0;
            const long Jn = (g + rn) % gs;
            uint4 pre = make_uint4(0u, 0u, 0u, 0u);
            if (rn < r1 && w == 0) pre = P[Jn * SYM_BLOCK + cn * 64 + lane];
            const ChunkRef cr = refs[J * SYM_WAVES + c];
            const int mask = (axis_float_safe(rwx, ulx, uhx, cr.ref.x, cr.lo.x, cr.hi.x) ? 0 : 1) |
                             (axis_float_safe(rwy, uly, uhy, cr.ref.y, cr.lo.y, cr.hi.y) ? 0 : 2) |
                             (axis_float_safe(rwz, ulz, uhz, cr.ref.z, cr.lo.z, cr.hi.z) ? 0 : 4);
            float ux[SYM_NS], uy[SYM_NS], uz[SYM_NS];
#pragma unroll
            for (int q = 0; q < SYM_NS; ++q) {
                ux[q] = (float)(int32_t)(xi[q] - cr.ref.x);
                uy[q] = (float)(int32_t)(yi[q] - cr.ref.y);
                uz[q] = (float)(int32_t)(zi[q] - cr.ref.z);
            }
            float tx[SYM_NS], ty[SYM_NS], tz[SYM_NS], tp[SYM_NS];
#pragma unroll
            for (int q = 0; q < SYM_NS; ++q) tx[q] = ty[q] = tz[q] = tp[q] = 0.f;
            float ax = 0.f, ay = 0.f, az = 0.f, ap = 0.f;
            const float4* Sl = Sf[b] + lane;
            const uint4* Il = Si[b] + lane;
#define CT2_SYM_CASE(M)                                                                                          \
    case M:                                                                                                      \
        sym_chunk<KERNEL, M>(Sl, Il, ux, uy, uz, xi, yi, zi, kc, tx, ty, tz, tp, ax, ay, az, ap);                \
        break;
            switch (mask) {
                CT2_SYM_CASE(0)
                CT2_SYM_CASE(1)
                CT2_SYM_CASE(2)
                CT2_SYM_CASE(3)
                CT2_SYM_CASE(4)
                CT2_SYM_CASE(5)
                CT2_SYM_CASE(6)
                CT2_SYM_CASE(7)
            }
#undef CT2_SYM_CASE
#pragma unroll
            for (int q = 0; q < SYM_NS; ++q) {
                two_sum(sx[q], cx[q], tx[q]);
                two_sum(sy[q], cy[q], ty[q]);
                two_sum(sz[q], cz[q], tz[q]);
                two_sum(sp[q], cp[q], tp[q]);
            }
            jr[b][w][(lane + 63) & 63] = make_float4(ax, ay, az, ap);
            // buffer b^1 held the previous chunk, whose steps ended before the last barrier
            if (rn < r1 && w == 0) sym_store_chunk(pre, refs[Jn * SYM_WAVES + cn], Sf[b ^ 1], Si[b ^ 1], lane);
            __syncthreads();
            if (tid < 256) {
                const int comp = tid >> 6, jj = tid & 63;
                double acc = 0.0;
#pragma unroll
                for (int v = 0; v < NWAVE; ++v) acc += (double)((const float*)&jr[b][v][jj])[comp];
                // the only update of this slot in this launch: a no-return FP64
                // atomic keeps the wave from waiting on the read-modify-write
                unsafeAtomicAdd(part + ((long)comp * ny + 1 + (r - r0)) * n + J * SYM_BLOCK + c * 64 + jj, acc);
            }
            b ^= 1;
        }
    }
    const long sp0 = blockIdx.y == 0 ? 0 : R + blockIdx.y;  // this group's sink-side plane
#pragma unroll
    for (int q = 0; q < SYM_NS; ++q) {
        const long i = (long)g * SYM_BLOCK + (w * SYM_NS + q) * 64 + lane;
        part[(0L * ny + sp0) * n + i] += (double)sx[q] + (double)cx[q];
        part[(1L * ny + sp0) * n + i] += (double)sy[q] + (double)cy[q];
        part[(2L * ny + sp0) * n + i] += (double)sz[q] + (double)cz[q];
        part[(3L * ny + sp0) * n + i] += (double)sp[q] + (double)cp[q];
    }
}

// Spatial order of the symmetric core: the key of particle s is the Morton
// index of its cell in a 1024^3 grid (10 bits per axis, x highest).
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(256) void sym_keys(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                const uint32_t* __restrict__ Z, int m, uint32_t* __restrict__ key,
                                                uint32_t* __restrict__ idx) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= m) return;
    key[s] = (spread3(X[s] >> 22) << 2) | (spread3(Y[s] >> 22) << 1) | spread3(Z[s] >> 22);
    idx[s] = (uint32_t)s;
}

// P[s] = (x, y, z, core index) of the particle at sorted slot s.
__global__ __launch_bounds__(256) void sym_gather(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                  const uint32_t* __restrict__ Z, const uint32_t* __restrict__ idx,
                                                  int m, uint4* __restrict__ P) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= m) return;
    const uint32_t i = idx[s];
    P[s] = make_uint4(X[i], Y[i], Z[i], i);
}

// One wave per chunk of 64: its bounding-box centre and the exact range of
// the offsets from it.
__global__ __launch_bounds__(64) void sym_chunk_refs(const uint4* __restrict__ P, ChunkRef* __restrict__ refs) {
    const uint4 p = P[(long)blockIdx.x * 64 + threadIdx.x];
    const uint32_t lx = wave_min_u32(p.x), hx = wave_max_u32(p.x);
    const uint32_t ly = wave_min_u32(p.y), hy = wave_max_u32(p.y);
    const uint32_t lz = wave_min_u32(p.z), hz = wave_max_u32(p.z);
    const uint4 ref = make_uint4(lx + (hx - lx) / 2, ly + (hy - ly) / 2, lz + (hz - lz) / 2, 0u);
    const int vx = (int32_t)(p.x - ref.x), vy = (int32_t)(p.y - ref.y), vz = (int32_t)(p.z - ref.z);
    const int4 lo = make_int4(wave_min_i32(vx), wave_min_i32(vy), wave_min_i32(vz), 0);
    const int4 hi = make_int4(wave_max_i32(vx), wave_max_i32(vy), wave_max_i32(vz), 0);
    if (threadIdx.x == 0) refs[blockIdx.x] = ChunkRef{ref, lo, hi};
}

// Sums the planes of the core's partial sums in plane order for each sorted
// slot and writes the totals to the core index (out[c * m + i]).
__global__ __launch_bounds__(256) void sym_unsort(const double* __restrict__ part, int ny, int m,
                                                  const uint4* __restrict__ P, double* __restrict__ out) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= m) return;
    const uint32_t i = P[s].w;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double v = 0.0;
        for (int y = 0; y < ny; ++y) v += part[((long)c * ny + y) * m + s];
        out[(long)c * m + i] = v;
    }
}

__global__ __launch_bounds__(256) void fill_f64(double* __restrict__ p, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// FP64 judge: block s handles sink idx[s]; 256 lanes stride the sources.
template <int KERNEL>
__global__ __launch_bounds__(256) void kick_sample_fp64(const uint32_t* __restrict__ X,
                                                        const uint32_t* __restrict__ Y,
                                                        const uint32_t* __restrict__ Z, int n,
                                                        const long* __restrict__ idx, int m, double soft,
                                                        double gm, double* __restrict__ out) {
    __shared__ double red[4][4];
    const int s = blockIdx.x;
    const long i = idx[s];
    const uint32_t xi = X[i], yi = Y[i], zi = Z[i];
    const double e2 = soft * soft, hinv = 1.0 / soft, hinv3 = hinv * hinv * hinv;
    double ax = 0.0, ay = 0.0, az = 0.0, ap = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double dx = (double)(int32_t)(X[j] - xi) * 0x1p-32;
        const double dy = (double)(int32_t)(Y[j] - yi) * 0x1p-32;
        const double dz = (double)(int32_t)(Z[j] - zi) * 0x1p-32;
        const double r2 = dx * dx + dy * dy + dz * dz;
        double k, p;
        if constexpr (KERNEL == CT2_KERNEL_PLUMMER) {
            const double w = 1.0 / sqrt(r2 + e2);
            k = w * w * w;
            p = w;
        } else {
            if (r2 < e2) {
                k = hinv3;
                p = 1.5 * hinv - 0.5 * r2 * hinv3;
            } else {
                const double w = 1.0 / sqrt(r2);
                k = w * w * w;
                p = w;
            }
        }
        ax += dx * k;
        ay += dy * k;
        az += dz * k;
        ap += p;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_xor(ax, off);
        ay += __shfl_xor(ay, off);
        az += __shfl_xor(az, off);
        ap += __shfl_xor(ap, off);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave][0] = ax;
        red[wave][1] = ay;
        red[wave][2] = az;
        red[wave][3] = ap;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int c = threadIdx.x;
        const double t = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
        out[(long)c * m + s] = gm * t;
    }
}

uint4* g_packed = nullptr;  // the symmetric core in spatial order: (x, y, z, core index)
size_t g_packed_bytes = 0;
ChunkRef* g_refs = nullptr;  // per chunk of 64 of g_packed
size_t g_refs_bytes = 0;
uint32_t* g_sort_kv = nullptr;  // keys in/out, indices in/out of the spatial sort
size_t g_sort_kv_bytes = 0;
void* g_sort_tmp = nullptr;
size_t g_sort_tmp_bytes = 0;
double* g_sym_out = nullptr;  // the core's summed partials, by core index
size_t g_sym_out_bytes = 0;
double* g_part_a = nullptr;  // ordered pass beside the symmetric core: core sinks x other sources
size_t g_part_a_bytes = 0;
double* g_part_b = nullptr;  // ordered pass of the sinks before the core x all sources
size_t g_part_b_bytes = 0;

// Ordered pass: sinks [i0, i1) against sources [0, nsrc), source chunks
// sized so that the grid has enough workgroups to fill 256 CUs; partial sums
// into *buf (grown as needed).  Returns the set's descriptor.
// With fixed != nullptr the partial sums go to that buffer of fixed_bytes
// instead (no allocation): the source split is reduced until they fit.
int ordered_pass(Context& c, int i0, int i1, int nsrc, bool naive, bool plummer, const KickConst& kc, double** buf,
                 size_t* bytes, PartialSums* out, double* fixed = nullptr, size_t fixed_bytes = 0) {
    const int ns = i1 - i0;
    const int nb = (ns + KICK_BLOCK - 1) / KICK_BLOCK;
    const int max_split = std::max(1, nsrc / (4 * KICK_TILE));
    int nsplit = std::min(max_split, std::max(1, (4096 + nb - 1) / nb));
    const size_t per_split = sizeof(double) * 4 * (size_t)ns;
    if (fixed) nsplit = (int)std::max<size_t>(1, std::min<size_t>(nsplit, fixed_bytes / per_split));
    int jchunk = (nsrc + nsplit - 1) / nsplit;
    jchunk = (jchunk + KICK_TILE - 1) / KICK_TILE * KICK_TILE;
    const int nsplit_eff = (nsrc + jchunk - 1) / jchunk;
    double* part = fixed;
    if (fixed) {
        if (per_split * nsplit_eff > fixed_bytes)
            return set_error(CT2_ERR_INVALID, "kick partial sums of %d sinks exceed the fixed %zu-byte buffer", ns,
                             fixed_bytes);
    } else {
        CT2_TRY(ensure_scratch((void**)buf, bytes, per_split * nsplit_eff, "kick partial sums"));
        part = *buf;
    }
    const dim3 grid(nb, nsplit_eff);
    if (plummer) {
        if (naive)
            kick_direct<CT2_KERNEL_PLUMMER, true><<<grid, KICK_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, i0, i1, nsrc, jchunk, kc, part);
        else
            kick_direct<CT2_KERNEL_PLUMMER, false><<<grid, KICK_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, i0, i1, nsrc, jchunk, kc, part);
    } else {
        if (naive)
            kick_direct<CT2_KERNEL_CT2, true><<<grid, KICK_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, i0, i1, nsrc, jchunk, kc, part);
        else
            kick_direct<CT2_KERNEL_CT2, false><<<grid, KICK_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, i0, i1, nsrc, jchunk, kc, part);
    }
    CT2_LAUNCH_CHECK();
    *out = PartialSums{part, nsplit_eff, i0, ns};
    return CT2_OK;
}

// Symmetric pass over the core [c0, c0 + m) (m a multiple of SYM_BLOCK):
// every unordered pair of the core once, partial sums of the core's sinks
// into c.partial.
int symmetric_pass(Context& c, int c0, int m, bool plummer, const KickConst& kc, PartialSums* out) {
    const int gs = m / SYM_BLOCK, rmax = gs / 2;
    const size_t plane = sizeof(double) * 4 * (size_t)m;
    const int rcap = (int)std::max<size_t>(1, ((size_t)4 << 30) / plane - 1);
    const int R = std::max(1, std::min({16, rmax, rcap}));
    // round groups per launch: enough workgroups (>= 1024) for small cores
    const int groups = std::max(1, std::min(R, (1024 + gs - 1) / gs));
    const int ny = R + groups;  // sink-side planes 0 and R+1 .. R+groups-1, source-side planes 1 .. R
    CT2_TRY(ensure_scratch((void**)&c.partial, &c.partial_bytes, plane * ny, "kick partial sums"));
    CT2_TRY(ensure_scratch((void**)&g_packed, &g_packed_bytes, sizeof(uint4) * (size_t)m, "packed positions"));
    CT2_TRY(ensure_scratch((void**)&g_refs, &g_refs_bytes, sizeof(ChunkRef) * (size_t)(m / 64), "chunk references"));
    CT2_TRY(ensure_scratch((void**)&g_sort_kv, &g_sort_kv_bytes, sizeof(uint32_t) * 4 * (size_t)m, "spatial sort"));
    CT2_TRY(ensure_scratch((void**)&g_sym_out, &g_sym_out_bytes, plane, "core partial sums"));
    // spatial order: (Morton key, core index) pairs sorted stably by key
    uint32_t *kin = g_sort_kv, *kout = g_sort_kv + m, *vin = g_sort_kv + 2L * m, *vout = g_sort_kv + 3L * m;
    const unsigned pb = (unsigned)((m + 255) / 256);
    sym_keys<<<pb, 256, 0, c.stream>>>(c.x + c0, c.y + c0, c.z + c0, m, kin, vin);
    CT2_LAUNCH_CHECK();
    size_t tmp = 0;
    CT2_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, m, 0, 30, c.stream));
    CT2_TRY(ensure_scratch(&g_sort_tmp, &g_sort_tmp_bytes, std::max<size_t>(tmp, 256), "spatial sort"));
    CT2_HIP(hipcub::DeviceRadixSort::SortPairs(g_sort_tmp, tmp, kin, kout, vin, vout, m, 0, 30, c.stream));
    sym_gather<<<pb, 256, 0, c.stream>>>(c.x + c0, c.y + c0, c.z + c0, vout, m, g_packed);
    CT2_LAUNCH_CHECK();
    sym_chunk_refs<<<m / 64, 64, 0, c.stream>>>(g_packed, g_refs);
    CT2_LAUNCH_CHECK();
    const long total = 4L * ny * m;
    fill_f64<<<(unsigned)((total + 255) / 256), 256, 0, c.stream>>>(c.partial, total);
    CT2_LAUNCH_CHECK();
    if (plummer)
        kick_sym_diag<CT2_KERNEL_PLUMMER><<<gs, SYM_BLOCK, 0, c.stream>>>(g_packed, m, ny, kc, c.partial);
    else
        kick_sym_diag<CT2_KERNEL_CT2><<<gs, SYM_BLOCK, 0, c.stream>>>(g_packed, m, ny, kc, c.partial);
    CT2_LAUNCH_CHECK();
    for (int r0 = 1; r0 <= rmax; r0 += R) {
        const int r1 = std::min(r0 + R, rmax + 1);
        if (plummer)
            kick_sym_pairs<CT2_KERNEL_PLUMMER><<<dim3(gs, groups), SYM_THREADS, 0, c.stream>>>(
                g_packed, g_refs, m, ny, gs, r0, r1, R, kc, c.partial);
        else
            kick_sym_pairs<CT2_KERNEL_CT2><<<dim3(gs, groups), SYM_THREADS, 0, c.stream>>>(
                g_packed, g_refs, m, ny, gs, r0, r1, R, kc, c.partial);
        CT2_LAUNCH_CHECK();
    }
    sym_unsort<<<pb, 256, 0, c.stream>>>(c.partial, ny, m, g_packed, g_sym_out);
    CT2_LAUNCH_CHECK();
    *out = PartialSums{g_sym_out, 1, c0, m};
    return CT2_OK;
}
// The no-allocation kick's scratch: partial sums of up to
// NOALLOC_MAX_SINKS sinks with one source split (8 MiB), then the max-rung
// counter.  One copy per GPU, held until the process ends; its address on
// each device is looked up once at ct2_init (g_scratch_ptr).
constexpr int NOALLOC_MAX_SINKS = CT2_NOALLOC_MAX_SINKS;
constexpr size_t NOALLOC_PART_BYTES = sizeof(double) * 4 * (size_t)NOALLOC_MAX_SINKS;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_PART_BYTES / sizeof(double) + 32];
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

double* g_sample_out = nullptr;
size_t g_sample_bytes = 0;
long* g_sample_idx = nullptr;
size_t g_sample_idx_bytes = 0;

KickConst kick_const(double h) {
    const double two32 = 4294967296.0;
    KickConst kc;
    kc.eps2u = (float)((h * two32) * (h * two32));
    kc.h2u = kc.eps2u;
    kc.hinvu = (float)(1.0 / (h * two32));
    kc.hinv3u = (float)(1.0 / ((h * two32) * (h * two32) * (h * two32)));
    return kc;
}

// Sums the partial-sum sets, stores the force and runs the epilogue, then
// hands back kick()'s return value from the device counter maxrung.
int finish_kick(Context& c, const ct2_kick_params* p, const PartialSums* sets, int n, int i0, int* maxrung,
                int* max_rung) {
    const double two32 = 4294967296.0;
    const EpiConst e = make_epilogue(p, c.opts.hsoft);
    // force in box units: sum(dxu * w3u) * 2^64 * GM; potential: sum(wu) * 2^32 * GM
    const double fscale = (double)c.opts.GM * two32 * two32;
    const double pscale = (double)c.opts.GM * two32;
    kick_finalize<<<(n - i0 + 255) / 256, 256, 0, c.stream>>>(sets[0], sets[1], sets[2], n, i0, fscale, pscale, e,
                                                              c.vx, c.vy, c.vz, c.rung, c.fx, c.fy, c.fz, c.pot,
                                                              maxrung);
    CT2_LAUNCH_CHECK();
    c.forces_valid = true;
    c.last_path = CT2_PATH_DIRECT;
    if (max_rung) {
        CT2_HIP(hipMemcpyAsync(max_rung, maxrung, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CT2_HIP(hipStreamSynchronize(c.stream));
    }
    return CT2_OK;
}

}  // namespace

int kick_noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return set_error(CT2_ERR_INVALID, "device %d beyond the %d devices the no-allocation kick supports", device,
                         NOALLOC_MAX_DEVICES);
    if (g_scratch_ptr[device]) return CT2_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return set_error(CT2_ERR_HIP, "hipGetSymbolAddress of the no-allocation kick scratch: %s",
                         hipGetErrorString(e));
    }
    g_scratch_ptr[device] = ptr;
    return CT2_OK;
}

void release_kick_buffers() {
    Context& c = ctx();
    dev_free(c.partial);
    c.partial = nullptr;
    c.partial_bytes = 0;
    dev_free(c.maxrung);
    c.maxrung = nullptr;
    for (void** p : {(void**)&g_packed, (void**)&g_refs, (void**)&g_sort_kv, &g_sort_tmp, (void**)&g_sym_out}) {
        dev_free(*p);
        *p = nullptr;
    }
    g_packed_bytes = g_refs_bytes = g_sort_kv_bytes = g_sort_tmp_bytes = g_sym_out_bytes = 0;
    dev_free(g_part_a);
    g_part_a = nullptr;
    g_part_a_bytes = 0;
    dev_free(g_part_b);
    g_part_b = nullptr;
    g_part_b_bytes = 0;
    dev_free(g_sample_out);
    g_sample_out = nullptr;
    g_sample_bytes = 0;
    dev_free(g_sample_idx);
    g_sample_idx = nullptr;
    g_sample_idx_bytes = 0;
}

int check_kick_params(const ct2_kick_params* p) {
    if (!p) return set_error(CT2_ERR_INVALID, "null kick_params");
    if (epilogue_active(p)) {
        if (p->rung < 0 || p->rung > 30) return set_error(CT2_ERR_INVALID, "rung %d out of [0,30]", p->rung);
        if (!(p->t0 > 0.0f)) return set_error(CT2_ERR_INVALID, "t0 must be positive");
    }
    return CT2_OK;
}

int run_epilogue(const ct2_kick_params* p, int* max_rung) {
    Context& c = ctx();
    if (!c.maxrung) CT2_TRY(dev_alloc((void**)&c.maxrung, sizeof(int), "max rung"));
    zero_int<<<1, 1, 0, c.stream>>>(c.maxrung);
    CT2_LAUNCH_CHECK();
    if (epilogue_active(p)) {
        const EpiConst e = make_epilogue(p, c.opts.hsoft);
        const int n = (int)c.n, i0 = (int)c.rung_begin;
        if (n > i0) {
            epilogue_kernel<<<(n - i0 + 255) / 256, 256, 0, c.stream>>>(c.fx, c.fy, c.fz, n, i0, e, c.vx, c.vy, c.vz,
                                                                         c.rung, c.maxrung);
            CT2_LAUNCH_CHECK();
        }
    }
    if (max_rung) {
        CT2_HIP(hipMemcpyAsync(max_rung, c.maxrung, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        CT2_HIP(hipStreamSynchronize(c.stream));
    }
    return CT2_OK;
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_kick(const ct2_kick_params* p, unsigned flags, int* max_rung) {
    CT2_TRY(require_particles());
    CT2_TRY(check_kick_params(p));
    if (flags & ~(unsigned)(CT2_KICK_NAIVE | CT2_KICK_STORE | CT2_KICK_ORDERED | CT2_KICK_SYMMETRIC))
        return set_error(CT2_ERR_INVALID, "unknown kick flags 0x%x", flags);
    if ((flags & CT2_KICK_ORDERED) && (flags & CT2_KICK_SYMMETRIC))
        return set_error(CT2_ERR_INVALID, "CT2_KICK_ORDERED and CT2_KICK_SYMMETRIC exclude each other");
    Context& c = ctx();
    const int n = (int)c.n;
    const int i0 = (int)c.rung_begin;  // sinks: the active rung range [i0, n)
    const int ns = n - i0;
    if (!c.maxrung) CT2_TRY(dev_alloc((void**)&c.maxrung, sizeof(int), "max rung"));
    if (ns <= 0) {  // no active sink: nothing to kick
        if (max_rung) *max_rung = 0;
        c.last_path = CT2_PATH_DIRECT;
        return CT2_OK;
    }
    const KickConst kc = kick_const(c.opts.hsoft);
    const bool naive = flags & CT2_KICK_NAIVE;
    const bool plummer = c.opts.kernel == CT2_KERNEL_PLUMMER;
    zero_int<<<1, 1, 0, c.stream>>>(c.maxrung);
    CT2_LAUNCH_CHECK();
    // Each unordered pair once where possible: the symmetric kernels take the
    // last m = floor(ns / 1024) * 1024 active sinks (the core [c0, n)); the
    // ordered kernel adds the core's pairs with the particles before it and
    // the few (< 1024) active sinks before the core against everything.  A
    // core of fewer than SYM_MIN_CORE sinks is latency-bound (a few rounds of
    // 1024 x 1024 pairs per workgroup), and the ordered kernel alone is faster.
    PartialSums sets[3] = {{nullptr, 0, 0, 0}, {nullptr, 0, 0, 0}, {nullptr, 0, 0, 0}};
    int m = (naive || (flags & CT2_KICK_ORDERED)) ? 0 : ns / SYM_BLOCK * SYM_BLOCK;
    if (m < SYM_MIN_CORE && !(flags & CT2_KICK_SYMMETRIC)) m = 0;
    if (m > 0) {
        const int c0 = n - m;
        CT2_TRY(symmetric_pass(c, c0, m, plummer, kc, &sets[0]));
        if (c0 > 0) CT2_TRY(ordered_pass(c, c0, n, c0, false, plummer, kc, &g_part_a, &g_part_a_bytes, &sets[1]));
        if (c0 > i0) CT2_TRY(ordered_pass(c, i0, c0, n, false, plummer, kc, &g_part_b, &g_part_b_bytes, &sets[2]));
    } else {
        CT2_TRY(ordered_pass(c, i0, n, n, naive, plummer, kc, &c.partial, &c.partial_bytes, &sets[0]));
    }

    CT2_TRY(finish_kick(c, p, sets, n, i0, c.maxrung, max_rung));
    (void)(flags & CT2_KICK_STORE);  // forces are always kept; the flag documents intent
    return CT2_OK;
}

// ct2_kick's ordered path with every buffer in the static scratch: no
// allocation.  Calls that overlap on one GPU share the scratch and race.
extern "C" int ct2_kick_noalloc(const ct2_kick_params* p, unsigned flags, int* max_rung) {
    CT2_TRY(require_particles());
    CT2_TRY(check_kick_params(p));
    if (flags & ~(unsigned)(CT2_KICK_NAIVE | CT2_KICK_STORE | CT2_KICK_ORDERED))
        return set_error(CT2_ERR_INVALID, "ct2_kick_noalloc takes CT2_KICK_NAIVE, CT2_KICK_STORE and CT2_KICK_ORDERED only"
                                          " (flags 0x%x)", flags);
    Context& c = ctx();
    void* base = (c.device >= 0 && c.device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[c.device] : nullptr;
    if (!base) return set_error(CT2_ERR_STATE, "no-allocation scratch not set up on device %d", c.device);
    const int n = (int)c.n;
    const int i0 = (int)c.rung_begin;
    const int ns = n - i0;
    if (ns > NOALLOC_MAX_SINKS)
        return set_error(CT2_ERR_INVALID, "%d active sinks exceed the no-allocation kick's maximum of %d", ns,
                         NOALLOC_MAX_SINKS);
    if (ns <= 0) {
        if (max_rung) *max_rung = 0;
        c.last_path = CT2_PATH_DIRECT;
        return CT2_OK;
    }
    int* maxrung = (int*)((char*)base + NOALLOC_PART_BYTES);
    zero_int<<<1, 1, 0, c.stream>>>(maxrung);
    CT2_LAUNCH_CHECK();
    PartialSums sets[3] = {{nullptr, 0, 0, 0}, {nullptr, 0, 0, 0}, {nullptr, 0, 0, 0}};
    CT2_TRY(ordered_pass(c, i0, n, n, flags & CT2_KICK_NAIVE, c.opts.kernel == CT2_KERNEL_PLUMMER,
                         kick_const(c.opts.hsoft), nullptr, nullptr, &sets[0], (double*)base, NOALLOC_PART_BYTES));
    return finish_kick(c, p, sets, n, i0, maxrung, max_rung);
}

extern "C" int ct2_kick_sample_fp64(const int64_t* idx, size_t m, int periodic, double* ax, double* ay,
                                    double* az, double* pot) {
    CT2_TRY(require_particles());
    Context& c = ctx();
    if (!idx || m == 0) return set_error(CT2_ERR_INVALID, "empty sample");
    for (size_t s = 0; s < m; ++s)
        if (idx[s] < 0 || (size_t)idx[s] >= c.n)
            return set_error(CT2_ERR_INVALID, "sample index %lld out of range", (long long)idx[s]);
    if (m > (size_t)INT32_MAX) return set_error(CT2_ERR_INVALID, "sample too large");
    CT2_TRY(ensure_scratch((void**)&g_sample_out, &g_sample_bytes, sizeof(double) * 4 * m, "judge output"));
    CT2_TRY(ensure_scratch((void**)&g_sample_idx, &g_sample_idx_bytes, sizeof(long) * m, "judge indices"));
    CT2_HIP(hipMemcpyAsync(g_sample_idx, idx, sizeof(long) * m, hipMemcpyHostToDevice, c.stream));
    const double soft = c.opts.hsoft, gm = c.opts.GM;
    if (periodic) {
        if (c.opts.kernel != CT2_KERNEL_PLUMMER)
            return set_error(CT2_ERR_INVALID, "the periodic judge uses Plummer softening only");
        CT2_TRY(ewald_judge(g_sample_idx, (int)m, g_sample_out));
    } else if (c.opts.kernel == CT2_KERNEL_PLUMMER)
        kick_sample_fp64<CT2_KERNEL_PLUMMER><<<(unsigned)m, 256, 0, c.stream>>>(c.x, c.y, c.z, (int)c.n, g_sample_idx,
                                                                               (int)m, soft, gm, g_sample_out);
    else
        kick_sample_fp64<CT2_KERNEL_CT2><<<(unsigned)m, 256, 0, c.stream>>>(c.x, c.y, c.z, (int)c.n, g_sample_idx,
                                                                           (int)m, soft, gm, g_sample_out);
    CT2_LAUNCH_CHECK();
    double* dst[4] = {ax, ay, az, pot};
    for (int a = 0; a < 4; ++a)
        if (dst[a])
            CT2_HIP(hipMemcpyAsync(dst[a], g_sample_out + (size_t)a * m, sizeof(double) * m, hipMemcpyDeviceToHost,
                                   c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}
