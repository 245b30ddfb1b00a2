// pm.hip -- the particle-mesh force on gfx950.
//
// Replaces the reference's multipole-mesh far field (particle_compute_multipoles
// and particle_compute_expansions, particle.cpp:180-271, evaluated by L2P at
// kick.cpp:29) with cloud-in-cell mass assignment, keeping its pipeline:
//   deposit       fixed-point CIC (default) or TSC weights, uint64 atomics:
//                 deterministic, bit-identical to the CPU oracle.  Grids of 32
//                 nodes a side or more: particles binned by 16 x 16 x 32-node
//                 tile, each tile summed in LDS, then one global atomic per
//                 touched node (dep_*); smaller grids: one global atomic per
//                 particle and stencil node (mesh_deposit)
//   fft3d_density forward FP32 transform of the density (fft.hip; its first
//                 pass reads the uint64 grid directly)
//   green_mul     rho_hat * G (FP64 Green's table, octant-symmetric), then
//                 the spectral gradient; two real fields share one complex
//                 grid: A = Fx + i Fy, B = Fz + i pot
//   fft3d x2      inverse transforms (two instead of four)
//   interp        force and potential at the particles (same weights): in
//                 deposit-tile order from LDS-staged half tiles (interp_tile)
//                 when tiles hold enough particles, else per particle
//                 (mesh_interp)
// Grid layout: (i*ng + j)*ng + k with i along x; node c at position c/ng.
#include <cmath>
#include <vector>

#include "epilogue.h"
#include "internal.h"
#include "wave.h"

namespace ct2 {
namespace {

constexpr int CIC_BITS = 21;  // weight resolution per axis
constexpr int PM_BLOCK = 256;

struct PmState {
    int ng = 0, g = 0;
    double rs_cells = 0.0;
    int order_table = 0;              // assignment order the table was built for
    float gm_table = 0.0f;  // GM the table was built for
    unsigned long long* W = nullptr;  // density, 2^-32 particle units
    float2* A = nullptr;              // rho_hat -> Fx + i Fy
    float2* B = nullptr;              // Fz + i pot
    double* G = nullptr;              // Green's table on the octant, (ng/2+1)^3
    // tiled deposit: particles binned by tile of 16 x 16 x 32 nodes
    int ntile = 0;
    int* tcount = nullptr;  // [ntile]
    int* tstart = nullptr;  // [ntile + 1]
    size_t cap = 0;         // particles slot / order hold
    int* slot = nullptr;    // [cap] position within the tile
    int* order = nullptr;   // [cap] particle indices grouped by tile
} g_pm;
int g_assign = 1;  // mass assignment: 1 CIC, 2 TSC (kept across ct2_pm_init)

// Per-axis weights, 21-bit fixed point summing to 2^21, and the first node:
//   ORDER 1 (CIC): nodes lo, lo+1 with lo = x >> s, a1 = fraction >> (s-21);
//   ORDER 2 (TSC): nearest node c = (x + 2^(s-1)) >> s, signed offset D in
//     [-2^20, 2^20) (2^-21 cells); nodes c-1, c, c+1 get round((2^20-D)^2/2^22),
//     the rest, round((2^20+D)^2/2^22): the quadratic spline (1/2)(1/2 -+ d)^2.
template <int ORDER>
__device__ __forceinline__ void assign_axis(uint32_t x, int s, uint32_t& lo, uint64_t (&w)[ORDER + 1]) {
    if constexpr (ORDER == 1) {
        lo = x >> s;
        const uint32_t t = x & ((1u << s) - 1u);
        w[1] = (uint64_t)(t >> (s - CIC_BITS));
        w[0] = (1ull << CIC_BITS) - w[1];
    } else {
        const uint32_t half = 1u << (s - 1);
        const uint32_t xs = x + half;
        const int32_t d = (int32_t)(xs & ((1u << s) - 1u)) - (int32_t)half;
        const int64_t D = (int64_t)(d >> (s - CIC_BITS));
        const uint64_t a = (uint64_t)((1 << 20) - D), b = (uint64_t)((1 << 20) + D);
        w[0] = (a * a + (1ull << 21)) >> 22;
        w[2] = (b * b + (1ull << 21)) >> 22;
        w[1] = (1ull << CIC_BITS) - w[0] - w[2];
        lo = (xs >> s) - 1u;
    }
}

__global__ __launch_bounds__(PM_BLOCK) void fill_u64(unsigned long long* __restrict__ p, long n) {
    const long i = (long)blockIdx.x * PM_BLOCK + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

template <int ORDER>
__global__ __launch_bounds__(PM_BLOCK) void mesh_deposit(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                        const uint32_t* __restrict__ Z, int n, int g,
                                                        unsigned long long* __restrict__ W) {
    const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
    if (p >= n) return;
    const int s = 32 - g;
    const uint32_t ng = 1u << g, mask = ng - 1;
    uint32_t i0, j0, k0;
    uint64_t ax[ORDER + 1], ay[ORDER + 1], az[ORDER + 1];
    assign_axis<ORDER>(X[p], s, i0, ax);
    assign_axis<ORDER>(Y[p], s, j0, ay);
    assign_axis<ORDER>(Z[p], s, k0, az);
#pragma unroll
    for (int dx = 0; dx <= ORDER; ++dx)
#pragma unroll
        for (int dy = 0; dy <= ORDER; ++dy) {
            const uint64_t axy = ax[dx] * ay[dy];
            const size_t row = ((size_t)((i0 + dx) & mask) * ng + ((j0 + dy) & mask)) * ng;
#pragma unroll
            for (int dz = 0; dz <= ORDER; ++dz)
                atomicAdd(W + row + ((k0 + dz) & mask), (unsigned long long)((axy * az[dz]) >> 31));
        }
}

// Tiled deposit for grids of 32 nodes and more a side.  The grid is cut into
// tiles of 16 x 16 x 32 nodes; a particle belongs to the tile of its anchor
// node (CIC: its first node; TSC: its nearest node), so its stencil lies in the
// tile grown by one node on every side.  Particles are binned by tile
// (dep_count, tile_scan, dep_scatter); one workgroup per tile then adds its
// particles' weights into an LDS copy of the grown tile (uint64 LDS atomics)
// and adds each nonzero LDS node into the grid with one global atomic.  The
// weights are those of mesh_deposit and integer sums do not depend on order,
// so the grid is bit-identical to mesh_deposit's.
constexpr int TILE_BX = 4, TILE_BY = 4, TILE_BZ = 5;  // log2 tile size per axis
constexpr int TLX = (1 << TILE_BX) + 2, TLY = (1 << TILE_BY) + 2, TLZ = (1 << TILE_BZ) + 2;
constexpr int TL_NODES = TLX * TLY * TLZ;  // 11,016 nodes: 88,128 B of LDS
constexpr int DEP_THREADS = 1024;
constexpr int TILE_MIN_G = TILE_BZ;  // grids of at least 2^5 nodes a side

template <int ORDER>
__device__ __forceinline__ uint32_t anchor_node(uint32_t x, int s) {
    return ORDER == 1 ? x >> s : (uint32_t)(x + (1u << (s - 1))) >> s;
}

template <int ORDER>
__device__ __forceinline__ int tile_of(uint32_t x, uint32_t y, uint32_t z, int g) {
    const int s = 32 - g;
    const int tx = (int)(anchor_node<ORDER>(x, s) >> TILE_BX);
    const int ty = (int)(anchor_node<ORDER>(y, s) >> TILE_BY);
    const int tz = (int)(anchor_node<ORDER>(z, s) >> TILE_BZ);
    return (tx << (2 * g - TILE_BY - TILE_BZ)) + (ty << (g - TILE_BZ)) + tz;
}

__global__ __launch_bounds__(PM_BLOCK) void fill_i32(int* __restrict__ p, int n) {
    const int i = blockIdx.x * PM_BLOCK + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int ORDER>
__global__ __launch_bounds__(PM_BLOCK) void dep_count(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                      const uint32_t* __restrict__ Z, int n, int g,
                                                      int* __restrict__ tcount, int* __restrict__ slot) {
    const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
    const int t = p < n ? tile_of<ORDER>(X[p], Y[p], Z[p], g) : -1;
    const int k = run_increment(tcount, t);  // every lane of the wave takes part
    if (p < n) slot[p] = k;
}

// Exclusive scan of len counts (one workgroup): out[0..len], out[len] = total.
__global__ __launch_bounds__(DEP_THREADS) void tile_scan(const int* __restrict__ in, int len, int* __restrict__ out) {
    __shared__ int part[DEP_THREADS];
    const int t = threadIdx.x;
    const int per = (len + DEP_THREADS - 1) / DEP_THREADS;
    const int b = min(len, t * per), e = min(len, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < DEP_THREADS; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (t == DEP_THREADS - 1) out[len] = part[t];
}

template <int ORDER>
__global__ __launch_bounds__(PM_BLOCK) void dep_scatter(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                        const uint32_t* __restrict__ Z, int n, int g,
                                                        const int* __restrict__ tstart, const int* __restrict__ slot,
                                                        int* __restrict__ order) {
    const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
    if (p >= n) return;
    order[tstart[tile_of<ORDER>(X[p], Y[p], Z[p], g)] + slot[p]] = p;
}

template <int ORDER>
__global__ __launch_bounds__(DEP_THREADS) void dep_tile(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                        const uint32_t* __restrict__ Z, const int* __restrict__ order,
                                                        const int* __restrict__ tstart, int g,
                                                        unsigned long long* __restrict__ W) {
    __shared__ unsigned long long acc[TL_NODES];
    const int tile = blockIdx.x;
    const int b = tstart[tile], e = tstart[tile + 1];
    if (b == e) return;  // uniform across the workgroup
    const int tid = threadIdx.x;
    const int s = 32 - g;
    const uint32_t mask = (1u << g) - 1u;
    const uint32_t ox = (uint32_t)(tile >> (2 * g - TILE_BY - TILE_BZ)) << TILE_BX;
    const uint32_t oy = (uint32_t)((tile >> (g - TILE_BZ)) & ((1 << (g - TILE_BY)) - 1)) << TILE_BY;
    const uint32_t oz = (uint32_t)(tile & ((1 << (g - TILE_BZ)) - 1)) << TILE_BZ;
    for (int i = tid; i < TL_NODES; i += DEP_THREADS) acc[i] = 0ull;
    __syncthreads();
    // local index of the stencil's first node: anchor - origin + (CIC: 1, TSC: 0)
    constexpr uint32_t SHIFT = ORDER == 1 ? 1u : 0u;
    for (int k = b + tid; k < e; k += DEP_THREADS) {
        const int p = order[k];
        const uint32_t x = X[p], y = Y[p], z = Z[p];
        uint32_t lo;
        uint64_t ax[ORDER + 1], ay[ORDER + 1], az[ORDER + 1];
        assign_axis<ORDER>(x, s, lo, ax);
        assign_axis<ORDER>(y, s, lo, ay);
        assign_axis<ORDER>(z, s, lo, az);
        const int lx = (int)(anchor_node<ORDER>(x, s) - ox + SHIFT);
        const int ly = (int)(anchor_node<ORDER>(y, s) - oy + SHIFT);
        const int lz = (int)(anchor_node<ORDER>(z, s) - oz + SHIFT);
#pragma unroll
        for (int dx = 0; dx <= ORDER; ++dx)
#pragma unroll
            for (int dy = 0; dy <= ORDER; ++dy) {
                const uint64_t axy = ax[dx] * ay[dy];
                unsigned long long* row = acc + ((lx + dx) * TLY + (ly + dy)) * TLZ + lz;
#pragma unroll
                for (int dz = 0; dz <= ORDER; ++dz)
                    atomicAdd(row + dz, (unsigned long long)((axy * az[dz]) >> 31));
            }
    }
    __syncthreads();
    for (int i = tid; i < TL_NODES; i += DEP_THREADS) {
        const unsigned long long v = acc[i];
        if (!v) continue;
        const int lx = i / (TLY * TLZ), ly = (i / TLZ) % TLY, lz = i % TLZ;
        const size_t gx = (ox + lx - 1u) & mask, gy = (oy + ly - 1u) & mask, gz = (oz + lz - 1u) & mask;
        atomicAdd(W + ((((gx << g) + gy) << g) + gz), v);
    }
}

// One thread per k-cell.  phi_hat = G rho_hat (FP64); F_hat = -i k phi_hat
// with the Nyquist component of k zeroed; pot_hat = -phi_hat.
__global__ __launch_bounds__(PM_BLOCK) void green_mul(float2* __restrict__ A, float2* __restrict__ B,
                                                      const double* __restrict__ G, int g) {
    const long c = (long)blockIdx.x * PM_BLOCK + threadIdx.x;
    const int ng = 1 << g, h = ng / 2 + 1;
    if (c >= (long)ng * ng * ng) return;
    const int l = (int)(c & (ng - 1)), j = (int)((c >> g) & (ng - 1)), i = (int)(c >> (2 * g));
    const int mi = i < ng / 2 ? i : i - ng, mj = j < ng / 2 ? j : j - ng, ml = l < ng / 2 ? l : l - ng;
    const double gv = G[((size_t)abs(mi) * h + abs(mj)) * h + abs(ml)];
    const float2 r = A[c];
    const double pr = gv * (double)r.x, pi = gv * (double)r.y;
    const double twopi = 6.283185307179586476925;
    const double kx = i == ng / 2 ? 0.0 : twopi * mi;
    const double ky = j == ng / 2 ? 0.0 : twopi * mj;
    const double kz = l == ng / 2 ? 0.0 : twopi * ml;
    // A = F_hat_x + i F_hat_y ; B = F_hat_z + i pot_hat
    A[c] = make_float2((float)(kx * pi + ky * pr), (float)(-kx * pr + ky * pi));
    B[c] = make_float2((float)((kz + 1.0) * pi), (float)(-(kz + 1.0) * pr));
}

template <int ORDER>
__global__ __launch_bounds__(PM_BLOCK) void mesh_interp(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                       const uint32_t* __restrict__ Z, int n, int g,
                                                       const float2* __restrict__ A, const float2* __restrict__ B,
                                                       float* __restrict__ fx, float* __restrict__ fy,
                                                       float* __restrict__ fz, float* __restrict__ pot) {
    const int p = blockIdx.x * PM_BLOCK + threadIdx.x;
    if (p >= n) return;
    const int s = 32 - g;
    const uint32_t ng = 1u << g, mask = ng - 1;
    uint32_t i0, j0, k0;
    uint64_t ax[ORDER + 1], ay[ORDER + 1], az[ORDER + 1];
    assign_axis<ORDER>(X[p], s, i0, ax);
    assign_axis<ORDER>(Y[p], s, j0, ay);
    assign_axis<ORDER>(Z[p], s, k0, az);
    const float q = 0x1p-21f;
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
#pragma unroll
    for (int dx = 0; dx <= ORDER; ++dx)
#pragma unroll
        for (int dy = 0; dy <= ORDER; ++dy) {
            const float wxy = ((float)ax[dx] * q) * ((float)ay[dy] * q);
            const size_t row = ((size_t)((i0 + dx) & mask) * ng + ((j0 + dy) & mask)) * ng;
#pragma unroll
            for (int dz = 0; dz <= ORDER; ++dz) {
                const size_t cc = row + ((k0 + dz) & mask);
                const float w = wxy * ((float)az[dz] * q);
                const float2 a = A[cc], b = B[cc];
                acc0 = fmaf(w, a.x, acc0);
                acc1 = fmaf(w, a.y, acc1);
                acc2 = fmaf(w, b.x, acc2);
                acc3 = fmaf(w, b.y, acc3);
            }
        }
    fx[p] = acc0;
    fy[p] = acc1;
    fz[p] = acc2;
    pot[p] = acc3;
}

// Interpolation in tile order, for grids with enough particles per tile: one
// workgroup per deposit tile (16 x 16 x 32 nodes) stages the force and
// potential of one 16-node half of the tile along z, grown by one node on
// every side (18^3 nodes of (Fx, Fy, Fz, pot), 93 KB), in LDS and
// interpolates the particles whose anchor node lies in that half; then the
// other half.  The same weights and summation order as mesh_interp, so the
// results are identical; each node is read from L2 once per tile half instead
// of once per particle and stencil node.
constexpr int IH = 1 << (TILE_BZ - 1);  // z nodes per half tile
constexpr int ILX = (1 << TILE_BX) + 2, ILY = (1 << TILE_BY) + 2, ILZ = IH + 2;
template <int ORDER>
__global__ __launch_bounds__(DEP_THREADS) void interp_tile(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                           const uint32_t* __restrict__ Z, const int* __restrict__ order,
                                                           const int* __restrict__ tstart, int g,
                                                           const float2* __restrict__ A, const float2* __restrict__ B,
                                                           float* __restrict__ fx, float* __restrict__ fy,
                                                           float* __restrict__ fz, float* __restrict__ pot) {
    __shared__ float4 nd[ILX * ILY * ILZ];
    const int tile = blockIdx.x;
    const int b = tstart[tile], e = tstart[tile + 1];
    if (b == e) return;  // uniform across the workgroup
    const int tid = threadIdx.x;
    const int s = 32 - g;
    const uint32_t mask = (1u << g) - 1u;
    const uint32_t ox = (uint32_t)(tile >> (2 * g - TILE_BY - TILE_BZ)) << TILE_BX;
    const uint32_t oy = (uint32_t)((tile >> (g - TILE_BZ)) & ((1 << (g - TILE_BY)) - 1)) << TILE_BY;
    const uint32_t oz = (uint32_t)(tile & ((1 << (g - TILE_BZ)) - 1)) << TILE_BZ;
    constexpr uint32_t SHIFT = ORDER == 1 ? 1u : 0u;
    const float q = 0x1p-21f;
    for (int h = 0; h < 2; ++h) {
        const uint32_t hz = oz + (uint32_t)(h * IH);
        __syncthreads();  // the previous half's reads are done
        for (int i = tid; i < ILX * ILY * ILZ; i += DEP_THREADS) {
            const int lx = i / (ILY * ILZ), ly = (i / ILZ) % ILY, lz = i % ILZ;
            const size_t gx = (ox + lx - 1u) & mask, gy = (oy + ly - 1u) & mask, gz = (hz + lz - 1u) & mask;
            const size_t cc = (((gx << g) + gy) << g) + gz;
            const float2 a = A[cc], bb = B[cc];
            nd[i] = make_float4(a.x, a.y, bb.x, bb.y);
        }
        __syncthreads();
        for (int k = b + tid; k < e; k += DEP_THREADS) {
            const int p = order[k];
            const uint32_t z = Z[p];
            const uint32_t lzs = anchor_node<ORDER>(z, s) - hz;  // anchor z within this half: [0, IH)
            if (lzs >= (uint32_t)IH) continue;
            const uint32_t x = X[p], y = Y[p];
            uint32_t lo;
            uint64_t ax[ORDER + 1], ay[ORDER + 1], az[ORDER + 1];
            assign_axis<ORDER>(x, s, lo, ax);
            assign_axis<ORDER>(y, s, lo, ay);
            assign_axis<ORDER>(z, s, lo, az);
            const int lx = (int)(anchor_node<ORDER>(x, s) - ox + SHIFT);
            const int ly = (int)(anchor_node<ORDER>(y, s) - oy + SHIFT);
            const int lz = (int)(lzs + SHIFT);
            float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
#pragma unroll
            for (int dx = 0; dx <= ORDER; ++dx)
#pragma unroll
                for (int dy = 0; dy <= ORDER; ++dy) {
                    const float wxy = ((float)ax[dx] * q) * ((float)ay[dy] * q);
                    const float4* row = nd + ((lx + dx) * ILY + (ly + dy)) * ILZ + lz;
#pragma unroll
                    for (int dz = 0; dz <= ORDER; ++dz) {
                        const float w = wxy * ((float)az[dz] * q);
                        const float4 v = row[dz];
                        acc0 = fmaf(w, v.x, acc0);
                        acc1 = fmaf(w, v.y, acc1);
                        acc2 = fmaf(w, v.z, acc2);
                        acc3 = fmaf(w, v.w, acc3);
                    }
                }
            fx[p] = acc0;
            fy[p] = acc1;
            fz[p] = acc2;
            pot[p] = acc3;
        }
    }
}

void free_pm() {
    dev_free(g_pm.W);
    dev_free(g_pm.A);
    dev_free(g_pm.B);
    dev_free(g_pm.G);
    dev_free(g_pm.tcount);
    dev_free(g_pm.tstart);
    dev_free(g_pm.slot);
    dev_free(g_pm.order);
    g_pm = PmState{};
}

// Per-particle buffers of the tiled deposit, grown to the container size.
int ensure_deposit(size_t n) {
    if (n <= g_pm.cap) return CT2_OK;
    dev_free(g_pm.slot);
    dev_free(g_pm.order);
    g_pm.slot = g_pm.order = nullptr;
    g_pm.cap = 0;
    CT2_TRY(dev_alloc((void**)&g_pm.slot, n * sizeof(int), "PM deposit slots"));
    CT2_TRY(dev_alloc((void**)&g_pm.order, n * sizeof(int), "PM deposit order"));
    g_pm.cap = n;
    return CT2_OK;
}

// Density of the container's particles into g_pm.W (zeroed here).
template <int ORDER>
int deposit(const Context& c) {
    const int n = (int)c.n, g = g_pm.g;
    const long n3 = 1L << (3 * g);
    const unsigned gb = (unsigned)((n3 + PM_BLOCK - 1) / PM_BLOCK), pb = (unsigned)((n + PM_BLOCK - 1) / PM_BLOCK);
    fill_u64<<<gb, PM_BLOCK, 0, c.stream>>>(g_pm.W, n3);
    CT2_LAUNCH_CHECK();
    if (g < TILE_MIN_G) {
        mesh_deposit<ORDER><<<pb, PM_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, n, g, g_pm.W);
        CT2_LAUNCH_CHECK();
        return CT2_OK;
    }
    CT2_TRY(ensure_deposit(c.n));
    const int nt = g_pm.ntile;
    fill_i32<<<(nt + PM_BLOCK - 1) / PM_BLOCK, PM_BLOCK, 0, c.stream>>>(g_pm.tcount, nt);
    CT2_LAUNCH_CHECK();
    dep_count<ORDER><<<pb, PM_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, n, g, g_pm.tcount, g_pm.slot);
    CT2_LAUNCH_CHECK();
    tile_scan<<<1, DEP_THREADS, 0, c.stream>>>(g_pm.tcount, nt, g_pm.tstart);
    CT2_LAUNCH_CHECK();
    dep_scatter<ORDER><<<pb, PM_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, n, g, g_pm.tstart, g_pm.slot, g_pm.order);
    CT2_LAUNCH_CHECK();
    dep_tile<ORDER><<<nt, DEP_THREADS, 0, c.stream>>>(c.x, c.y, c.z, g_pm.order, g_pm.tstart, g, g_pm.W);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

// FP64 Green's table on the octant |m_d| <= ng/2 (host, libm):
//   G(m) = -4 pi GM exp(-k^2 rs^2) / (k^2 W(m)^2), W = prod_d sinc^(order+1)(pi m_d/ng),
//   G(0) = 0.
std::vector<double> green_table(int ng, double gm, double rs_cells, int order) {
    const int h = ng / 2 + 1;
    const double rs = rs_cells / ng;
    std::vector<double> t((size_t)h * h * h);
    std::vector<double> sinc2(h), k2(h);
    for (int m = 0; m < h; ++m) {
        const double u = M_PI * m / ng;
        const double sc = m ? std::sin(u) / u : 1.0;
        sinc2[m] = order == 1 ? sc * sc : sc * sc * sc;
        k2[m] = (2.0 * M_PI * m) * (2.0 * M_PI * m);
    }
    for (int a = 0; a < h; ++a)
        for (int b = 0; b < h; ++b)
            for (int c = 0; c < h; ++c) {
                double v = 0.0;
                if (a || b || c) {
                    const double w = sinc2[a] * sinc2[b] * sinc2[c];
                    const double kk = k2[a] + k2[b] + k2[c];
                    v = -4.0 * M_PI * gm * std::exp(-kk * rs * rs) / (kk * w * w);
                }
                t[((size_t)a * h + b) * h + c] = v;
            }
    return t;
}

int ensure_table() {
    Context& c = ctx();
    if (g_pm.gm_table == c.opts.GM && g_pm.order_table == g_assign) return CT2_OK;
    const std::vector<double> t = green_table(g_pm.ng, (double)c.opts.GM, g_pm.rs_cells, g_assign);
    CT2_HIP(hipMemcpyAsync(g_pm.G, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    g_pm.gm_table = c.opts.GM;
    g_pm.order_table = g_assign;
    return CT2_OK;
}

}  // namespace

int pm_require() {
    CT2_TRY(require_particles());
    if (!g_pm.ng) return set_error(CT2_ERR_STATE, "ct2_pm_init has not been called");
    return CT2_OK;
}

// The PM force and potential of the container into the context force arrays.
int pm_forces() {
    CT2_TRY(pm_require());
    CT2_TRY(ensure_table());
    Context& c = ctx();
    const int n = (int)c.n, g = g_pm.g;
    const long n3 = 1L << (3 * g);
    const unsigned gb = (unsigned)((n3 + PM_BLOCK - 1) / PM_BLOCK), pb = (unsigned)((n + PM_BLOCK - 1) / PM_BLOCK);
    CT2_TRY(g_assign == 1 ? deposit<1>(c) : deposit<2>(c));
    CT2_TRY(fft3d_density(g_pm.W, g_pm.A, g_pm.ng));  // density (real) -> rho_hat
    green_mul<<<gb, PM_BLOCK, 0, c.stream>>>(g_pm.A, g_pm.B, g_pm.G, g);
    CT2_LAUNCH_CHECK();
    CT2_TRY(fft3d(g_pm.A, g_pm.ng, -1, false));
    CT2_TRY(fft3d(g_pm.B, g_pm.ng, -1, false));
    // tiles hold n / ntile particles on average: stage a tile's nodes once when
    // that beats one L2 read per particle and stencil node.  Measured: faster
    // on 256^3 grids (0.19 -> 0.16 ms, TSC, 2^21 particles), slower on 512^3
    // (1.0 -> 1.4 ms, 2^24 particles), where the per-particle gathers of
    // particles in lattice order already hit L2 and the staged workgroup (one
    // per CU) cannot overlap its loads with its arithmetic.
    const bool tiled = g >= TILE_MIN_G && g <= 8 &&
                       (long)n * (g_assign == 1 ? 8 : 27) > 2L * 2 * ILX * ILY * ILZ * g_pm.ntile;
    if (tiled && g_assign == 1)
        interp_tile<1><<<g_pm.ntile, DEP_THREADS, 0, c.stream>>>(c.x, c.y, c.z, g_pm.order, g_pm.tstart, g, g_pm.A,
                                                                 g_pm.B, c.fx, c.fy, c.fz, c.pot);
    else if (tiled)
        interp_tile<2><<<g_pm.ntile, DEP_THREADS, 0, c.stream>>>(c.x, c.y, c.z, g_pm.order, g_pm.tstart, g, g_pm.A,
                                                                 g_pm.B, c.fx, c.fy, c.fz, c.pot);
    else if (g_assign == 1)
        mesh_interp<1><<<pb, PM_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, n, g, g_pm.A, g_pm.B, c.fx, c.fy, c.fz, c.pot);
    else
        mesh_interp<2><<<pb, PM_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, n, g, g_pm.A, g_pm.B, c.fx, c.fy, c.fz, c.pot);
    CT2_LAUNCH_CHECK();
    if (near_enabled()) CT2_TRY(near_forces(g_pm.rs_cells / g_pm.ng));
    c.forces_valid = true;
    return CT2_OK;
}

// Device grids for the learned correction's feature gather.
void pm_config(int* ng, double* rs_cells, int* order, double* rcut_rs) {
    *ng = g_pm.ng;
    *rs_cells = g_pm.rs_cells;
    *order = g_assign;
    *rcut_rs = near_rcut_rs();
}

void pm_grids(int* g, const unsigned long long** W, const float2** A, const float2** B) {
    *g = g_pm.g;
    *W = g_pm.W;
    *A = g_pm.A;
    *B = g_pm.B;
}

void release_pm() { free_pm(); }

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_pm_init(int ng, double rs_cells) {
    CT2_TRY(require_init());
    if (!is_pow2(ng) || ng < 8 || ng > 1024)
        return set_error(CT2_ERR_INVALID, "PM grid %d is not a power of two in [8, 1024]", ng);
    if (!(rs_cells >= 0.0) || rs_cells > 64.0) return set_error(CT2_ERR_INVALID, "bad split radius %g", rs_cells);
    if (g_pm.ng == ng && g_pm.rs_cells == rs_cells) return CT2_OK;
    free_pm();
    const size_t n3 = (size_t)ng * ng * ng, h = ng / 2 + 1;
    int rc;
    if ((rc = dev_alloc((void**)&g_pm.W, n3 * sizeof(unsigned long long), "PM density")) ||
        (rc = dev_alloc((void**)&g_pm.A, n3 * sizeof(float2), "PM grid A")) ||
        (rc = dev_alloc((void**)&g_pm.B, n3 * sizeof(float2), "PM grid B")) ||
        (rc = dev_alloc((void**)&g_pm.G, h * h * h * sizeof(double), "PM Green's table"))) {
        free_pm();
        return rc;
    }
    g_pm.ng = ng;
    g_pm.g = ilog2(ng);
    if (g_pm.g >= TILE_MIN_G) {
        g_pm.ntile = 1 << (3 * g_pm.g - TILE_BX - TILE_BY - TILE_BZ);
        if ((rc = dev_alloc((void**)&g_pm.tcount, (size_t)g_pm.ntile * sizeof(int), "PM tile counts")) ||
            (rc = dev_alloc((void**)&g_pm.tstart, (size_t)(g_pm.ntile + 1) * sizeof(int), "PM tile starts"))) {
            free_pm();
            return rc;
        }
    }
    g_pm.rs_cells = rs_cells;
    g_pm.gm_table = 0.0f;
    return CT2_OK;
}

extern "C" int ct2_pm_near(double rcut_rs) {
    CT2_TRY(require_init());
    if (!g_pm.ng) return set_error(CT2_ERR_STATE, "ct2_pm_init has not been called");
    if (rcut_rs > 0.0 && !(g_pm.rs_cells > 0.0))
        return set_error(CT2_ERR_INVALID, "the near field needs a PM split radius rs > 0");
    if (rcut_rs > 0.0 && rcut_rs * g_pm.rs_cells / g_pm.ng > 1.0 / 3.0)
        return set_error(CT2_ERR_INVALID, "near-field cutoff %g rs exceeds a third of the box", rcut_rs);
    return near_set(rcut_rs);
}

extern "C" int ct2_pm_near_cells(void) { return near_cells(); }

extern "C" int ct2_pm_assignment(int order) {
    CT2_TRY(require_init());
    if (order != 1 && order != 2) return set_error(CT2_ERR_INVALID, "assignment order %d is not 1 (CIC) or 2 (TSC)", order);
    g_assign = order;
    return CT2_OK;
}

extern "C" int ct2_pm_forces(void) { return pm_forces(); }

extern "C" int ct2_kick_pm(const ct2_kick_params* p, unsigned flags, int* max_rung) {
    CT2_TRY(pm_require());
    CT2_TRY(check_kick_params(p));
    if (flags & ~(unsigned)CT2_KICK_STORE) return set_error(CT2_ERR_INVALID, "unknown flags 0x%x", flags);
    CT2_TRY(pm_forces());
    ctx().last_path = near_enabled() ? CT2_PATH_PM_NEAR : CT2_PATH_PM;
    return run_epilogue(p, max_rung);
}

extern "C" int ct2_pm_density_download(uint64_t* host) {
    CT2_TRY(require_init());
    if (!g_pm.ng) return set_error(CT2_ERR_STATE, "ct2_pm_init has not been called");
    if (!host) return set_error(CT2_ERR_INVALID, "null output");
    Context& c = ctx();
    const size_t n3 = (size_t)g_pm.ng * g_pm.ng * g_pm.ng;
    CT2_HIP(hipMemcpyAsync(host, g_pm.W, n3 * 8, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

extern "C" int ct2_pm_grid_download(int which, float* host) {
    CT2_TRY(require_init());
    if (!g_pm.ng) return set_error(CT2_ERR_STATE, "ct2_pm_init has not been called");
    if (!host || (which != 0 && which != 1)) return set_error(CT2_ERR_INVALID, "bad grid request");
    Context& c = ctx();
    const size_t n3 = (size_t)g_pm.ng * g_pm.ng * g_pm.ng;
    CT2_HIP(hipMemcpyAsync(host, which ? g_pm.B : g_pm.A, n3 * 8, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}
