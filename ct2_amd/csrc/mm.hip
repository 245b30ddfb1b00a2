// mm.hip -- ct2's multipole-mesh force (particle.cpp:116-271, kick.cpp:11-29,
// 105-125) on gfx950.
//
// ct2's kick() sums, for the particles of each cell, direct pairs with the
// cells of its neighbour list and a local expansion of everything else about
// the cell centre; the expansions are a convolution over the cell grid done
// with one FFT per multipole component.  The reference's P2M, MG2L and L2P
// come from sfmm (spherical harmonics, PORDER = 8), which is not in the tree;
// this module uses Cartesian Taylor expansions of total order P <= 8.
//
//   cells      nc^3, nc a power of two in [8, 256]; lengths below in cells;
//   P2M        M_l(c) = sum_{j in c} (-b_j)^l / l!   (b = x - centre, unit mass);
//   table      D_n(rho) = d^n G_far(rho) at the cell offsets rho, G_far the
//              periodic Green's function minus the nearest image 1/|rho| on the
//              near set |rho|^2 <= 10 (particle.cpp:136), transformed once:
//              Dt[n][q] = D^_n(q) / nc^3 (real for even |n|, i x real for odd);
//   MG2L       L^_k(q) = sum_{|k|+|l|<=P} M^_l(q) D^_{k+l}(q), per wave vector;
//   L2P        pot = sum_k a^k/k! L_k,  F_q = sum_k a^{k-e_q}/(k-e_q)! L_k;
//   near       softened direct pairs with the particles of the near set's cells.
//
// Box units: force GM (near + nc^2 far), potential GM (near + nc far).
//
// The table: G = G_short + G_long (Ewald split beta = 0.5 per cell).
// G_short = sum_m erfc(beta |rho + nc m|)/|rho + nc m| is expanded on the host in
// FP64: for a radial f(s), s = |rho + h|^2, the Taylor coefficients in h are
// sum_j f^(j)(s0)/j! [h^n] (2 rho.h + h.h)^j with f^(j) = (-1/2)^j B_j and
// B_0 = erfc(beta r)/r, B_j = ((2j-1) B_{j-1} + (2 beta^2)^j e^{-beta^2 r^2}/(beta sqrt pi))/r^2
// (Smith 1998); a near offset's own image is -erf(beta r)/r, a power series in s.
// Its samples are transformed with the FP64 3-D FFT.  G_long's sampled
// transform is analytic: sum over aliases kappa = q/nc + m of
// (2 pi i kappa)^n e^{-pi^2 kappa^2/beta^2} / (pi kappa^2), and -pi/beta^2 at 0.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "epilogue.h"
#include "internal.h"

namespace ct2 {
namespace {

constexpr int MM_PMAX = 8;
constexpr int MM_KMAX = 165;       // components at P = 8
constexpr int MM_NEAR2 = 10;       // |cell offset|^2 of the direct set (particle.cpp:136)
constexpr double MM_BETA = 0.5;    // Ewald split per cell
constexpr int MM_QB = 16;          // wave vectors per MG2L workgroup
constexpr int MM_SLOTS = 16;       // MG2L work lists per wave vector (256 threads)
constexpr int MM_SCHED_MAX = 384;  // entries per list
// list stride in words: rows of whole 16-byte groups, the lists of a wave 16 bytes apart modulo 128
constexpr int mg2l_stride(int slen) { return (slen + 31) / 32 * 32 + 4; }
constexpr int MM_COLS_MAX = 48;    // (dx, dy) columns of the near set (37 at |d|^2 <= 10)
constexpr int MM_WAVES = 4;        // waves per near-field workgroup

// MG2L list entry: flags | l << 7 | j << 15 | n << 22, so that l * 128 (the
// byte offset of row l of the LDS multipoles) is a mask of the word and n * 64
// (row n of the LDS table) a shift.  Entries come in groups of four products
// M^_l D^_n of one output component and one parity of |n| (D^_n is real for
// even |n|, i times real for odd); the group's last entry carries its flags.
// The components 2j and 2j+1 share a list: at the last group of 2j (save) the
// sum is set aside, at the last of 2j+1 (flush) L^_2j + i L^_2j+1 goes out.
// Pads: l = 0 and n = K, a zero row.
constexpr uint32_t G_ODD = 1u, G_SAVE = 2u, G_FLUSH = 4u;
__host__ __device__ constexpr uint32_t mg2l_entry(int l, int n, int j, uint32_t flags) {
    return flags | (uint32_t)l << 7 | (uint32_t)j << 15 | (uint32_t)n << 22;
}

__constant__ int c_e[MM_KMAX][3];         // exponents of component c
__constant__ float c_ifact[MM_KMAX];      // 1 / (a! b! c!)
__constant__ int c_cols[MM_COLS_MAX][3];  // near set by column: dx, dy, half-height w in z

struct MMIndex {
    int P = 0, K = 0;
    int e[MM_KMAX][3];
    int id[MM_PMAX + 1][MM_PMAX + 1][MM_PMAX + 1];
    double fact[MM_KMAX];
    std::vector<int> pa, pb, pc;  // products within degree P: (i, j) -> i + j
    void init(int p) {
        P = p;
        K = (P + 1) * (P + 2) * (P + 3) / 6;
        std::memset(id, -1, sizeof id);
        int c = 0;
        for (int d = 0; d <= P; ++d)
            for (int a = d; a >= 0; --a)
                for (int b = d - a; b >= 0; --b) {
                    e[c][0] = a;
                    e[c][1] = b;
                    e[c][2] = d - a - b;
                    id[a][b][d - a - b] = c;
                    fact[c] = std::tgamma(a + 1.0) * std::tgamma(b + 1.0) * std::tgamma(d - a - b + 1.0);
                    ++c;
                }
        pa.clear();
        pb.clear();
        pc.clear();
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) {
                const int a = e[i][0] + e[j][0], b = e[i][1] + e[j][1], d = e[i][2] + e[j][2];
                if (a + b + d > P) continue;
                pa.push_back(i);
                pb.push_back(j);
                pc.push_back(id[a][b][d]);
            }
    }
    int deg(int c) const { return e[c][0] + e[c][1] + e[c][2]; }
};

struct MMState {
    int nc = 0, P = 0, K = 0, nz = 0;  // nz = complex grids (K/2 rounded up)
    int slen = 0, ncols = 0;
    double* Dt = nullptr;  // [K][nc^3]
    size_t Dt_bytes = 0;
    float* Dt32 = nullptr;  // the same rounded to FP32, the rows MG2L reads
    size_t Dt32_bytes = 0;
    float2* Zm = nullptr;  // multipole grids [nz][nc^3], then their transforms
    float2* Zl = nullptr;  // local-expansion grids
    size_t Z_bytes = 0;
    uint32_t* keys = nullptr;  // sort keys / values in and out (4 n)
    size_t keys_bytes = 0;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    uint4* Ps = nullptr;  // sorted (x, y, z, index)
    size_t Ps_bytes = 0;
    int* start = nullptr;  // cell starts (nc^3 + 1)
    size_t start_bytes = 0;
    int* items = nullptr;  // near-field work items (row of cells, first sink, sinks) and their count
    size_t items_bytes = 0;
    int* nitems = nullptr;
    int* cnt = nullptr;  // items per row, then the rows' first item (exclusive sum)
    size_t cnt_bytes = 0;
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    double2* tabz = nullptr;  // one FP64 complex grid while the table is built
    size_t tabz_bytes = 0;
    uint32_t* sched = nullptr;  // the MG2L lists [MM_SLOTS][mg2l_stride(slen)]
    size_t sched_bytes = 0;
};
MMState g_mm;

// ------------------------------------------------------------ the table (host, FP64)

// s-derivatives of erfc(beta sqrt s)/sqrt s at s = r^2
void erfc_sderiv(int P, double r, double* fs) {
    const double b = MM_BETA, r2 = r * r, ex = std::exp(-b * b * r2) / (b * std::sqrt(M_PI));
    double B = std::erfc(b * r) / r, tb = 1.0, sgn = 1.0;
    fs[0] = B;
    for (int j = 1; j <= P; ++j) {
        tb *= 2.0 * b * b;
        B = ((2 * j - 1) * B + tb * ex) / r2;
        sgn *= -0.5;
        fs[j] = sgn * B;
    }
}

// s-derivatives of -erf(beta sqrt s)/sqrt s = -(2 beta/sqrt pi) sum_i (-beta^2 s)^i/(i! (2i+1))
void erf_sderiv(int P, double s, double* fs) {
    const double b2 = MM_BETA * MM_BETA, c0 = -2.0 * MM_BETA / std::sqrt(M_PI);
    for (int j = 0; j <= P; ++j) {
        double sum = 0.0, term = std::pow(-b2, j);
        for (int i = j; i < j + 400; ++i) {
            const double t = term / (2 * i + 1);
            sum += t;
            if (std::fabs(t) < 1e-19 * std::fabs(sum) && i > j + 4) break;
            term *= -b2 * s / (double)(i - j + 1);
        }
        fs[j] = c0 * sum;
    }
}

// T += Taylor coefficients (d^n f / n!) at rho of f(|rho + h|^2), fs = f^(j)(|rho|^2)
void radial_taylor_add(const MMIndex& I, const double* rho, const double* fs, double* T) {
    double delta[MM_KMAX] = {}, pw[MM_KMAX] = {}, nx[MM_KMAX];
    for (int q = 0; q < 3; ++q) {
        int a[3] = {0, 0, 0};
        a[q] = 1;
        delta[I.id[a[0]][a[1]][a[2]]] = 2.0 * rho[q];
        a[q] = 2;
        if (I.P >= 2) delta[I.id[a[0]][a[1]][a[2]]] = 1.0;
    }
    pw[0] = 1.0;
    T[0] += fs[0];
    double jf = 1.0;
    const size_t np = I.pa.size();
    for (int j = 1; j <= I.P; ++j) {
        std::fill(nx, nx + I.K, 0.0);
        for (size_t p = 0; p < np; ++p) nx[I.pc[p]] += pw[I.pa[p]] * delta[I.pb[p]];
        std::copy(nx, nx + I.K, pw);
        jf *= j;
        const double c = fs[j] / jf;
        for (int k = 0; k < I.K; ++k) T[k] += c * pw[k];
    }
}

int wrap_offset(int i, int nc) { return i < nc / 2 ? i : i - nc; }

// D_n of G_short (less the near set's nearest image) at every cell, [K][nc^3];
// cells beyond the cutoff stay zero.  Split over host threads.
void short_samples(const MMIndex& I, int nc, std::vector<double>& S) {
    const size_t n3 = (size_t)nc * nc * nc;
    S.assign(n3 * I.K, 0.0);
    const double rc = 8.0 / MM_BETA;  // erfc(beta r)/r and its derivatives < 1e-16 of their size beyond
    const int mr = (int)std::ceil(rc / nc) + 1;
    auto work = [&](size_t c0, size_t c1) {
        double fs[MM_PMAX + 1], T[MM_KMAX];
        for (size_t cell = c0; cell < c1; ++cell) {
            const int o[3] = {wrap_offset((int)(cell / ((size_t)nc * nc)), nc), wrap_offset((int)(cell / nc % nc), nc),
                              wrap_offset((int)(cell % nc), nc)};
            const bool near = o[0] * o[0] + o[1] * o[1] + o[2] * o[2] <= MM_NEAR2;
            bool any = false;
            std::fill(T, T + I.K, 0.0);
            for (int m0 = -mr; m0 <= mr; ++m0)
                for (int m1 = -mr; m1 <= mr; ++m1)
                    for (int m2 = -mr; m2 <= mr; ++m2) {
                        const double r[3] = {(double)o[0] + nc * m0, (double)o[1] + nc * m1, (double)o[2] + nc * m2};
                        const double s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                        if (near && !m0 && !m1 && !m2) {
                            erf_sderiv(I.P, s, fs);
                        } else {
                            if (s >= rc * rc) continue;
                            erfc_sderiv(I.P, std::sqrt(s), fs);
                        }
                        radial_taylor_add(I, r, fs, T);
                        any = true;
                    }
            if (any)
                for (int k = 0; k < I.K; ++k) S[(size_t)k * n3 + cell] = T[k] * I.fact[k];
        }
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, n3 * t / nt, n3 * (t + 1) / nt);
    for (auto& t : th) t.join();
}

// Dt[c][q] = (parity part of the transformed short samples + the analytic
// long-range transform) / nc^3 for one component c of exponents (a, b, d);
// Dt32 the same rounded to FP32.
__global__ __launch_bounds__(256) void mm_table_finish(const double2* __restrict__ Z, int nc, int a, int b, int d,
                                                       double* __restrict__ Dt, float* __restrict__ Dt32) {
    const size_t n3 = (size_t)nc * nc * nc;
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n3) return;
    const int deg = a + b + d;
    const int q0 = (int)(q / ((size_t)nc * nc)), q1 = (int)(q / nc % nc), q2 = (int)(q % nc);
    const double k0 = (double)(q0 < nc / 2 ? q0 : q0 - nc) / nc, k1 = (double)(q1 < nc / 2 ? q1 : q1 - nc) / nc,
                 k2 = (double)(q2 < nc / 2 ? q2 : q2 - nc) / nc;
    const double b2 = MM_BETA * MM_BETA, tp = 2.0 * M_PI;
    double lr = 0.0;
    for (int m0 = -1; m0 <= 1; ++m0)
        for (int m1 = -1; m1 <= 1; ++m1)
            for (int m2 = -1; m2 <= 1; ++m2) {
                const double x = k0 + m0, y = k1 + m1, z = k2 + m2, kk = x * x + y * y + z * z;
                if (kk == 0.0) continue;
                double mono = exp(-M_PI * M_PI * kk / b2) / (M_PI * kk);
                for (int i = 0; i < a; ++i) mono *= tp * x;
                for (int i = 0; i < b; ++i) mono *= tp * y;
                for (int i = 0; i < d; ++i) mono *= tp * z;
                lr += mono;
            }
    const double s = ((deg >> 1) & 1) ? -1.0 : 1.0;  // i^deg = s or i s
    double v = (deg & 1) ? Z[q].y + s * lr : Z[q].x + s * lr;
    if (deg == 0 && q == 0) v -= M_PI / b2;
    Dt[q] = v / (double)n3;
    Dt32[q] = (float)(v / (double)n3);
}

// ------------------------------------------------------------ binning

__global__ __launch_bounds__(256) void mm_keys(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                               const uint32_t* __restrict__ Z, int n, int lg, uint32_t* __restrict__ key,
                                               uint32_t* __restrict__ idx) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int sh = 32 - lg;
    key[i] = (((X[i] >> sh) << lg | (Y[i] >> sh)) << lg) | (Z[i] >> sh);
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void mm_gather(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                 const uint32_t* __restrict__ Z, const uint32_t* __restrict__ idx, int n,
                                                 uint4* __restrict__ P) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = idx[s];
    P[s] = make_uint4(X[i], Y[i], Z[i], i);
}

// start[c] = first sorted slot of cell c (lower bound of key c), start[n3] = n
__global__ __launch_bounds__(256) void mm_bounds(const uint32_t* __restrict__ key, int n, int n3,
                                                 int* __restrict__ start) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c > n3) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] < (uint32_t)c)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[c] = lo;
}

// Work items of the near-set kernel: the sorted particles of each row of
// cells (fixed x and y cell) cut into runs of at most 64 that span fewer than
// zspan z cells.  One thread per row walks the row's cells; pass 0 counts the
// row's items into cnt, pass 1 writes (row, first slot, sinks) at first[row].
__global__ __launch_bounds__(256) void mm_row_items(const int* __restrict__ start, int lg, int zspan, int pass,
                                                    int* __restrict__ cnt, const int* __restrict__ first,
                                                    int* __restrict__ items, int* __restrict__ nitems) {
    const int nc = 1 << lg, rows = nc * nc;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    int it = pass ? first[r] : 0;
    int f = start[r * nc], c = 0, z0 = 0;
    for (int z = 0; z < nc; ++z) {
        int left = start[r * nc + z + 1] - start[r * nc + z];
        while (left > 0) {
            if (c == 64 || (c > 0 && z - z0 >= zspan)) {
                if (pass) {
                    items[3 * it] = r;
                    items[3 * it + 1] = f;
                    items[3 * it + 2] = c;
                }
                ++it;
                f += c;
                c = 0;
            }
            if (c == 0) z0 = z;
            const int t = min(left, 64 - c);
            c += t;
            left -= t;
        }
    }
    if (c > 0) {
        if (pass) {
            items[3 * it] = r;
            items[3 * it + 1] = f;
            items[3 * it + 2] = c;
        }
        ++it;
    }
    if (!pass)
        cnt[r] = it;
    else if (r == rows - 1)
        *nitems = it;
}

// ------------------------------------------------------------ P2M

// MM_P2M_CELLS consecutive cells per workgroup, MM_P2M_CELLS / 4 of them per
// wave in turn: the powers of -b (offset from the cell centre, in cells) of up
// to 64 particles in the wave's LDS slot, lane c (and c + 64, c + 128) sums
// component c over them in FP64; the moments of the block's cells collect in
// LDS and leave as rows of MM_P2M_CELLS consecutive cells per grid.  Component
// c goes to grid c/2, real part for even c, imaginary for odd (an odd K leaves
// the last imaginary part 0).
constexpr int MM_P2M_CELLS = 32;
__global__ __launch_bounds__(256) void mm_p2m(const uint4* __restrict__ P, const int* __restrict__ start, int lg, int Pord,
                                              int K, size_t n3, float2* __restrict__ Zm) {
    __shared__ float pw[4][64][3][MM_PMAX + 1];
    __shared__ float Mo[MM_KMAX + 1][MM_P2M_CELLS + 1];  // + 1: a cell's column spreads over the banks
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const size_t cell0 = (size_t)blockIdx.x * MM_P2M_CELLS;
    const int sh = 32 - lg;
    const uint32_t mask = (1u << sh) - 1u, half = 1u << (sh - 1);
    const float inv = ldexpf(1.0f, -sh);
    int ex[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int c = lane + 64 * r;
        ex[r][0] = c < K ? c_e[c][0] : 0;
        ex[r][1] = c < K ? c_e[c][1] : 0;
        ex[r][2] = c < K ? c_e[c][2] : 0;
    }
    for (int lc = w; lc < MM_P2M_CELLS; lc += 4) {
        const size_t cell = cell0 + lc;
        const int s0 = start[cell], s1 = start[cell + 1];
        double acc[3] = {0.0, 0.0, 0.0};
        for (int base = s0; base < s1; base += 64) {
            if (base + lane < s1) {
                const uint4 p = P[base + lane];
                const uint32_t v[3] = {p.x, p.y, p.z};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const float mb = -(float)(int32_t)((v[q] & mask) - half) * inv;  // -b in cells
                    float pv = 1.0f;
                    pw[w][lane][q][0] = 1.0f;
                    for (int e = 1; e <= Pord; ++e) {
                        pv *= mb;
                        pw[w][lane][q][e] = pv;
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int cnt = min(64, s1 - base);
#pragma unroll
            for (int r = 0; r < 3; ++r)
                if (lane + 64 * r < K)
                    for (int j = 0; j < cnt; ++j)
                        acc[r] += (double)(pw[w][j][0][ex[r][0]] * pw[w][j][1][ex[r][1]] * pw[w][j][2][ex[r][2]]);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int c = lane + 64 * r;
            if (c < K) Mo[c][lc] = (float)(acc[r] * (double)c_ifact[c]);
            else if (c == K) Mo[c][lc] = 0.0f;  // the imaginary part of an unpaired last component
        }
    }
    __syncthreads();
    // grid j, cells cell0 .. cell0 + MM_P2M_CELLS - 1: (M_2j, M_2j+1)
    const int nz = (K + 1) / 2;
    for (int e = t; e < nz * MM_P2M_CELLS; e += 256) {
        const int jz = e / MM_P2M_CELLS, lc = e % MM_P2M_CELLS;
        Zm[(size_t)jz * n3 + cell0 + lc] = make_float2(Mo[2 * jz][lc], Mo[2 * jz + 1][lc]);
    }
}

// ------------------------------------------------------------ MG2L

// MM_QB wave vectors per workgroup: the transformed multipoles unpacked from
// the paired grids (Z(q) and Z(-q)) and the table rows (rounded to FP32) in
// LDS; thread (slot, qq) walks list `slot` (read through L1: every workgroup
// reads the same lists) four products at a time, all lists of equal length,
// each pair of components (2j, 2j+1) in one list, and writes L^_2j + i L^_2j+1
// (paired like the multipoles, for the inverse transforms).  A group's four
// products sum in FP32 (the multipoles are FP32 already), groups in FP64:
// those with real D^_n into (ar, ai), those with imaginary D^_n (as M^ times
// the real coefficient) into (br, bi); a component is (ar - bi) + i (ai + br).
static_assert(MM_QB * sizeof(float2) == 128 && MM_QB * sizeof(float) == 64, "LDS rows of 128 and 64 bytes");
__global__ __launch_bounds__(256) void mm_mg2l(const float2* __restrict__ Zm, const float* __restrict__ Dt32,
                                               const uint32_t* __restrict__ sched, int nc, int K, int nz, int slen,
                                               float2* __restrict__ Zl) {
    __shared__ float2 Mh[MM_KMAX][MM_QB];
    __shared__ float Dq[MM_KMAX + 1][MM_QB];
    const size_t n3 = (size_t)nc * nc * nc;
    const size_t q0 = (size_t)blockIdx.x * MM_QB;
    const int t = threadIdx.x;
    const int m = nc - 1;
    for (int e = t; e < nz * MM_QB; e += 256) {
        const int j = e / MM_QB, qq = e % MM_QB;
        const size_t q = q0 + qq;
        const int i0 = (int)(q / ((size_t)nc * nc)), i1 = (int)(q / nc % nc), i2 = (int)(q % nc);
        const size_t qn = ((size_t)((-i0) & m) * nc + (size_t)((-i1) & m)) * nc + (size_t)((-i2) & m);
        const float2 a = Zm[(size_t)j * n3 + q], b = Zm[(size_t)j * n3 + qn];
        Mh[2 * j][qq] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));          // (Z(q) + conj Z(-q)) / 2
        if (2 * j + 1 < K) Mh[2 * j + 1][qq] = make_float2(0.5f * (a.y + b.y), 0.5f * (b.x - a.x));  // / 2i
    }
    for (int e = t; e < K * MM_QB; e += 256) {
        const int c = e / MM_QB, qq = e % MM_QB;
        Dq[c][qq] = Dt32[(size_t)c * n3 + q0 + qq];
    }
    if (t < MM_QB) Dq[K][t] = 0.0f;  // the pads' row
    __syncthreads();
    const int qq = t & (MM_QB - 1), slot = t / MM_QB;
    const uint32_t* L = sched + slot * mg2l_stride(slen);
    const char* mb = (const char*)&Mh[0][qq];
    const char* db = (const char*)&Dq[0][qq];
    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0, er = 0.0, ei = 0.0;
    uint4 gn = *(const uint4*)L;  // the next group's entries, read one group ahead
    for (int i = 0; i < slen; i += 4) {
        const uint4 g = gn;
        if (i + 4 < slen) gn = *(const uint4*)(L + i + 4);
        const uint32_t w[4] = {g.x, g.y, g.z, g.w};
        float2 mv[4];
        float d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            mv[u] = *(const float2*)(mb + (w[u] & 0x7f80u));
            d[u] = *(const float*)(db + ((w[u] >> 22) << 6));
        }
        float tr = 0.0f, ti = 0.0f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tr = fmaf(mv[u].x, d[u], tr);
            ti = fmaf(mv[u].y, d[u], ti);
        }
        const uint32_t f = g.w;
        if (f & G_ODD) {
            br += (double)tr;
            bi += (double)ti;
        } else {
            ar += (double)tr;
            ai += (double)ti;
        }
        if (f & (G_SAVE | G_FLUSH)) {
            double lr = ar - bi, li = ai + br;
            ar = ai = br = bi = 0.0;
            if (f & G_SAVE) {
                er = lr;
                ei = li;
                lr = li = 0.0;
            }
            if (f & G_FLUSH)  // L_2j + i L_2j+1
                Zl[(size_t)((f >> 15) & 127) * n3 + q0 + qq] = make_float2((float)(er - li), (float)(ei + lr));
        }
    }
}

// ------------------------------------------------------------ near pairs + L2P

struct NearK {
    float eps2u, h2u, hinvu, hinv3u;  // as the direct kick, in 2^-32 units
};

__device__ __forceinline__ void two_sum_f(float& s, float& c, float v) {
#pragma clang fp contract(off)
    const float t = s + v;
    const float bp = t - s;
    c += (s - (t - bp)) + (v - bp);
    s = t;
}

// The pair term of the direct set: ddx, ddy, ddz in 2^-32 box units.
template <int KERNEL>
__device__ __forceinline__ void near_pair(float ddx, float ddy, float ddz, const NearK& kc, float& tx, float& ty,
                                          float& tz, float& tp) {
    float k, p;
    if constexpr (KERNEL == CT2_KERNEL_PLUMMER) {
        const float r2 = fmaf(ddz, ddz, fmaf(ddy, ddy, fmaf(ddx, ddx, kc.eps2u)));
        const float wv = __builtin_amdgcn_rsqf(r2);
        k = (wv * wv) * wv;
        p = wv;
    } else {
        const float r2 = fmaf(ddz, ddz, fmaf(ddy, ddy, ddx * ddx));
        const bool in = r2 < kc.h2u;
        const float wv = __builtin_amdgcn_rsqf(in ? kc.h2u : r2);
        k = in ? kc.hinv3u : (wv * wv) * wv;
        p = in ? fmaf(-0.5f * kc.hinv3u, r2, 1.5f * kc.hinvu) : wv;
    }
    tx = fmaf(ddx, k, tx);
    ty = fmaf(ddy, k, ty);
    tz = fmaf(ddz, k, tz);
    tp += p;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int MM_STAGE = 256;  // sources per wave staged in LDS at a time
#ifndef MM_L2P_CHUNK
#define MM_L2P_CHUNK 4  // local-expansion grids read ahead at a time by L2P
#endif
#ifndef MM_NEAR_UNROLL
#define MM_NEAR_UNROLL 4  // sources per step of a lane's run (LDS reads in flight)
#endif

// One wave per work item (up to 64 consecutive sinks of a row of cells,
// spanning fewer than zspan z cells); a wave takes items in turn, about one
// when the grid covers the item count.  For each (dx, dy) column of the near set (half-height w
// in z) the sources of the column's cells z_first - w .. z_last + w (the
// window) are staged in the wave's LDS slot in window order, up to MM_STAGE
// at a time.  FP32 path (nc >= 16, where the window is shorter than the
// column): a sink's own cells z - w .. z + w are one contiguous run of the
// window, so each lane loops over exactly its run; the separations are FP32
// offsets from the centre of the item's first cell, which are the minimum
// image when zspan + 3.5 < nc / 2 (a near pair is less than 4 cells apart on
// every axis).  EXACT path (8^3 cells, where the window can cover the whole
// column and a sink's run wrap inside it): every lane walks the window and
// takes the sources within w of its own cell, with separations from the
// fixed-point positions.  Each column's sum is folded in with TwoSum; then
// the local expansion of the sink's cell.
template <int KERNEL, bool EXACT>
__global__ __launch_bounds__(64 * MM_WAVES) void mm_near_l2p(const uint4* __restrict__ P, const int* __restrict__ start,
                                                             const int* __restrict__ items, const int* __restrict__ nitems,
                                                             const float2* __restrict__ Zl, int lg, int Pord, int K,
                                                             int ncols, NearK kc, double fnear, double pnear,
                                                             double ffar, double pfar, float* __restrict__ fx,
                                                             float* __restrict__ fy, float* __restrict__ fz,
                                                             float* __restrict__ pot) {
    __shared__ uint4 S[MM_WAVES][MM_STAGE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nc = 1 << lg, sh = 32 - lg, m = nc - 1;
    const size_t n3 = (size_t)nc * nc * nc;
    const uint32_t half = 1u << (sh - 1);
    const int total = *nitems;
    for (int item = blockIdx.x * MM_WAVES + w; item < total; item += gridDim.x * MM_WAVES) {
        const int row = items[3 * item], first = items[3 * item + 1], cnt = items[3 * item + 2];
        const int cx = row >> lg, cy = row & m;
        const bool live = lane < cnt;
        const uint4 me = P[first + (live ? lane : cnt - 1)];
        const int zs = (int)(me.z >> sh);
        const int zl = __builtin_amdgcn_readfirstlane(zs);
        const int zh = __shfl(zs, cnt - 1);
        const uint32_t rx = ((uint32_t)cx << sh) + half, ry = ((uint32_t)cy << sh) + half,
                       rz = ((uint32_t)zl << sh) + half;
        const float ux = (float)(int32_t)(me.x - rx), uy = (float)(int32_t)(me.y - ry), uz = (float)(int32_t)(me.z - rz);
        float sx = 0.f, sy = 0.f, sz = 0.f, sp = 0.f, cxs = 0.f, cys = 0.f, czs = 0.f, cps = 0.f;
        for (int col = 0; col < ncols; ++col) {
            const int dx = c_cols[col][0], dy = c_cols[col][1], hw = c_cols[col][2];
            const int srow = (((cx + dx) & m) << lg) | ((cy + dy) & m);
            const int* st = start + (size_t)srow * nc;  // st[nc] is the end of the column
            const int span = min(zh - zl + 2 * hw + 1, nc);
            const int zlo = span == nc ? 0 : (zl - hw) & m;
            // the window in order: cells zlo .. min(zlo + span, nc) - 1, then 0 .. zlo + span - nc - 1
            const int s0 = st[zlo], w0 = st[min(zlo + span, nc)] - s0, s1 = st[0];
            const int W = w0 + (zlo + span > nc ? st[zlo + span - nc] - s1 : 0);
            int lo = 0, hi = 0;
            if constexpr (!EXACT) {  // this sink's run: window cells zs - zl .. zs - zl + 2w
                const int za = zlo + zs - zl, zb = za + 2 * hw + 1;
                lo = za <= nc ? st[za] - s0 : w0 + st[za - nc] - s1;
                hi = zb <= nc ? st[zb] - s0 : w0 + st[zb - nc] - s1;
            }
            const int zsw = zs - hw;  // EXACT: a source of cell zc counts iff (zc - zsw) mod nc <= 2 hw
            float tx = 0.f, ty = 0.f, tz = 0.f, tp = 0.f;
            for (int jb = 0; jb < W; jb += MM_STAGE) {
                const int nj = min(MM_STAGE, W - jb);
                // EXACT: 16-byte records (fixed-point position, cell z); FP32 path:
                // planes of x, y and z offsets (4-byte entries: the runs of a wave's
                // cells start at different banks)
                float* const sf = (float*)&S[w][0];
                for (int k = lane; k < nj; k += 64) {
                    const int p = jb + k;
                    const uint4 q = P[p < w0 ? s0 + p : s1 + (p - w0)];
                    if constexpr (EXACT) {
                        S[w][k] = make_uint4(q.x, q.y, q.z, q.z >> sh);
                    } else {
                        sf[k] = (float)(int32_t)(q.x - rx);
                        sf[MM_STAGE + k] = (float)(int32_t)(q.y - ry);
                        sf[2 * MM_STAGE + k] = (float)(int32_t)(q.z - rz);
                    }
                }
                wave_sync();
                if constexpr (EXACT) {
                    for (int j = 0; j < nj; ++j) {
                        const uint4 q = S[w][j];
                        if ((((int)q.w - zsw) & m) > 2 * hw) continue;
                        near_pair<KERNEL>((float)(int32_t)(q.x - me.x), (float)(int32_t)(q.y - me.y),
                                          (float)(int32_t)(q.z - me.z), kc, tx, ty, tz, tp);
                    }
                } else {
                    const int a = max(lo - jb, 0), b = min(hi - jb, nj);
                    int j = a;
                    for (; j + MM_NEAR_UNROLL - 1 < b; j += MM_NEAR_UNROLL) {
                        float qx[MM_NEAR_UNROLL], qy[MM_NEAR_UNROLL], qz[MM_NEAR_UNROLL];
#pragma unroll
                        for (int u = 0; u < MM_NEAR_UNROLL; ++u) {
                            qx[u] = sf[j + u];
                            qy[u] = sf[MM_STAGE + j + u];
                            qz[u] = sf[2 * MM_STAGE + j + u];
                        }
#pragma unroll
                        for (int u = 0; u < MM_NEAR_UNROLL; ++u)
                            near_pair<KERNEL>(qx[u] - ux, qy[u] - uy, qz[u] - uz, kc, tx, ty, tz, tp);
                    }
                    for (; j < b; ++j)
                        near_pair<KERNEL>(sf[j] - ux, sf[MM_STAGE + j] - uy, sf[2 * MM_STAGE + j] - uz, kc, tx, ty,
                                          tz, tp);
                }
                wave_sync();
            }
            two_sum_f(sx, cxs, tx);
            two_sum_f(sy, cys, ty);
            two_sum_f(sz, czs, tz);
            two_sum_f(sp, cps, tp);
        }
        // the local expansion of the sink's own cell at a = offset from its centre, in cells
        const size_t cell = ((size_t)row << lg) | (size_t)zs;
        const float inv = ldexpf(1.0f, -sh);
        const float ax = (float)(int32_t)(me.x - rx) * inv, ay = (float)(int32_t)(me.y - ry) * inv,
                    az = (float)(int32_t)(me.z - (((uint32_t)zs << sh) + half)) * inv;
        float pa[MM_PMAX + 1], pb[MM_PMAX + 1], pc[MM_PMAX + 1];
        pa[0] = pb[0] = pc[0] = 1.0f;
        for (int e = 1; e <= Pord; ++e) {
            pa[e] = pa[e - 1] * ax;
            pb[e] = pb[e - 1] * ay;
            pc[e] = pc[e - 1] * az;
        }
        float phi = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
        // the component pairs (grids) read MM_L2P_CHUNK at a time, all in flight
        // together, instead of one load per component on the critical path
        const int nzg = (K + 1) >> 1;
        for (int j0 = 0; j0 < nzg; j0 += MM_L2P_CHUNK) {
            float2 zc[MM_L2P_CHUNK];
#pragma unroll
            for (int j = 0; j < MM_L2P_CHUNK; ++j)
                if (j0 + j < nzg) zc[j] = Zl[(size_t)(j0 + j) * n3 + cell];
#pragma unroll
            for (int j = 0; j < MM_L2P_CHUNK; ++j) {
#pragma unroll
                for (int hb = 0; hb < 2; ++hb) {
                    const int c = 2 * (j0 + j) + hb;
                    if (c >= K) continue;
                    const float L = (hb ? zc[j].y : zc[j].x) * c_ifact[c];
                    const int a = c_e[c][0], b = c_e[c][1], d = c_e[c][2];
                    phi = fmaf(L, pa[a] * pb[b] * pc[d], phi);
                    if (a) gx = fmaf(L * (float)a, pa[a - 1] * pb[b] * pc[d], gx);
                    if (b) gy = fmaf(L * (float)b, pa[a] * pb[b - 1] * pc[d], gy);
                    if (d) gz = fmaf(L * (float)d, pa[a] * pb[b] * pc[d - 1], gz);
                }
            }
        }
        if (live) {
            const uint32_t i = me.w;
            fx[i] = (float)(((double)sx + (double)cxs) * fnear + (double)gx * ffar);
            fy[i] = (float)(((double)sy + (double)cys) * fnear + (double)gy * ffar);
            fz[i] = (float)(((double)sz + (double)czs) * fnear + (double)gz * ffar);
            pot[i] = (float)(((double)sp + (double)cps) * pnear + (double)phi * pfar);
        }
    }
}

int mm_require() {
    CT2_TRY(require_particles());
    if (!g_mm.nc) return set_error(CT2_ERR_STATE, "ct2_mm_init has not been called");
    return CT2_OK;
}

int mm_forces() {
    Context& c = ctx();
    MMState& s = g_mm;
    const int n = (int)c.n, nc = s.nc, lg = ilog2(nc);
    const int n3 = nc * nc * nc;
    hipStream_t st = c.stream;
    CT2_TRY(ensure_scratch((void**)&s.keys, &s.keys_bytes, sizeof(uint32_t) * 4 * (size_t)n, "multipole-mesh keys"));
    CT2_TRY(ensure_scratch((void**)&s.Ps, &s.Ps_bytes, sizeof(uint4) * (size_t)n, "multipole-mesh particles"));
    CT2_TRY(ensure_scratch((void**)&s.start, &s.start_bytes, sizeof(int) * ((size_t)n3 + 1), "cell starts"));
    const int rows = nc * nc;
    CT2_TRY(ensure_scratch((void**)&s.cnt, &s.cnt_bytes, sizeof(int) * 2 * (size_t)rows, "item counts"));
    const int zspan = nc >= 32 ? 8 : nc >= 16 ? 4 : nc;  // the FP32 offsets stay minimum images
    const size_t max_items = (size_t)n / 64 + (size_t)n3 / (size_t)std::min(zspan, nc) + (size_t)rows + 1;
    CT2_TRY(ensure_scratch((void**)&s.items, &s.items_bytes, sizeof(int) * (3 * max_items + 1), "work items"));
    s.nitems = s.items + 3 * max_items;
    uint32_t *kin = s.keys, *kout = s.keys + n, *vin = s.keys + 2L * n, *vout = s.keys + 3L * n;
    const unsigned pb = (unsigned)((n + 255) / 256);
    mm_keys<<<pb, 256, 0, st>>>(c.x, c.y, c.z, n, lg, kin, vin);
    CT2_LAUNCH_CHECK();
    size_t tmp = 0;
    CT2_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, n, 0, 3 * lg, st));
    CT2_TRY(ensure_scratch(&s.sort_tmp, &s.sort_tmp_bytes, std::max<size_t>(tmp, 256), "multipole-mesh sort"));
    CT2_HIP(hipcub::DeviceRadixSort::SortPairs(s.sort_tmp, tmp, kin, kout, vin, vout, n, 0, 3 * lg, st));
    mm_gather<<<pb, 256, 0, st>>>(c.x, c.y, c.z, vout, n, s.Ps);
    CT2_LAUNCH_CHECK();
    mm_bounds<<<(unsigned)((n3 + 256) / 256), 256, 0, st>>>(kout, n, n3, s.start);
    CT2_LAUNCH_CHECK();
    int* first = s.cnt + rows;
    const unsigned rb = (unsigned)((rows + 255) / 256);
    mm_row_items<<<rb, 256, 0, st>>>(s.start, lg, zspan, 0, s.cnt, first, s.items, s.nitems);
    CT2_LAUNCH_CHECK();
    tmp = 0;
    CT2_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, s.cnt, first, rows, st));
    CT2_TRY(ensure_scratch(&s.scan_tmp, &s.scan_tmp_bytes, std::max<size_t>(tmp, 256), "item scan"));
    CT2_HIP(hipcub::DeviceScan::ExclusiveSum(s.scan_tmp, tmp, s.cnt, first, rows, st));
    mm_row_items<<<rb, 256, 0, st>>>(s.start, lg, zspan, 1, s.cnt, first, s.items, s.nitems);
    CT2_LAUNCH_CHECK();
    // far field
    mm_p2m<<<(unsigned)(n3 / MM_P2M_CELLS), 256, 0, st>>>(s.Ps, s.start, lg, s.P, s.K, (size_t)n3, s.Zm);
    CT2_LAUNCH_CHECK();
    CT2_TRY(fft3d_batch(s.Zm, nc, 1, false, s.nz));
    mm_mg2l<<<(unsigned)(n3 / MM_QB), 256, 0, st>>>(s.Zm, s.Dt32, s.sched, nc, s.K, s.nz, s.slen, s.Zl);
    CT2_LAUNCH_CHECK();
    CT2_TRY(fft3d_batch(s.Zl, nc, -1, false, s.nz));
    // near pairs + L2P
    const double two32 = 4294967296.0, h = c.opts.hsoft, hu = h * two32, gm = c.opts.GM;
    NearK kc;
    kc.eps2u = (float)(hu * hu);
    kc.h2u = kc.eps2u;
    kc.hinvu = (float)(1.0 / hu);
    kc.hinv3u = (float)(1.0 / (hu * hu * hu));
    const double fnear = gm * two32 * two32, pnear = gm * two32;
    const double ffar = gm * (double)nc * nc, pfar = gm * (double)nc;
    // about one work item per wave (max_items bounds the count the device
    // computes; waves past it exit at once): measured faster than 2,048 to
    // 8,192 persistent workgroups taking items in turn (unequal items)
    const unsigned grid = (unsigned)std::min<size_t>(65535, (max_items + MM_WAVES - 1) / MM_WAVES);
    const bool plummer = c.opts.kernel == CT2_KERNEL_PLUMMER, exact = nc < 16;
#define CT2_MM_NEAR(KER, EX)                                                                                    \
    mm_near_l2p<KER, EX><<<grid, 64 * MM_WAVES, 0, st>>>(s.Ps, s.start, s.items, s.nitems, s.Zl, lg, s.P, s.K, \
                                                         s.ncols, kc, fnear, pnear, ffar, pfar, c.fx, c.fy, c.fz, c.pot)
    if (plummer && exact)
        CT2_MM_NEAR(CT2_KERNEL_PLUMMER, true);
    else if (plummer)
        CT2_MM_NEAR(CT2_KERNEL_PLUMMER, false);
    else if (exact)
        CT2_MM_NEAR(CT2_KERNEL_CT2, true);
    else
        CT2_MM_NEAR(CT2_KERNEL_CT2, false);
#undef CT2_MM_NEAR
    CT2_LAUNCH_CHECK();
    c.forces_valid = true;
    return CT2_OK;
}

}  // namespace

void release_mm() {
    MMState& s = g_mm;
    for (void* p : {(void*)s.Dt, (void*)s.Dt32, (void*)s.Zm, (void*)s.Zl, (void*)s.keys, s.sort_tmp, (void*)s.Ps, (void*)s.start,
                    (void*)s.items, (void*)s.cnt, s.scan_tmp, (void*)s.tabz, (void*)s.sched})
        dev_free(p);
    s = MMState();
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_mm_init(int nc, int order) {
    CT2_TRY(require_init());
    if (!is_pow2(nc) || nc < 8 || nc > 256) return set_error(CT2_ERR_INVALID, "cell grid %d: a power of two in [8, 256]", nc);
    if (order < 1 || order > MM_PMAX) return set_error(CT2_ERR_INVALID, "expansion order %d outside [1, %d]", order, MM_PMAX);
    MMIndex I;
    I.init(order);
    const size_t n3 = (size_t)nc * nc * nc;
    const size_t dt_bytes = sizeof(double) * n3 * I.K;
    if (dt_bytes > ((size_t)8 << 30))
        return set_error(CT2_ERR_INVALID, "the derivative table of order %d on %d^3 cells (%zu bytes) exceeds 8 GiB", order,
                         nc, dt_bytes);
    Context& c = ctx();
    MMState& s = g_mm;
    hipStream_t st = c.stream;
    s.nc = 0;  // unusable until the table is complete
    // the MG2L lists: per component, its products split by the parity of |n|
    // into groups of four (padded), grouped by component pair (2j, 2j+1); the
    // pairs spread over MM_SLOTS lists by descending length (each to the
    // shortest list), padded with empty groups to equal length
    const int nz = (I.K + 1) / 2;
    std::vector<std::vector<uint32_t>> byj(nz);
    for (int k = 0; k < I.K; ++k) {
        std::vector<uint32_t>& out = byj[k >> 1];
        for (int odd = 0; odd < 2; ++odd) {
            std::vector<uint32_t> grp;
            for (size_t p = 0; p < I.pa.size(); ++p)
                if (I.pa[p] == k && (I.deg(I.pc[p]) & 1) == odd) grp.push_back(mg2l_entry(I.pb[p], I.pc[p], k >> 1, 0));
            if (grp.empty()) continue;
            while (grp.size() % 4) grp.push_back(mg2l_entry(0, I.K, k >> 1, 0));
            for (size_t g = 3; g < grp.size(); g += 4) grp[g] |= odd ? G_ODD : 0u;
            out.insert(out.end(), grp.begin(), grp.end());
        }
        if (!(k & 1)) out.back() |= G_SAVE;  // the end of the even component
    }
    for (int j = 0; j < nz; ++j) byj[j].back() |= G_FLUSH;
    std::vector<int> jorder(nz);
    for (int j = 0; j < nz; ++j) jorder[j] = j;
    std::stable_sort(jorder.begin(), jorder.end(), [&](int a, int b) { return byj[a].size() > byj[b].size(); });
    std::vector<std::vector<uint32_t>> lists(MM_SLOTS);
    for (int j : jorder) {
        auto it = std::min_element(lists.begin(), lists.end(),
                                   [](const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) { return a.size() < b.size(); });
        it->insert(it->end(), byj[j].begin(), byj[j].end());
    }
    size_t slen = 0;
    for (auto& l : lists) slen = std::max(slen, l.size());
    if (slen > (size_t)MM_SCHED_MAX) return set_error(CT2_ERR_INVALID, "MG2L list of %zu entries exceeds %d", slen, MM_SCHED_MAX);
    const size_t sstride = (size_t)mg2l_stride((int)slen);
    std::vector<uint32_t> sched((size_t)MM_SLOTS * sstride, mg2l_entry(0, I.K, 0, 0));
    for (int sl = 0; sl < MM_SLOTS; ++sl) std::copy(lists[sl].begin(), lists[sl].end(), sched.begin() + (size_t)sl * sstride);
    int ce[MM_KMAX][3] = {};
    float cf[MM_KMAX] = {};
    for (int k = 0; k < I.K; ++k) {
        for (int q = 0; q < 3; ++q) ce[k][q] = I.e[k][q];
        cf[k] = (float)(1.0 / I.fact[k]);
    }
    // the near set |offset|^2 <= 10 as columns: (dx, dy) with dz in [-w, w]
    int cn[MM_COLS_MAX][3] = {};
    int ncols = 0;
    for (int a = -3; a <= 3; ++a)
        for (int b = -3; b <= 3; ++b)
            if (a * a + b * b <= MM_NEAR2) {
                int hw = 0;
                while (a * a + b * b + (hw + 1) * (hw + 1) <= MM_NEAR2) ++hw;
                cn[ncols][0] = a;
                cn[ncols][1] = b;
                cn[ncols][2] = hw;
                ++ncols;
            }
    CT2_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_e), ce, sizeof ce, 0, hipMemcpyHostToDevice, st));
    CT2_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_ifact), cf, sizeof cf, 0, hipMemcpyHostToDevice, st));
    CT2_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_cols), cn, sizeof cn, 0, hipMemcpyHostToDevice, st));
    // buffers
    CT2_TRY(ensure_scratch((void**)&s.sched, &s.sched_bytes, sizeof(uint32_t) * sched.size(), "MG2L lists"));
    CT2_HIP(hipMemcpyAsync(s.sched, sched.data(), sizeof(uint32_t) * sched.size(), hipMemcpyHostToDevice, st));
    CT2_TRY(ensure_scratch((void**)&s.Dt, &s.Dt_bytes, dt_bytes, "multipole-mesh derivative table"));
    CT2_TRY(ensure_scratch((void**)&s.Dt32, &s.Dt32_bytes, dt_bytes / 2, "multipole-mesh FP32 table"));
    CT2_TRY(ensure_scratch((void**)&s.Zm, &s.Z_bytes, sizeof(float2) * n3 * nz, "multipole grids"));
    if (s.Zl) dev_free(s.Zl);
    s.Zl = nullptr;
    CT2_TRY(dev_alloc((void**)&s.Zl, sizeof(float2) * n3 * nz, "local-expansion grids"));
    CT2_TRY(ensure_scratch((void**)&s.tabz, &s.tabz_bytes, sizeof(double2) * n3, "table transform"));
    // the table: short-range samples on the host, one FP64 transform per component
    std::vector<double> S;
    short_samples(I, nc, S);
    std::vector<double> z(2 * n3);
    for (int k = 0; k < I.K; ++k) {
        for (size_t q = 0; q < n3; ++q) {
            z[2 * q] = S[(size_t)k * n3 + q];
            z[2 * q + 1] = 0.0;
        }
        CT2_HIP(hipMemcpyAsync(s.tabz, z.data(), sizeof(double2) * n3, hipMemcpyHostToDevice, st));
        CT2_TRY(fft3d(s.tabz, nc, 1, true));
        mm_table_finish<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(s.tabz, nc, I.e[k][0], I.e[k][1], I.e[k][2],
                                                                       s.Dt + (size_t)k * n3, s.Dt32 + (size_t)k * n3);
        CT2_LAUNCH_CHECK();
        CT2_HIP(hipStreamSynchronize(st));  // z is refilled next
    }
    s.nc = nc;
    s.P = order;
    s.K = I.K;
    s.nz = nz;
    s.slen = (int)slen;
    s.ncols = ncols;
    return CT2_OK;
}

extern "C" int ct2_mm_cells_for(size_t n, int particles_per_cell, int* nc) {
    if (!nc) return set_error(CT2_ERR_INVALID, "null output");
    if (particles_per_cell < 1) return set_error(CT2_ERR_INVALID, "particles per cell %d < 1", particles_per_cell);
    size_t g = 8;
    while (g < 256 && (2 * g) * (2 * g) * (2 * g) * (size_t)particles_per_cell <= n) g *= 2;
    *nc = (int)g;
    return CT2_OK;
}

extern "C" int ct2_mm_forces(void) {
    CT2_TRY(mm_require());
    CT2_TRY(mm_forces());
    ctx().last_path = CT2_PATH_MM;
    return CT2_OK;
}

extern "C" int ct2_kick_mm(const ct2_kick_params* p, unsigned flags, int* max_rung) {
    CT2_TRY(mm_require());
    CT2_TRY(check_kick_params(p));
    if (flags & ~(unsigned)CT2_KICK_STORE) return set_error(CT2_ERR_INVALID, "unknown flags 0x%x", flags);
    CT2_TRY(mm_forces());
    ctx().last_path = CT2_PATH_MM;
    return run_epilogue(p, max_rung);
}

extern "C" int ct2_mm_table_download(double* host) {
    CT2_TRY(require_init());
    if (!g_mm.nc) return set_error(CT2_ERR_STATE, "ct2_mm_init has not been called");
    if (!host) return set_error(CT2_ERR_INVALID, "null output");
    Context& c = ctx();
    CT2_HIP(hipMemcpyAsync(host, g_mm.Dt, sizeof(double) * (size_t)g_mm.K * g_mm.nc * g_mm.nc * g_mm.nc,
                           hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}
