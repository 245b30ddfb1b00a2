// mlp.hip -- the learned short-range force correction, BF16 MFMA on gfx950.
//
// No ct2 counterpart: the reference evaluates its far field with L2P
// (kick.cpp:29) and has no learned part.  Here a small network maps local
// PM-grid features to the force the PM misses (the periodic FP64 direct sum
// minus the PM force):
//   input   108 = 27 nodes (3x3x3 around the particle's nearest node) x
//           (density / mean, PM Fx, Fy, Fz / F_unit) + 3 offsets from that
//           node in cell units; padded to KIN = 128 with zeros
//   hidden  256 -> GELU -> 256 -> GELU
//   output  3 force components in units of F_unit = GM ng^2 (padded to 32)
// Weights are FP32 masters (Adam) with BF16 copies; every product runs on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  Flat weight layout:
//   W1[128][256] b1[256] W2[256][256] b2[256] W3[256][32] b3[32]
// (row = input, column = output).  GELU is the tanh form.
#include <cmath>
#include <cstring>
#include <vector>

#include "epilogue.h"
#include "internal.h"

namespace ct2 {
namespace {

constexpr int KIN = 128, HID = 256, KOUT = 32, NOUT = 3, NFEAT = 111;
constexpr int SPLITK_MAX = 256;  // K slices of the weight-gradient products at most
constexpr int LOSS_CHUNK = 4096;  // training steps whose losses are copied to the host at once
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + KIN * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
                 OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID * KOUT, NPARAM = OFF_B3 + KOUT;
constexpr float GELU_C = 0.7978845608028654f;

typedef unsigned short bf16_t;  // storage
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline bf16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
// Two floats to two bf16 (round to nearest even) in one word, low half first:
// one v_cvt_pk_bf16_f32 on gfx950.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

__host__ __device__ inline float bf2f(bf16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// GELU, tanh form, branch-free and without the library tanh or a division:
// 0.5 z (1 + tanh u) = z s, s = 1 / (1 + e^{-2u}), u = c (z + 0.044715 z^3),
// from the hardware exp2 and reciprocal (each within about 1 ulp); within
// ~1e-6 of the tanh expression, far below the bf16 rounding that follows.
__device__ __forceinline__ float gelu_sig(float z) {
    constexpr float K1 = -2.0f * GELU_C * 1.4426950408889634f;  // -2 c log2(e)
    constexpr float K3 = K1 * 0.044715f;
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z * z, K3, K1)));
}
__device__ __forceinline__ float gelu_fast(float z) { return z * gelu_sig(z); }
// its derivative from s = gelu_sig(z): s + z s (1 - s) 2 u' (1 - s, not e s:
// finite where e^{-2u} overflows)
__device__ __forceinline__ float gelu_d(float z, float s) {
    return fmaf(z * s * (1.0f - s), 2.0f * GELU_C * fmaf(3.0f * 0.044715f, z * z, 1.0f), s);
}

// ------------------------------------------------------------------ GEMM
// C[M][N] = A[M][K] * B[K][N]; A row-major (K contiguous, lda), B given as
// Bt[N][K] (K contiguous, ldb).  One wave per 32x32 output tile (2x2 MFMA
// 16x16x32 tiles); 4 waves per workgroup along M.  M % 128 == 0,
// N % 32 == 0, K % 32 == 0.
enum Epi { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

struct EpiArgs {
    float* f32;      // fp32 output [M][ldc] (STORE, BIAS: result; unused otherwise)
    int ldc;
    bf16_t* h;       // bf16 output [M][ldc] (BIAS_GELU: gelu(z); DGELU: dz)
    bf16_t* ht;      // bf16 transposed output [N][ldt] (optional)
    int ldt;
    const float* bias;  // [N]
    const bf16_t* gin;  // DGELU: the saved gelu'(z) [M][ldc]
    float* cpart = nullptr;  // DGELU: column sums of dz over each 128-row block [M / 128][N] (optional)
    bf16_t* gd = nullptr;    // BIAS_GELU: gelu'(z) for the backward pass [M][ldc] (optional)
};

__device__ __forceinline__ bf16x8 load8(const bf16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }

// Workgroup id -> (row block, column block, K slice) of a one-dimensional
// grid of (M/128) (N/32) S workgroups.  Workgroups go to the 8 XCDs in turn
// (id mod 8).  When the row blocks come in eights, every column block and K
// slice of a row block runs on the same XCD, so its A rows are fetched into
// that XCD's L2 once instead of once per column block; otherwise, when the K
// slices come in eights, every tile of a slice runs on the same XCD, so the
// slice's A and B rows are fetched there once.
struct GemmBlock {
    int rb, cb, z;
};
__device__ __forceinline__ GemmBlock gemm_block(int id, int nrb, int ncb, int S) {
    GemmBlock g;
    if (nrb % 8 == 0) {
        const int xcd = id & 7, local = id >> 3, groups = nrb >> 3;
        g.cb = local % ncb;
        const int t = local / ncb;
        g.rb = (t % groups) * 8 + xcd;
        g.z = t / groups;
    } else if (S % 8 == 0) {
        const int xcd = id & 7, local = id >> 3, tiles = nrb * ncb;
        const int t = local % tiles;
        g.rb = t % nrb;
        g.cb = t / nrb;
        g.z = (local / tiles) * 8 + xcd;
    } else {
        g.rb = id % nrb;
        const int t = id / nrb;
        g.cb = t % ncb;
        g.z = t / ncb;
    }
    return g;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave's LDS for the GELU / DGELU epilogue of one 32 x 32 tile.
struct TileLds {
    bf16_t tg[32][40];  // gelu'(z) rows (out of BIAS_GELU, into DGELU)
    bf16_t th[32][40];  // bf16 rows
    bf16_t tt[32][72];  // bf16 transposed, up to 64 rows (144-byte rows: the 16 columns a write spans hit 16 banks)
};

// The GELU / DGELU epilogue of one wave's 32 x 32 tile at (tm, tn), whose
// lane holds C[tm + 16 i + 4 (lane/16) + j][tn + 16 jj + lane%16] in
// acc[i][jj][j].  The tile passes through LDS, so that its bf16 rows leave
// (and DGELU's saved gelu'(z) arrive) as whole 64-byte rows, 16 bytes per
// lane and instruction; its transpose goes to columns roff .. roff + 31 of
// L.tt, which store_ht writes out.  The forward pass keeps gelu'(z) in bf16 for the backward one (the
// product dz = dh gelu'(z) is rounded to bf16 in any case) rather than z in
// FP32: half the bytes written and read again.  DGELU adds this lane's dz of
// column block jj to csum[jj].
template <int EPI>
__device__ __forceinline__ void gelu_tile(const EpiArgs& ep, int tm, int tn, const floatx4 (&acc)[2][2], int lane,
                                          TileLds& L, float* csum, int roff) {
    const int r16 = lane & 15;
    if constexpr (EPI == EPI_DGELU) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = q * 64 + lane, r = e >> 2, part = e & 3;
            *(uint4*)&L.tg[r][8 * part] = *(const uint4*)(ep.gin + (size_t)(tm + r) * ep.ldc + tn + 8 * part);
        }
        wave_sync();
    }
    // rows r, r + 1 of a column in pairs: one v_cvt_pk_bf16_f32 and one 4-byte
    // write of the transposed tile per pair
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
            const int c = 16 * jj + r16;
#pragma unroll
            for (int jp = 0; jp < 2; ++jp) {
                const int r = 16 * i + 4 * (lane >> 4) + 2 * jp;
                const float v0 = acc[i][jj][2 * jp], v1 = acc[i][jj][2 * jp + 1];
                uint32_t hw;
                if constexpr (EPI == EPI_BIAS_GELU) {
                    const float bz = ep.bias[tn + c];
                    const float z0 = v0 + bz, z1 = v1 + bz;
                    const float s0 = gelu_sig(z0), s1 = gelu_sig(z1);
                    hw = pack_bf16x2(z0 * s0, z1 * s1);
                    if (ep.gd) {
                        const uint32_t gw = pack_bf16x2(gelu_d(z0, s0), gelu_d(z1, s1));
                        L.tg[r][c] = (bf16_t)(gw & 0xffffu);
                        L.tg[r + 1][c] = (bf16_t)(gw >> 16);
                    }
                } else {
                    const float d0 = v0 * bf2f(L.tg[r][c]), d1 = v1 * bf2f(L.tg[r + 1][c]);
                    csum[jj] += d0;
                    csum[jj] += d1;
                    hw = pack_bf16x2(d0, d1);
                }
                L.th[r][c] = (bf16_t)(hw & 0xffffu);
                L.th[r + 1][c] = (bf16_t)(hw >> 16);
                *(uint32_t*)&L.tt[c][roff + r] = hw;
            }
        }
    wave_sync();
    if (EPI == EPI_BIAS_GELU && ep.gd) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = q * 64 + lane, r = e >> 2, part = e & 3;
            *(uint4*)(ep.gd + (size_t)(tm + r) * ep.ldc + tn + 8 * part) = *(const uint4*)&L.tg[r][8 * part];
        }
    }
    if (ep.h) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = q * 64 + lane, r = e >> 2, part = e & 3;
            *(uint4*)(ep.h + (size_t)(tm + r) * ep.ldc + tn + 8 * part) = *(const uint4*)&L.th[r][8 * part];
        }
    }
    wave_sync();  // the next tile may reuse L.th, L.tg
}

// The transposed copy of rows tm .. tm + ROWS - 1, columns tn .. tn + 31
// (ROWS = 32 or 64), from L.tt: 32 rows of 2 ROWS contiguous bytes, 16 per
// lane and store.  A 64-row tile leaves as whole 128-byte lines (HT rows are
// M entries apart, so 64-byte pieces cost a partial line each).
template <int ROWS>
__device__ __forceinline__ void store_ht(const EpiArgs& ep, int tm, int tn, int lane, TileLds& L) {
    constexpr int PARTS = ROWS / 8;
#pragma unroll
    for (int q = 0; q < 32 * PARTS / 64; ++q) {
        const int e = q * 64 + lane, c = e / PARTS, part = e % PARTS;
        *(uint4*)(ep.ht + (size_t)(tn + c) * ep.ldt + tm + 8 * part) = *(const uint4*)&L.tt[c][8 * part];
    }
    wave_sync();  // the next tile may reuse L.tt
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(const bf16_t* __restrict__ A, int lda, const bf16_t* __restrict__ Bt,
                                                 int ldb, int M, int N, int K, int S, EpiArgs ep) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const GemmBlock gb = gemm_block((int)blockIdx.x, M / 128, N / 32, S);
    const int tm = (gb.rb * 4 + wave) * 32, tn = gb.cb * 32;
    if (tm >= M) return;
    // split K: slice z of S sums k in [kb, ke) (EPI_STORE only; the slice's
    // partial product goes to plane z of ep.f32)
    const int kc = ((K / 32 + S - 1) / S) * 32;
    const int kb = gb.z * kc, ke = min(K, kb + kc);
    if constexpr (EPI == EPI_STORE) ep.f32 += (size_t)gb.z * M * ep.ldc;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    floatx4 acc[2][2] = {};
    const bf16_t* a0 = A + (size_t)(tm + r16) * lda + kq;
    const bf16_t* a1 = a0 + (size_t)16 * lda;
    const bf16_t* b0 = Bt + (size_t)(tn + r16) * ldb + kq;
    const bf16_t* b1 = b0 + (size_t)16 * ldb;
    for (int k = kb; k < ke; k += 32) {
        const bf16x8 fa0 = load8(a0 + k), fa1 = load8(a1 + k), fb0 = load8(b0 + k), fb1 = load8(b1 + k);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa0, fb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa0, fb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa1, fb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa1, fb1, acc[1][1], 0, 0, 0);
    }
    // lane holds C[tm + 16 i + 4 (lane/16) + j][tn + 16 jj + lane%16]
    if constexpr (EPI == EPI_BIAS_GELU || EPI == EPI_DGELU) {
        __shared__ TileLds lds[4];
        float csum[2] = {0.0f, 0.0f};  // DGELU: this lane's sums of dz per column block jj
        gelu_tile<EPI>(ep, tm, tn, acc, lane, lds[wave], csum, 0);
        if (ep.ht) store_ht<32>(ep, tm, tn, lane, lds[wave]);
        if (EPI == EPI_DGELU && ep.cpart) {
            // the bias gradient's share of this block: column sums over its
            // 128 rows (lane groups, then the four waves in order)
            __shared__ float cs[4][32];
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                float v = csum[jj];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (lane < 16) cs[wave][16 * jj + lane] = v;
            }
            __syncthreads();
            if (threadIdx.x < 32) {
                const int c = threadIdx.x;
                ep.cpart[(size_t)gb.rb * N + tn + c] = ((cs[0][c] + cs[1][c]) + cs[2][c]) + cs[3][c];
            }
        }
    } else {  // STORE, BIAS: 64-byte row pieces straight from the accumulators
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                const int col = tn + 16 * jj + r16;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = tm + 16 * i + 4 * (lane >> 4) + j;
                    const size_t o = (size_t)row * ep.ldc + col;
                    float v = acc[i][jj][j];
                    if constexpr (EPI == EPI_STORE)
                        ep.f32[o] = v;
                    else
                        ep.f32[o] = v + ep.bias[col];
                }
            }
    }
}

// The products with M and N multiples of 128 (the hidden layers and their
// weight gradients): one wave per 64 x 64 output tile (4 x 4 MFMA 16x16x32
// tiles, so that each operand fragment feeds four MFMAs), RW x 2 waves per
// workgroup tile of 64 RW x 128 (RW = 4 where that still makes a workgroup
// per CU, 2 otherwise).  B (the weights, or dZ^T in the weight gradients) passes
// through LDS in 64-k chunks, double buffered: the workgroup loads a chunk
// once (128 columns x 128 bytes, by direct memory-to-LDS loads that use no
// registers) while the current chunk's MFMAs run, instead of every pair of
// waves sharing its columns loading it from L2.  A column's eight 16-byte
// pieces sit in its 128-byte LDS row at slots p ^ (c & 7), so that the 16
// columns a fragment read spans fall on different banks.  A fragments come
// straight from memory: within a chunk, lane group g takes k 16 g + 8 h .. + 7
// in half h (the two halves of a 32-byte piece of its row, both in flight in
// the same chunk, so whole 128-byte lines are consumed at once), and B the
// same k; a 32-k chunk (K % 64 remainder) takes 8 g .. 8 g + 7.  Scheduling
// barriers keep the compiler from sinking each load to just before its first
// MFMA.  The staging buffers and the epilogue tiles share the LDS (39 or 78
// KB: four or two workgroups per CU at 128 registers).  EPI_STORE with S K
// slices writes slice z's partial product to plane z of ep.f32.
typedef __attribute__((address_space(3))) void lds_void;
template <int NW>
union GemmLds {
    bf16_t b[2][128][64];  // staged B chunks: column c's 16-byte piece p at slot p ^ (c & 7)
    TileLds t[NW];
};
// RW waves along M (4: 256-row tiles of 8 waves; 2: 128-row tiles of 4 waves)
template <int EPI, int RW>
__global__ __launch_bounds__(128 * RW) __attribute__((amdgpu_waves_per_eu(4))) void gemm_lds(
    const bf16_t* __restrict__ A, int lda, const bf16_t* __restrict__ Bt, int ldb, int M, int N, int K, int S,
    EpiArgs ep) {
    static_assert(EPI == EPI_BIAS_GELU || EPI == EPI_DGELU || EPI == EPI_STORE, "no staged BIAS epilogue");
    static_assert(RW == 2 || RW == 4, "128- or 256-row tiles");
    constexpr int NB = 8 / RW;  // 8-column blocks of a chunk each wave stages
    __shared__ GemmLds<2 * RW> L;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const GemmBlock gb = gemm_block((int)blockIdx.x, M / (64 * RW), N / 128, S);
    const int wr = wave % RW, wc = wave / RW;
    const int tm = gb.rb * 64 * RW + wr * 64, tn = gb.cb * 128 + wc * 64;
    const int kc = S == 1 ? K : (((K + 63) / 64 + S - 1) / S) * 64;
    const int kb = gb.z * kc, ke = min(K, kb + kc);
    const int r16 = lane & 15, g4 = lane >> 4;
    floatx4 acc[4][4] = {};
    const bf16_t* a = A + (size_t)(tm + r16) * lda;
    // staging: wave w moves column blocks NB w .. NB w + NB - 1 (8 columns
    // each) of a chunk straight from memory into LDS (global_load_lds_dwordx4:
    // lane l's 16 bytes land at the block's base + 16 l, i.e. column l / 8,
    // slot l % 8, which holds piece (l % 8) ^ (l / 8) of the column)
    const int sp = (lane & 7) ^ (lane >> 3);
    const bf16_t* bsrc = Bt + (size_t)(gb.cb * 128 + 8 * NB * wave + (lane >> 3)) * ldb;
    // whole 64-k chunks, then at most one 32-k chunk (K % 64 remainder),
    // outside the loop so that the accumulators flow through one path
    const int nfull = ke > kb ? (ke - kb) / 64 : 0, nch = ke > kb ? (ke - kb + 63) / 64 : 0;
    auto stage = [&](int c) {
        const int k0 = kb + 64 * c, p = ke - k0 >= 64 ? sp : (sp & 3);  // a 32-k chunk reads pieces 0..3 only
#pragma unroll
        for (int q = 0; q < NB; ++q)
            __builtin_amdgcn_global_load_lds((const void*)(bsrc + (size_t)8 * q * ldb + k0 + 8 * p),
                                             (lds_void*)&L.b[c & 1][8 * NB * wave + 8 * q][0], 16, 0, 0);
    };
    if (nch > 0) stage(0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int c = 0; c < nfull; ++c) {
        const int k0 = kb + 64 * c;
        const bf16_t(*Bl)[64] = L.b[c & 1];
        // each half's B fragments before its A loads; the next chunk's loads
        // into the other buffer start once this chunk's last LDS reads are
        // issued, before the second half's MFMAs
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 fa[4], fb[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = wc * 64 + 16 * j + r16;
                fb[j] = *(const bf16x8*)&Bl[col][8 * ((2 * g4 + h) ^ (col & 7))];
            }
            __builtin_amdgcn_sched_barrier(0);
            if (h == 1 && c + 1 < nch) stage(c + 1);
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[i] = load8(a + (size_t)16 * i * lda + k0 + 16 * g4 + 8 * h);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        __builtin_amdgcn_s_waitcnt(0);  // the next chunk has landed in LDS
        __syncthreads();
    }
    if (nch > nfull) {  // the 32-k chunk: k 8 g .. 8 g + 7
        const int k0 = kb + 64 * nfull;
        const bf16_t(*Bl)[64] = L.b[nfull & 1];
        bf16x8 fa[4], fb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = wc * 64 + 16 * j + r16;
            fb[j] = *(const bf16x8*)&Bl[col][8 * (g4 ^ (col & 7))];
            fa[j] = load8(a + (size_t)16 * j * lda + k0 + 8 * g4);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        __syncthreads();  // the epilogue reuses the staging buffers
    }
    if constexpr (EPI == EPI_STORE) {
        float* out = ep.f32 + (size_t)gb.z * M * ep.ldc;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    out[(size_t)(tm + 16 * i + 4 * (lane >> 4) + j) * ep.ldc + tn + 16 * jj + r16] = acc[i][jj][j];
    } else {
        // the staging buffers are free (the loop ended on a barrier)
        float csum[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // DGELU: this lane's sums of dz per 16-column block
#pragma unroll
        for (int j2 = 0; j2 < 2; ++j2) {
#pragma unroll
            for (int i2 = 0; i2 < 2; ++i2) {
                const floatx4 sub[2][2] = {{acc[2 * i2][2 * j2], acc[2 * i2][2 * j2 + 1]},
                                           {acc[2 * i2 + 1][2 * j2], acc[2 * i2 + 1][2 * j2 + 1]}};
                gelu_tile<EPI>(ep, tm + 32 * i2, tn + 32 * j2, sub, lane, L.t[wave], csum + 2 * j2, 32 * i2);
            }
            if (ep.ht) store_ht<64>(ep, tm, tn + 32 * j2, lane, L.t[wave]);
        }
        if (EPI == EPI_DGELU && ep.cpart) {
            // the bias gradient's share of each 128-row block (waves wr = 2 hb,
            // 2 hb + 1): each wave's column sums over its 64 rows, then the
            // two waves of a block and column half in order
            __shared__ float cs[2 * RW][64];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float v = csum[q];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (lane < 16) cs[wave][16 * q + lane] = v;
            }
            __syncthreads();
            if (threadIdx.x < 64 * RW) {
                const int hb = threadIdx.x >> 7, cc = threadIdx.x & 127, w0 = 2 * hb + RW * (cc >> 6), c = cc & 63;
                ep.cpart[(size_t)(gb.rb * (RW / 2) + hb) * N + gb.cb * 128 + cc] = cs[w0][c] + cs[w0 + 1][c];
            }
        }
    }
}

template <int EPI>
int gemm(hipStream_t st, const bf16_t* A, int lda, const bf16_t* Bt, int ldb, int M, int N, int K, const EpiArgs& ep) {
    if (M % 128 || N % 32 || K % 32) return set_error(CT2_ERR_INVALID, "gemm shape %dx%dx%d not tiled", M, N, K);
    if constexpr (EPI == EPI_BIAS_GELU || EPI == EPI_DGELU) {
        // the staged tiles where they make at least one workgroup per CU
        // (measured slower with fewer: batch 16,384 took 0.138 instead of
        // 0.128 ms per step)
        if (N % 128 == 0 && M % 256 == 0 && (M / 256) * (N / 128) >= 256) {
            gemm_lds<EPI, 4><<<(M / 256) * (N / 128), 512, 0, st>>>(A, lda, Bt, ldb, M, N, K, 1, ep);
            CT2_LAUNCH_CHECK();
            return CT2_OK;
        }
        if (N % 128 == 0) {
            gemm_lds<EPI, 2><<<(M / 128) * (N / 128), 256, 0, st>>>(A, lda, Bt, ldb, M, N, K, 1, ep);
            CT2_LAUNCH_CHECK();
            return CT2_OK;
        }
    }
    gemm_bf16<EPI><<<(M / 128) * (N / 32), 256, 0, st>>>(A, lda, Bt, ldb, M, N, K, 1, ep);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

// out[i] = sum over z of part[z][i] (len % 4 == 0) in a fixed order: 32
// thread groups each sum every 32nd plane (z = g, g + 32, ...) for 32
// consecutive entries (a float4 per thread), then group 0 adds the 32 partial
// sums in order.  Thirty-two groups of 8 threads rather than 8 of 32: the
// outputs of the weight gradients are small (8,192 to 65,536 entries), and the
// planes, not the entries, give the parallelism (256 slices of the layer-3
// gradient took 10 us with 8 groups).
constexpr int SK_GROUPS = 32, SK_ENTRIES = 256 / SK_GROUPS;
__global__ __launch_bounds__(256) void splitk_reduce(const float* __restrict__ part, int S, int len,
                                                     float* __restrict__ out) {
    __shared__ float4 red[SK_GROUPS][SK_ENTRIES];
    const int g = threadIdx.x / SK_ENTRIES, e = threadIdx.x % SK_ENTRIES;
    const int i = (blockIdx.x * SK_ENTRIES + e) * 4;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < len)
        for (int z = g; z < S; z += SK_GROUPS) {
            const float4 v = *(const float4*)(part + (size_t)z * len + i);
            s.x += v.x;
            s.y += v.y;
            s.z += v.z;
            s.w += v.w;
        }
    red[g][e] = s;
    __syncthreads();
    if (g == 0 && i < len) {
        float4 t = red[0][e];
#pragma unroll
        for (int q = 1; q < SK_GROUPS; ++q) {
            t.x += red[q][e].x;
            t.y += red[q][e].y;
            t.z += red[q][e].z;
            t.w += red[q][e].w;
        }
        *(float4*)(out + i) = t;
    }
}

// C = A B with a long K (the weight gradients, K = batch rows): K is cut into S
// slices so that the grid fills the GPU; each slice's product goes to its own
// plane of `scratch` (S * M * N floats) and splitk_reduce adds the planes in
// slice order, so the result does not depend on scheduling.
int gemm_splitk(hipStream_t st, const bf16_t* A, int lda, const bf16_t* Bt, int ldb, int M, int N, int K,
                float* scratch, int smax, float* out) {
    if (M % 128 || N % 32 || K % 32) return set_error(CT2_ERR_INVALID, "gemm shape %dx%dx%d not tiled", M, N, K);
    // 256 x 128 staged tiles, about one workgroup per CU, where K is long
    // enough for that many slices
    if (N % 128 == 0 && M % 256 == 0 && (M / 256) * (N / 128) * std::min(smax, K / 256) >= 256) {
        const int tiles = (M / 256) * (N / 128);
        const int S = std::max(1, std::min({smax, K / 256, (256 + tiles - 1) / tiles}));
        EpiArgs ep{S > 1 ? scratch : out, N, nullptr, nullptr, 0, nullptr, nullptr};
        gemm_lds<EPI_STORE, 4><<<tiles * S, 512, 0, st>>>(A, lda, Bt, ldb, M, N, K, S, ep);
        CT2_LAUNCH_CHECK();
        if (S > 1) {
            splitk_reduce<<<(M * N + 4 * SK_ENTRIES - 1) / (4 * SK_ENTRIES), 256, 0, st>>>(scratch, S, M * N, out);
            CT2_LAUNCH_CHECK();
        }
        return CT2_OK;
    }
    if (N % 128 == 0) {  // 128 x 128 workgroup tiles, at least 512 workgroups
        const int tiles = (M / 128) * (N / 128);
        const int S = std::max(1, std::min({smax, K / 256, (512 + tiles - 1) / tiles}));
        EpiArgs ep{S > 1 ? scratch : out, N, nullptr, nullptr, 0, nullptr, nullptr};
        gemm_lds<EPI_STORE, 2><<<tiles * S, 256, 0, st>>>(A, lda, Bt, ldb, M, N, K, S, ep);
        CT2_LAUNCH_CHECK();
        if (S > 1) {
            splitk_reduce<<<(M * N + 4 * SK_ENTRIES - 1) / (4 * SK_ENTRIES), 256, 0, st>>>(scratch, S, M * N, out);
            CT2_LAUNCH_CHECK();
        }
        return CT2_OK;
    }
    const int tiles = (M / 128) * (N / 32);
    const int S = std::max(1, std::min({smax, K / 256, (1024 + tiles - 1) / tiles}));
    EpiArgs ep{S > 1 ? scratch : out, N, nullptr, nullptr, 0, nullptr, nullptr};
    gemm_bf16<EPI_STORE><<<(M / 128) * (N / 32) * S, 256, 0, st>>>(A, lda, Bt, ldb, M, N, K, S, ep);
    CT2_LAUNCH_CHECK();
    if (S > 1) {
        splitk_reduce<<<(M * N + 4 * SK_ENTRIES - 1) / (4 * SK_ENTRIES), 256, 0, st>>>(scratch, S, M * N, out);
        CT2_LAUNCH_CHECK();
    }
    return CT2_OK;
}

// ------------------------------------------------------------------ weights
// bf16 copies in the layouts the GEMMs read: transposed [out][in] for the
// forward products, natural [in][out] for the input-gradient products.
// Parameter i's bf16 copies in the GEMM layouts: W1t [HID][KIN], W2t / W2n
// (transposed / as stored) and W3t / W3n; biases have none.
struct PackedW {
    bf16_t *W1t, *W2t, *W3t, *W2n, *W3n;
};
__device__ __forceinline__ void pack_param(int i, float wv, const PackedW& P) {
    if (i < OFF_W1 + KIN * HID) {  // W1[k][n] -> W1t[n][k]
        const int t = i - OFF_W1, k = t / HID, n = t % HID;
        P.W1t[n * KIN + k] = f2bf(wv);
    } else if (i >= OFF_W2 && i < OFF_W2 + HID * HID) {
        const int t = i - OFF_W2, k = t / HID, n = t % HID;
        const bf16_t v = f2bf(wv);
        P.W2t[n * HID + k] = v;
        P.W2n[t] = v;
    } else if (i >= OFF_W3 && i < OFF_W3 + HID * KOUT) {
        const int t = i - OFF_W3, k = t / KOUT, n = t % KOUT;
        const bf16_t v = f2bf(wv);
        P.W3t[n * HID + k] = v;
        P.W3n[t] = v;
    }
}
static_assert(OFF_W1 == 0, "W1 leads the parameter vector");

__global__ void pack_weights(const float* __restrict__ w, int n, PackedW P) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) pack_param(i, w[i], P);
}

// One Adam step of every parameter; the updated weights' bf16 copies are
// written at once (no separate packing pass per step).
__global__ void adam_step(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                          const float* __restrict__ g, int n, float lr, float b1, float b2, float c1, float c2, float eps,
                          PackedW P) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wn = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wn;
    pack_param(i, wn, P);
}

// ------------------------------------------------------------------ features
struct FeatGrid {
    int g;
    const unsigned long long* W;
    const float2* A;  // Fx + i Fy
    const float2* B;  // Fz + i pot
    float dscale;     // ng^3 / N: density in units of the mean
    float fscale;     // 1 / F_unit
};

__device__ __forceinline__ void gather_features(uint32_t x, uint32_t y, uint32_t z, const FeatGrid& fg, float* f) {
    const int s = 32 - fg.g;
    const uint32_t ng = 1u << fg.g, mask = ng - 1, half = 1u << (s - 1);
    const uint32_t ic = ((x + half) >> s) & mask, jc = ((y + half) >> s) & mask, kc = ((z + half) >> s) & mask;
    const float inv = 1.0f / (float)(1u << s);
    f[108] = (float)(int32_t)(x - (ic << s)) * inv;
    f[109] = (float)(int32_t)(y - (jc << s)) * inv;
    f[110] = (float)(int32_t)(z - (kc << s)) * inv;
    int q = 0;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b) {
            const size_t row = ((size_t)((ic + a) & mask) * ng + ((jc + b) & mask)) * ng;
            for (int c = -1; c <= 1; ++c, ++q) {
                const size_t cc = row + ((kc + c) & mask);
                const float2 av = fg.A[cc], bv = fg.B[cc];
                f[4 * q + 0] = (float)((double)fg.W[cc] * 0x1p-32) * fg.dscale;
                f[4 * q + 1] = av.x * fg.fscale;
                f[4 * q + 2] = av.y * fg.fscale;
                f[4 * q + 3] = bv.x * fg.fscale;
            }
        }
}

// X[r][128] (bf16, optional XT[128][ldt] and fp32 copy) for particles idx[r] (or r0 + r)
__global__ __launch_bounds__(256) void features_kernel(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                       const uint32_t* __restrict__ Z, const long* __restrict__ idx,
                                                       long r0, int m, FeatGrid fg, bf16_t* __restrict__ out,
                                                       float* __restrict__ outf) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const long p = idx ? idx[r] : r0 + r;
    float f[KIN];
    gather_features(X[p], Y[p], Z[p], fg, f);
#pragma unroll
    for (int k = NFEAT; k < KIN; ++k) f[k] = 0.0f;
    uint4* o = (uint4*)(out + (size_t)r * KIN);
#pragma unroll
    for (int v = 0; v < KIN / 8; ++v) {
        uint4 w;
        w.x = (uint32_t)f2bf(f[8 * v + 0]) | ((uint32_t)f2bf(f[8 * v + 1]) << 16);
        w.y = (uint32_t)f2bf(f[8 * v + 2]) | ((uint32_t)f2bf(f[8 * v + 3]) << 16);
        w.z = (uint32_t)f2bf(f[8 * v + 4]) | ((uint32_t)f2bf(f[8 * v + 5]) << 16);
        w.w = (uint32_t)f2bf(f[8 * v + 6]) | ((uint32_t)f2bf(f[8 * v + 7]) << 16);
        o[v] = w;
    }
    if (outf)
        for (int k = 0; k < KIN; ++k) outf[(size_t)r * KIN + k] = f[k];
}

// The per-node part of the features for the fused inference kernel: node c's
// four features (density / mean, PM force / F_unit) rounded to bf16 once, as
// 8 bytes, so that a particle's 27-node gather reads 8 B per node from one
// array.  The values are features_kernel's, bit for bit.
__global__ __launch_bounds__(256) void node_grid(FeatGrid fg, long n3, uint2* __restrict__ node) {
    const long cc = (long)blockIdx.x * 256 + threadIdx.x;
    if (cc >= n3) return;
    const float2 av = fg.A[cc], bv = fg.B[cc];
    const float d = (float)((double)fg.W[cc] * 0x1p-32) * fg.dscale;
    node[cc] = make_uint2(pack_bf16x2(d, av.x * fg.fscale), pack_bf16x2(av.y * fg.fscale, bv.x * fg.fscale));
}

// ------------------------------------------------------------------ training data
// target = (F_judge - F_pm) / F_unit; column 3 = the sample's loss weight
// 1 / (|F_judge| / F_unit)^2, so that the loss measures relative force error
__global__ void targets_kernel(const double* __restrict__ judge, const long* __restrict__ idx, int m,
                               const float* __restrict__ fx, const float* __restrict__ fy, const float* __restrict__ fz,
                               double inv_unit, float* __restrict__ T) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const long p = idx[r];
    const double jx = judge[r], jy = judge[(long)m + r], jz = judge[2L * m + r];
    T[4 * (size_t)r + 0] = (float)((jx - fx[p]) * inv_unit);
    T[4 * (size_t)r + 1] = (float)((jy - fy[p]) * inv_unit);
    T[4 * (size_t)r + 2] = (float)((jz - fz[p]) * inv_unit);
    const double f2 = (jx * jx + jy * jy + jz * jz) * inv_unit * inv_unit;
    T[4 * (size_t)r + 3] = (float)(1.0 / (f2 + 1e-12));
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Batch of M rows drawn uniformly (with replacement) from the data set:
// X[M][128], XT[128][M], T[M][4].
// Block b draws rows [64 b, 64 b + 64) (M % 64 == 0); XT is written through an
// LDS transpose, 64 contiguous rows (128 bytes) per feature.
constexpr int BATCH_ROWS = 64, BATCH_THREADS = 256;
// A training batch: row r takes sample mix64(key + r) mod size.  16 threads
// per row move its 256 bytes of features as 16-byte pieces (and one the
// 16-byte target); the block's 64 rows pass through LDS so that XT leaves as
// 128 rows of 128 contiguous bytes.
__global__ __launch_bounds__(BATCH_THREADS) void batch_kernel(const bf16_t* __restrict__ DX,
                                                              const float* __restrict__ DT, long size, uint64_t key,
                                                              int M, bf16_t* __restrict__ X, bf16_t* __restrict__ XT,
                                                              float* __restrict__ T) {
    static_assert(KIN == 16 * 8 && BATCH_ROWS * KIN == BATCH_THREADS * 32, "16 pieces of 8 features per row");
    __shared__ __align__(16) bf16_t tile[KIN][BATCH_ROWS + 8];
    const int piece = threadIdx.x & 15, rb = blockIdx.x * BATCH_ROWS;
#pragma unroll
    for (int pass = 0; pass < BATCH_ROWS / 16; ++pass) {
        const int q = pass * 16 + (threadIdx.x >> 4), r = rb + q;
        const long p = (long)(mix64(key + (uint64_t)r) % (uint64_t)size);
        const uint4 v = *(const uint4*)(DX + (size_t)p * KIN + 8 * piece);
        *(uint4*)(X + (size_t)r * KIN + 8 * piece) = v;
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
        // row q's 8-row chunk sits at chunk (q / 8) ^ (piece % 8) of the
        // feature's LDS row, so that the 16 pieces a write spans fall on 8 bank
        // groups instead of 2
        const int col = (((q >> 3) ^ (piece & 7)) << 3) | (q & 7);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            tile[8 * piece + 2 * j][col] = (bf16_t)(w[j] & 0xffffu);
            tile[8 * piece + 2 * j + 1][col] = (bf16_t)(w[j] >> 16);
        }
        if (piece == 0) *(float4*)(T + 4 * (size_t)r) = *(const float4*)(DT + 4 * (size_t)p);
    }
    __syncthreads();
    // feature f of the block's rows: 64 contiguous elements of XT row f
#pragma unroll
    for (int e = threadIdx.x; e < KIN * BATCH_ROWS / 8; e += BATCH_THREADS) {
        const int f = e >> 3, part = e & 7;
        *(uint4*)(XT + (size_t)f * M + rb + 8 * part) = *(const uint4*)&tile[f][8 * (part ^ ((f >> 3) & 7))];
    }
}

// dY = 2 w (Y - T) / M on the first 3 outputs (w = T[r][3]), 0 elsewhere,
// LOSS_ROWS rows per block (M % LOSS_ROWS == 0): dY row-major and dYT
// through LDS as KOUT rows of LOSS_ROWS contiguous entries; per-block sums
// of w |Y - T|^2 (partial[b]) and of each column of dY over the block's rows
// in order (cpart[b][o], the bias gradient's share)
constexpr int LOSS_ROWS = 64;
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ Yo, const float* __restrict__ T, int M,
                                                   bf16_t* __restrict__ dY,
                                                   bf16_t* __restrict__ dYT, float* __restrict__ partial,
                                                   float* __restrict__ cpart) {
    __shared__ float td[LOSS_ROWS][KOUT + 1];
    __shared__ __align__(16) bf16_t tt[KOUT][LOSS_ROWS + 8];
    __shared__ float red[4];
    const int t = threadIdx.x, o = t & 31, rq = t >> 5;
    const int rb = blockIdx.x * LOSS_ROWS;
    float e2 = 0.0f;
#pragma unroll
    for (int it = 0; it < LOSS_ROWS / 8; ++it) {
        const int rl = it * 8 + rq, r = rb + rl;
        float d = 0.0f;
        if (o < NOUT) {
            const float wr = T[4 * (size_t)r + 3];
            const float e = Yo[(size_t)r * KOUT + o] - T[4 * (size_t)r + o];
            e2 += wr * e * e;
            d = 2.0f * wr * e / (float)M;
        }
        const bf16_t h = f2bf(d);
        dY[(size_t)r * KOUT + o] = h;
        td[rl][o] = d;
        tt[o][rl] = h;
    }
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off);
    if ((t & 63) == 0) red[t >> 6] = e2;
    __syncthreads();
    if (t == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
    if (t < KOUT) {
        float cs = 0.0f;
        for (int rl = 0; rl < LOSS_ROWS; ++rl) cs += td[rl][t];
        cpart[(size_t)blockIdx.x * KOUT + t] = cs;
    }
    static_assert(KOUT * LOSS_ROWS == 256 * 8, "one 16-byte piece of dYT per thread");
    const int o2 = t >> 3, part = t & 7;
    *(uint4*)(dYT + (size_t)o2 * M + rb + 8 * part) = *(const uint4*)&tt[o2][8 * part];
}

// Per-block partials of a gradient vector: db[n] = sum over b < nb of
// part[b][n].
struct ColSums {
    const float* part;
    int nb, N;
    float* db;
};
// The final reductions of the bias gradients and the loss, one launch:
// block n < a.N sums column n of a, the next b.N blocks (b.N may be 0) the
// columns of b, each with 256 strided sums and an LDS tree in a fixed
// pattern; the last block, when lout is given, writes (sum of the lnb loss
// partials) / M to *lout, in FP64.  The results do not depend on scheduling.
__global__ __launch_bounds__(256) void colsums(ColSums a, ColSums b, const float* __restrict__ lp, int lnb, int M,
                                               float* __restrict__ lout) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    int n = blockIdx.x;
    if (n >= a.N + b.N) {
        double s = 0.0;
        for (int q = t; q < lnb; q += 256) s += lp[q];
        red[t] = s;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (t < h) red[t] += red[t + h];
            __syncthreads();
        }
        if (t == 0) *lout = (float)(red[0] / M);
        return;
    }
    const ColSums c = n < a.N ? a : b;
    if (n >= a.N) n -= a.N;
    float s = 0.0f;
    for (int q = t; q < c.nb; q += 256) s += c.part[(size_t)q * c.N + n];
    float* fr = (float*)red;
    fr[t] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) fr[t] += fr[t + h];
        __syncthreads();
    }
    if (t == 0) c.db[n] = fr[0];
}

// F[p] += F_unit * Y[r][0..2] for p = r0 + r
__global__ void apply_kernel(const float* __restrict__ Yo, long r0, int m, float unit, float* __restrict__ fx,
                             float* __restrict__ fy, float* __restrict__ fz) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    const long p = r0 + r;
    fx[p] += unit * Yo[(size_t)r * KOUT + 0];
    fy[p] += unit * Yo[(size_t)r * KOUT + 1];
    fz[p] += unit * Yo[(size_t)r * KOUT + 2];
}

// ------------------------------------------------------------------ fused inference
// The three layers of one tile of particles in one kernel, activations kept
// in registers.  Products are taken transposed, Z^T = W^T X^T: the weights are
// the MFMA A operand (rows = output units), the particles the B operand
// (columns), so that a lane's 16x16x32 output -- units 4 (lane/16) + j of
// particle lane%16 -- can be fed to the next layer's B operand without moving:
// of a 32-wide k-step, element i of lane group g is unit 16 (i/4) + 4 g + i%4
// of the step, and the next layer's weights are stored with their k columns
// in that order (kperm).  Eight waves per workgroup, INF_CB column blocks of
// 16 particles per wave; W2 (128 KB) sits in LDS for the kernel's life, W1
// and W3 come from L2; persistent workgroups walk the tiles.
#ifndef INF_CB_N
#define INF_CB_N 2
#endif
// INF_CB column blocks of 16 particles per wave: each weight fragment loaded
// (W1 from L2, W2 from LDS) feeds INF_CB MFMAs
#ifndef INF_WAVES_N
#define INF_WAVES_N 8
#endif
constexpr int INF_WAVES = INF_WAVES_N, INF_CB = INF_CB_N, INF_PW = 16 * INF_CB, INF_TILE = INF_WAVES * INF_PW;

__host__ __device__ constexpr int kperm(int q) { return 16 * ((q & 7) >> 2) + 4 * (q >> 3) + (q & 3); }

// W2p[n][32 s + q] = W2[32 s + kperm(q)][n];  W3p[n][32 s + q] = W3[32 s + kperm(q)][n], n < 16
__global__ void pack_infer(const float* __restrict__ w, bf16_t* __restrict__ W2p, bf16_t* __restrict__ W3p) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < HID * HID) {
        const int n = t / HID, c = t % HID;
        const int k = (c & ~31) + kperm(c & 31);
        W2p[t] = f2bf(w[OFF_W2 + (size_t)k * HID + n]);
    }
    if (t < 16 * HID) {
        const int n = t / HID, c = t % HID;
        const int k = (c & ~31) + kperm(c & 31);
        W3p[t] = f2bf(w[OFF_W3 + (size_t)k * KOUT + n]);
    }
}

// bias + GELU of 8 unit blocks (128 units) of a column block, packed as the
// next layer's B operand (bias points at the lane's first unit, 4 * (lane /
// 16)): units 4 of block r go to words 2 (r & 1), 2 (r & 1) + 1 of h[r / 2]
__device__ __forceinline__ void bias_gelu_pack8(const floatx4 (&acc)[8], const float* __restrict__ bias,
                                                uint4 (&h)[4]) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const float4 b = *(const float4*)(bias + 16 * r);
        const uint32_t lo = pack_bf16x2(gelu_fast(acc[r][0] + b.x), gelu_fast(acc[r][1] + b.y));
        const uint32_t hi = pack_bf16x2(gelu_fast(acc[r][2] + b.z), gelu_fast(acc[r][3] + b.w));
        if (r & 1) {
            h[r >> 1].z = lo;
            h[r >> 1].w = hi;
        } else {
            h[r >> 1].x = lo;
            h[r >> 1].y = hi;
        }
    }
}

// Grid index of stencil node q (0..26: (a, b, c) = (q / 9, q / 3 % 3, q % 3) - 1)
// around the nearest node (ic, jc, kc): the order of gather_features.
__device__ __forceinline__ size_t stencil_node(int q, uint32_t ic, uint32_t jc, uint32_t kc, int g, uint32_t mask) {
    const uint32_t a = (uint32_t)(q / 9) - 1u, b = (uint32_t)(q / 3 % 3) - 1u, c = (uint32_t)(q % 3) - 1u;
    return ((size_t)((ic + a) & mask) << (2 * g)) + ((size_t)((jc + b) & mask) << g) + ((kc + c) & mask);
}

__global__ __launch_bounds__(INF_WAVES * 64) void mlp_infer(const uint32_t* __restrict__ PX,
                                                           const uint32_t* __restrict__ PY,
                                                           const uint32_t* __restrict__ PZ,
                                                           const uint2* __restrict__ node, int gbits, int ntiles,
                                                           const bf16_t* __restrict__ W1t,
                                                           const bf16_t* __restrict__ W2p,
                                                           const bf16_t* __restrict__ W3p, const float* __restrict__ w,
                                                           long r0, int m, float unit, float* __restrict__ fx,
                                                           float* __restrict__ fy, float* __restrict__ fz) {
    // W2 in LDS (128 KB): 16-byte chunk j of row n at chunk j ^ (n & 15) of
    // the row, so that the 16 rows a lane group reads together (rows 512 B
    // apart) fall on 16 different 4-bank groups instead of the same one
    __shared__ uint4 sW2[HID * HID / 8];
    for (int e = threadIdx.x; e < HID * HID / 8; e += INF_WAVES * 64) {
        const int n = e / (HID / 8), j = e % (HID / 8);
        sW2[n * (HID / 8) + (j ^ (n & 15))] = ((const uint4*)W2p)[e];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, g = lane >> 4;
    const float* b1 = w + OFF_B1;
    const float* b2 = w + OFF_B2;
    const float* b3 = w + OFF_B3;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        // the weight addresses are rebuilt per tile from an opaque zero: hoisted
        // out of the loop they would take ~150 registers and spill
        int zero;
        asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
        const bf16_t* w1 = W1t + zero + l16 * KIN + 8 * g;
        const bf16_t* w3 = W3p + zero + l16 * HID + 8 * g;
        const float* bb1 = b1 + zero + 4 * g;
        const float* bb2 = b2 + zero + 4 * g;
        // this lane's particles (columns): one in each of the wave's INF_CB blocks
        const int p0 = tile * INF_TILE + wave * INF_PW + l16;
        // their nearest nodes and in-cell offsets (features 108..110); the
        // padded tail of the last tile repeats particle m - 1
        const int s = 32 - gbits;
        const uint32_t mask = (1u << gbits) - 1u, half = 1u << (s - 1);
        const float inv = 1.0f / (float)(1u << s);
        uint32_t ic[INF_CB], jc[INF_CB], kc[INF_CB], oxy[INF_CB], oz0[INF_CB];
#pragma unroll
        for (int c = 0; c < INF_CB; ++c) {
            const long q = r0 + min(p0 + 16 * c, m - 1);
            const uint32_t x = PX[q], y = PY[q], z = PZ[q];
            ic[c] = ((x + half) >> s) & mask;
            jc[c] = ((y + half) >> s) & mask;
            kc[c] = ((z + half) >> s) & mask;
            oxy[c] = pack_bf16x2((float)(int32_t)(x - (ic[c] << s)) * inv, (float)(int32_t)(y - (jc[c] << s)) * inv);
            oz0[c] = pack_bf16x2((float)(int32_t)(z - (kc[c] << s)) * inv, 0.0f);
        }
        // every layer in two halves of 128 output units, so that the live
        // accumulators are half a layer: the features of all four k-steps are
        // gathered first (the operand of lane group g at k-step ks is features
        // 32 ks + 8 g .. + 7: stencil nodes 8 ks + 2 g and + 1, 4 features
        // each; at ks = 3 group 1 holds node 26 and the offsets, groups 2, 3
        // zeros), layer 3 sums over the halves of layer 2 as they come
        bf16x8 xf[INF_CB][KIN / 32];
#pragma unroll
        for (int ks = 0; ks < KIN / 32; ++ks)
#pragma unroll
            for (int c = 0; c < INF_CB; ++c) {
                const int q = 8 * ks + 2 * g;
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if (ks < 3 || g == 0) {
                    const uint2 n0 = node[stencil_node(q, ic[c], jc[c], kc[c], gbits, mask)];
                    const uint2 n1 = node[stencil_node(q + 1, ic[c], jc[c], kc[c], gbits, mask)];
                    v = make_uint4(n0.x, n0.y, n1.x, n1.y);
                } else if (g == 1) {
                    const uint2 n0 = node[stencil_node(26, ic[c], jc[c], kc[c], gbits, mask)];
                    v = make_uint4(n0.x, n0.y, oxy[c], oz0[c]);
                }
                xf[c][ks] = __builtin_bit_cast(bf16x8, v);
            }
        __builtin_amdgcn_sched_barrier(0);
        uint4 h[INF_CB][8];
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            floatx4 acc[INF_CB][8];
#pragma unroll
            for (int c = 0; c < INF_CB; ++c)
#pragma unroll
                for (int r = 0; r < 8; ++r) acc[c][r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KIN / 32; ++ks) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const bf16x8 a = load8(w1 + 16 * (8 * hf + r) * KIN + 32 * ks);
#pragma unroll
                    for (int c = 0; c < INF_CB; ++c)
                        acc[c][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xf[c][ks], acc[c][r], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int c = 0; c < INF_CB; ++c) {
                uint4 hh[4];
                bias_gelu_pack8(acc[c], bb1 + 128 * hf, hh);
#pragma unroll
                for (int k = 0; k < 4; ++k) h[c][4 * hf + k] = hh[k];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        floatx4 y[INF_CB];
#pragma unroll
        for (int c = 0; c < INF_CB; ++c) y[c] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            floatx4 acc[INF_CB][8];
#pragma unroll
            for (int c = 0; c < INF_CB; ++c)
#pragma unroll
                for (int r = 0; r < 8; ++r) acc[c][r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < HID / 32; ++ks) {
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int row = 16 * (8 * hf + r) + l16;
                    const bf16x8 a = __builtin_bit_cast(bf16x8, sW2[row * (HID / 8) + ((4 * ks + g) ^ l16)]);
#pragma unroll
                    for (int c = 0; c < INF_CB; ++c)
                        acc[c][r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8, h[c][ks]),
                                                                            acc[c][r], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int c = 0; c < INF_CB; ++c) {
                uint4 h2[4];
                bias_gelu_pack8(acc[c], bb2 + 128 * hf, h2);
#pragma unroll
                for (int ks = 0; ks < 4; ++ks)
                    y[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(load8(w3 + 32 * (4 * hf + ks)),
                                                                   __builtin_bit_cast(bf16x8, h2[ks]), y[c], 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // lanes of group 0 hold outputs 0..3 of their particles
#pragma unroll
        for (int c = 0; c < INF_CB; ++c) {
            const int p = p0 + 16 * c;
            if (g == 0 && p < m) {
                fx[r0 + p] += unit * (y[c][0] + b3[0]);
                fy[r0 + p] += unit * (y[c][1] + b3[1]);
                fz[r0 + p] += unit * (y[c][2] + b3[2]);
            }
        }
    }
}

// ------------------------------------------------------------------ state
struct Mlp {
    bool ready = false;
    // the weights are trained (ct2_mlp_set_weights or ct2_mlp_train), and the
    // mesh they were trained to correct, when known
    bool trained = false;
    // (the features are in grid units, so the grid side itself may differ)
    bool mesh_known = false;
    int mesh_order = 0;
    double mesh_rs = 0.0, mesh_rcut = 0.0, mesh_ppn = 0.0;
    // the mesh of the training set (data_order 0: unknown or mixed)
    int data_order = 0;
    double data_rs = 0.0, data_rcut = 0.0, data_ppn = 0.0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    bf16_t *W1t = nullptr, *W2t = nullptr, *W3t = nullptr, *W2n = nullptr, *W3n = nullptr;
    bf16_t *W2p = nullptr, *W3p = nullptr;  // k-permuted copies for the fused inference kernel
    uint2* node = nullptr;                   // bf16 per-node features of the fused inference kernel
    long node_cap = 0;
    long adam_t = 0;
    // work buffers for up to cap rows
    int cap = 0;
    bf16_t *X = nullptr, *XT = nullptr, *H1 = nullptr, *H1T = nullptr, *H2 = nullptr, *H2T = nullptr;
    bf16_t *dY = nullptr, *dYT = nullptr, *dZ2 = nullptr, *dZ2T = nullptr, *dZ1T = nullptr;
    bf16_t *G1 = nullptr, *G2 = nullptr;  // gelu'(z) of the hidden layers, kept for the backward pass
    float *Yo = nullptr, *T = nullptr;
    float* lpart = nullptr;
    float* lsteps = nullptr;  // per-step losses of a training call, LOSS_CHUNK at a time
    float* gsplit = nullptr;  // split-K planes of the weight-gradient products
    float* cspart = nullptr;  // per-block column sums of the bias gradients
    float* Xf = nullptr;  // fp32 features for host download
    long* idx = nullptr;
    double* judge = nullptr;
    // training set
    long dcap = 0, dsize = 0;
    bf16_t* DX = nullptr;
    float* DT = nullptr;
} g_mlp;

void free_work() {
    for (void* p : {(void*)g_mlp.X, (void*)g_mlp.XT, (void*)g_mlp.H1, (void*)g_mlp.H1T, (void*)g_mlp.H2, (void*)g_mlp.H2T,
                    (void*)g_mlp.dY, (void*)g_mlp.dYT, (void*)g_mlp.dZ2, (void*)g_mlp.dZ2T, (void*)g_mlp.dZ1T,
                    (void*)g_mlp.G1, (void*)g_mlp.G2, (void*)g_mlp.Yo, (void*)g_mlp.T,
                    (void*)g_mlp.lpart, (void*)g_mlp.lsteps, (void*)g_mlp.gsplit,
                    (void*)g_mlp.cspart, (void*)g_mlp.Xf, (void*)g_mlp.idx,
                    (void*)g_mlp.judge})
        dev_free(p);
    g_mlp.X = g_mlp.XT = g_mlp.H1 = g_mlp.H1T = g_mlp.H2 = g_mlp.H2T = nullptr;
    g_mlp.dY = g_mlp.dYT = g_mlp.dZ2 = g_mlp.dZ2T = g_mlp.dZ1T = nullptr;
    g_mlp.G1 = g_mlp.G2 = nullptr;
    g_mlp.Yo = g_mlp.T = g_mlp.lpart = g_mlp.lsteps = g_mlp.Xf = nullptr;
    g_mlp.gsplit = g_mlp.cspart = nullptr;
    g_mlp.idx = nullptr;
    g_mlp.judge = nullptr;
    g_mlp.cap = 0;
}

// Work buffers for `rows` rows (rounded up to 128); training buffers only when asked.
int ensure_work(int rows, bool training) {
    constexpr int ROUND = INF_TILE > 256 ? INF_TILE : 256;  // whole inference tiles and GEMM tiles
    rows = (rows + ROUND - 1) / ROUND * ROUND;
    const bool have_train = g_mlp.G1 != nullptr;
    if (g_mlp.cap >= rows && (!training || have_train)) return CT2_OK;
    const int cap = std::max(rows, g_mlp.cap);
    free_work();
    const size_t R = (size_t)cap;
    int rc;
#define ALLOC(p, bytes, what) \
    if ((rc = dev_alloc((void**)&(p), (bytes), what))) { free_work(); return rc; }
    ALLOC(g_mlp.X, R * KIN * 2, "mlp X");
    ALLOC(g_mlp.H1, R * HID * 2, "mlp H1");
    ALLOC(g_mlp.H2, R * HID * 2, "mlp H2");
    ALLOC(g_mlp.Yo, R * KOUT * 4, "mlp Y");
    ALLOC(g_mlp.idx, R * sizeof(long), "mlp index");
    if (training) {
        ALLOC(g_mlp.XT, R * KIN * 2, "mlp XT");
        ALLOC(g_mlp.H1T, R * HID * 2, "mlp H1T");
        ALLOC(g_mlp.H2T, R * HID * 2, "mlp H2T");
        ALLOC(g_mlp.dY, R * KOUT * 2, "mlp dY");
        ALLOC(g_mlp.dYT, R * KOUT * 2, "mlp dYT");
        ALLOC(g_mlp.dZ2, R * HID * 2, "mlp dZ2");
        ALLOC(g_mlp.dZ2T, R * HID * 2, "mlp dZ2T");
        ALLOC(g_mlp.dZ1T, R * HID * 2, "mlp dZ1T");
        ALLOC(g_mlp.G1, R * HID * 2, "mlp gelu' 1");
        ALLOC(g_mlp.G2, R * HID * 2, "mlp gelu' 2");
        ALLOC(g_mlp.T, R * 4 * 4, "mlp T");
        ALLOC(g_mlp.lpart, (R / 8 + 1) * 4, "mlp loss");
        ALLOC(g_mlp.lsteps, LOSS_CHUNK * 4, "mlp step losses");
        ALLOC(g_mlp.gsplit, (size_t)SPLITK_MAX * HID * HID * 4, "mlp split-K planes");
        ALLOC(g_mlp.cspart, (R / LOSS_ROWS + 1) * HID * 4, "mlp column sums");
        ALLOC(g_mlp.Xf, R * KIN * 4, "mlp X fp32");
        ALLOC(g_mlp.judge, R * 4 * 8, "mlp judge");
    }
#undef ALLOC
    g_mlp.cap = cap;
    return CT2_OK;
}

int require_mlp() {
    CT2_TRY(require_init());
    if (!g_mlp.ready) return set_error(CT2_ERR_STATE, "ct2_mlp_init has not been called");
    return CT2_OK;
}

PackedW packed() { return PackedW{g_mlp.W1t, g_mlp.W2t, g_mlp.W3t, g_mlp.W2n, g_mlp.W3n}; }

int repack(hipStream_t st) {
    pack_weights<<<(int)((NPARAM + 255) / 256), 256, 0, st>>>(g_mlp.w, (int)NPARAM, packed());
    CT2_LAUNCH_CHECK();
    pack_infer<<<(HID * HID + 255) / 256, 256, 0, st>>>(g_mlp.w, g_mlp.W2p, g_mlp.W3p);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

// Forward of M rows already in g_mlp.X; keeps Z/H for training when train.
int forward(hipStream_t st, int M, bool train) {
    EpiArgs e1{nullptr, HID, g_mlp.H1, train ? g_mlp.H1T : nullptr, M, g_mlp.w + OFF_B1, nullptr, nullptr,
               train ? g_mlp.G1 : nullptr};
    CT2_TRY(gemm<EPI_BIAS_GELU>(st, g_mlp.X, KIN, g_mlp.W1t, KIN, M, HID, KIN, e1));
    EpiArgs e2{nullptr, HID, g_mlp.H2, train ? g_mlp.H2T : nullptr, M, g_mlp.w + OFF_B2, nullptr, nullptr,
               train ? g_mlp.G2 : nullptr};
    CT2_TRY(gemm<EPI_BIAS_GELU>(st, g_mlp.H1, HID, g_mlp.W2t, HID, M, HID, HID, e2));
    EpiArgs e3{g_mlp.Yo, KOUT, nullptr, nullptr, 0, g_mlp.w + OFF_B3, nullptr};
    CT2_TRY(gemm<EPI_BIAS>(st, g_mlp.H2, HID, g_mlp.W3t, HID, M, KOUT, HID, e3));
    return CT2_OK;
}

// Backward of the MSE loss for M rows with targets in g_mlp.T; gradient
// into g_mlp.grad, per-block loss partials into g_mlp.lpart and, when lout
// is given, the loss itself into *lout (device memory).
int backward(hipStream_t st, int M, float* lout) {
    float* g = g_mlp.grad;
    loss_kernel<<<M / LOSS_ROWS, 256, 0, st>>>(g_mlp.Yo, g_mlp.T, M, g_mlp.dY, g_mlp.dYT, g_mlp.lpart,
                                                g_mlp.cspart);
    CT2_LAUNCH_CHECK();
    const ColSums none{nullptr, 0, 0, nullptr};
    colsums<<<KOUT + (lout ? 1 : 0), 256, 0, st>>>(ColSums{g_mlp.cspart, M / LOSS_ROWS, KOUT, g + OFF_B3}, none,
                                                   g_mlp.lpart, M / LOSS_ROWS, M, lout);
    CT2_LAUNCH_CHECK();
    // the hidden layers' bias partials: dZ2's in the first half of cspart, dZ1's in the second
    float* cs2 = g_mlp.cspart;
    float* cs1 = g_mlp.cspart + (size_t)(M / 128) * HID;
    // dW3 = H2^T dY, db3
    CT2_TRY(gemm_splitk(st, g_mlp.H2T, M, g_mlp.dYT, M, HID, KOUT, M, g_mlp.gsplit, SPLITK_MAX, g + OFF_W3));
    // dZ2 = (dY W3^T) * gelu'(Z2)
    EpiArgs d2{nullptr, HID, g_mlp.dZ2, g_mlp.dZ2T, M, nullptr, g_mlp.G2, cs2};
    CT2_TRY(gemm<EPI_DGELU>(st, g_mlp.dY, KOUT, g_mlp.W3n, KOUT, M, HID, KOUT, d2));
    // dW2 = H1^T dZ2, db2
    CT2_TRY(gemm_splitk(st, g_mlp.H1T, M, g_mlp.dZ2T, M, HID, HID, M, g_mlp.gsplit, SPLITK_MAX, g + OFF_W2));
    // dZ1 = (dZ2 W2^T) * gelu'(Z1)
    EpiArgs d1{nullptr, HID, nullptr, g_mlp.dZ1T, M, nullptr, g_mlp.G1, cs1};
    CT2_TRY(gemm<EPI_DGELU>(st, g_mlp.dZ2, HID, g_mlp.W2n, HID, M, HID, HID, d1));
    // dW1 = X^T dZ1, db1
    CT2_TRY(gemm_splitk(st, g_mlp.XT, M, g_mlp.dZ1T, M, KIN, HID, M, g_mlp.gsplit, SPLITK_MAX, g + OFF_W1));
    // db2, db1
    colsums<<<2 * HID, 256, 0, st>>>(ColSums{cs2, M / 128, HID, g + OFF_B2}, ColSums{cs1, M / 128, HID, g + OFF_B1},
                                     nullptr, 0, M, nullptr);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

FeatGrid feat_grid() {
    Context& c = ctx();
    FeatGrid fg;
    pm_grids(&fg.g, &fg.W, &fg.A, &fg.B);
    const double ng = (double)(1L << fg.g);
    fg.dscale = (float)(ng * ng * ng / (double)c.n);
    fg.fscale = (float)(1.0 / ((double)c.opts.GM * ng * ng));
    return fg;
}

double force_unit() {
    int g;
    const unsigned long long* W;
    const float2 *A, *B;
    pm_grids(&g, &W, &A, &B);
    const double ng = (double)(1L << g);
    return (double)ctx().opts.GM * ng * ng;
}

}  // namespace

// PM force plus the learned correction for every particle, in chunks.
int mlp_correct_forces() {
    CT2_TRY(require_mlp());
    Context& c = ctx();
    hipStream_t st = c.stream;
    const long n = (long)c.n;
    const int chunk = (int)std::min<long>(n, 1L << 22);
    const FeatGrid fg = feat_grid();
    const float unit = (float)force_unit();
    const long n3 = 1L << (3 * fg.g);
    if (n3 > g_mlp.node_cap) {
        dev_free(g_mlp.node);
        g_mlp.node = nullptr;
        g_mlp.node_cap = 0;
        CT2_TRY(dev_alloc((void**)&g_mlp.node, (size_t)n3 * sizeof(uint2), "mlp node grid"));
        g_mlp.node_cap = n3;
    }
    node_grid<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>(fg, n3, g_mlp.node);
    CT2_LAUNCH_CHECK();
    for (long r0 = 0; r0 < n; r0 += chunk) {
        const int m = (int)std::min<long>(chunk, n - r0);
        const int M = (m + INF_TILE - 1) / INF_TILE * INF_TILE;
        const int ntiles = M / INF_TILE;
        mlp_infer<<<std::min(ntiles, 256), INF_WAVES * 64, 0, st>>>(c.x, c.y, c.z, g_mlp.node, fg.g, ntiles, g_mlp.W1t,
                                                                    g_mlp.W2p, g_mlp.W3p, g_mlp.w, r0, m, unit, c.fx,
                                                                    c.fy, c.fz);
        CT2_LAUNCH_CHECK();
    }
    return CT2_OK;
}

void release_mlp() {
    free_work();
    for (void* p : {(void*)g_mlp.w, (void*)g_mlp.m, (void*)g_mlp.v, (void*)g_mlp.grad, (void*)g_mlp.W1t,
                    (void*)g_mlp.W2t, (void*)g_mlp.W3t, (void*)g_mlp.W2n, (void*)g_mlp.W3n, (void*)g_mlp.W2p,
                    (void*)g_mlp.W3p, (void*)g_mlp.DX, (void*)g_mlp.DT, (void*)g_mlp.node})
        dev_free(p);
    g_mlp = Mlp{};
}

}  // namespace ct2

using namespace ct2;

extern "C" size_t ct2_mlp_num_params(void) { return NPARAM; }

extern "C" int ct2_mlp_init(uint64_t seed) {
    CT2_TRY(require_init());
    release_mlp();
    int rc;
    const size_t P = NPARAM;
    if ((rc = dev_alloc((void**)&g_mlp.w, P * 4, "mlp weights")) || (rc = dev_alloc((void**)&g_mlp.m, P * 4, "adam m")) ||
        (rc = dev_alloc((void**)&g_mlp.v, P * 4, "adam v")) || (rc = dev_alloc((void**)&g_mlp.grad, P * 4, "mlp grad")) ||
        (rc = dev_alloc((void**)&g_mlp.W1t, KIN * HID * 2, "W1t")) ||
        (rc = dev_alloc((void**)&g_mlp.W2t, HID * HID * 2, "W2t")) ||
        (rc = dev_alloc((void**)&g_mlp.W3t, HID * KOUT * 2, "W3t")) ||
        (rc = dev_alloc((void**)&g_mlp.W2n, HID * HID * 2, "W2n")) ||
        (rc = dev_alloc((void**)&g_mlp.W3n, HID * KOUT * 2, "W3n")) ||
        (rc = dev_alloc((void**)&g_mlp.W2p, HID * HID * 2, "W2p")) ||
        (rc = dev_alloc((void**)&g_mlp.W3p, 16 * HID * 2, "W3p"))) {
        release_mlp();
        return rc;
    }
    // Glorot-uniform weights on the used rows/columns, zero padding and biases
    std::vector<float> w(P, 0.0f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto uni = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((double)(z >> 11) + 0.5) * 0x1p-53;
    };
    auto fill = [&](size_t off, int rows, int cols, int used_rows, int used_cols) {
        const double lim = std::sqrt(6.0 / (used_rows + used_cols));
        for (int r = 0; r < used_rows; ++r)
            for (int c = 0; c < used_cols; ++c) w[off + (size_t)r * cols + c] = (float)((2.0 * uni() - 1.0) * lim);
        (void)rows;
    };
    fill(OFF_W1, KIN, HID, NFEAT, HID);
    fill(OFF_W2, HID, HID, HID, HID);
    fill(OFF_W3, HID, KOUT, HID, NOUT);
    Context& c = ctx();
    CT2_HIP(hipMemcpyAsync(g_mlp.w, w.data(), P * 4, hipMemcpyHostToDevice, c.stream));
    CT2_HIP(hipMemsetAsync(g_mlp.m, 0, P * 4, c.stream));
    CT2_HIP(hipMemsetAsync(g_mlp.v, 0, P * 4, c.stream));
    CT2_TRY(repack(c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    g_mlp.adam_t = 0;
    g_mlp.ready = true;
    return CT2_OK;
}

extern "C" int ct2_mlp_get_weights(float* w, size_t n) {
    CT2_TRY(require_mlp());
    if (!w || n != NPARAM) return set_error(CT2_ERR_INVALID, "weight buffer must hold %zu floats", NPARAM);
    Context& c = ctx();
    CT2_HIP(hipMemcpyAsync(w, g_mlp.w, NPARAM * 4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

extern "C" int ct2_mlp_set_weights(const float* w, size_t n) {
    CT2_TRY(require_mlp());
    if (!w || n != NPARAM) return set_error(CT2_ERR_INVALID, "weight buffer must hold %zu floats", NPARAM);
    Context& c = ctx();
    CT2_HIP(hipMemcpyAsync(g_mlp.w, w, NPARAM * 4, hipMemcpyHostToDevice, c.stream));
    CT2_HIP(hipMemsetAsync(g_mlp.m, 0, NPARAM * 4, c.stream));
    CT2_HIP(hipMemsetAsync(g_mlp.v, 0, NPARAM * 4, c.stream));
    CT2_TRY(repack(c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    g_mlp.adam_t = 0;
    g_mlp.trained = true;
    g_mlp.mesh_known = false;  // until ct2_mlp_set_mesh says what these weights were trained for
    return CT2_OK;
}

extern "C" int ct2_mlp_set_mesh(double rs_cells, int order, double rcut_rs, double particles_per_node) {
    CT2_TRY(require_mlp());
    if (!(rs_cells > 0.0) || (order != 1 && order != 2) || !(rcut_rs >= 0.0) || !(particles_per_node >= 0.0))
        return set_error(CT2_ERR_INVALID, "bad mesh description");
    g_mlp.mesh_known = true;
    g_mlp.mesh_rs = rs_cells;
    g_mlp.mesh_order = order;
    g_mlp.mesh_rcut = rcut_rs;
    g_mlp.mesh_ppn = particles_per_node;
    return CT2_OK;
}

// Host features -> outputs (the GPU forward on given inputs; for tests).
extern "C" int ct2_mlp_forward(const float* x, size_t m, float* y) {
    CT2_TRY(require_mlp());
    if (!x || !y || m == 0 || m > (1u << 24)) return set_error(CT2_ERR_INVALID, "bad forward request");
    CT2_TRY(ensure_work((int)m, false));
    Context& c = ctx();
    const int M = ((int)m + 127) / 128 * 128;
    std::vector<bf16_t> hx((size_t)M * KIN, 0);
    for (size_t i = 0; i < m * KIN; ++i) hx[i] = f2bf(x[i]);
    CT2_HIP(hipMemcpyAsync(g_mlp.X, hx.data(), hx.size() * 2, hipMemcpyHostToDevice, c.stream));
    CT2_TRY(forward(c.stream, M, false));
    std::vector<float> hy((size_t)M * KOUT);
    CT2_HIP(hipMemcpyAsync(hy.data(), g_mlp.Yo, hy.size() * 4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    for (size_t r = 0; r < m; ++r)
        for (int o = 0; o < NOUT; ++o) y[r * NOUT + o] = hy[r * KOUT + o];
    return CT2_OK;
}

// Loss and gradient over m given samples (no update); t[m][4] holds the
// three targets and the sample weight.
extern "C" int ct2_mlp_grad(const float* x, const float* t, size_t m, float* grad, float* loss) {
    CT2_TRY(require_mlp());
    if (!x || !t || !grad || m == 0 || m % 128 || m > (1u << 20))
        return set_error(CT2_ERR_INVALID, "gradient batch must be a positive multiple of 128 (<= 2^20)");
    CT2_TRY(ensure_work((int)m, true));
    Context& c = ctx();
    const int M = (int)m;
    std::vector<bf16_t> hx((size_t)M * KIN), hxt((size_t)M * KIN);
    for (int r = 0; r < M; ++r)
        for (int k = 0; k < KIN; ++k) {
            const bf16_t v = f2bf(x[(size_t)r * KIN + k]);
            hx[(size_t)r * KIN + k] = v;
            hxt[(size_t)k * M + r] = v;
        }
    std::vector<float> ht(t, t + (size_t)M * 4);
    CT2_HIP(hipMemcpyAsync(g_mlp.X, hx.data(), hx.size() * 2, hipMemcpyHostToDevice, c.stream));
    CT2_HIP(hipMemcpyAsync(g_mlp.XT, hxt.data(), hxt.size() * 2, hipMemcpyHostToDevice, c.stream));
    CT2_HIP(hipMemcpyAsync(g_mlp.T, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, c.stream));
    CT2_TRY(forward(c.stream, M, true));
    CT2_TRY(backward(c.stream, M, nullptr));
    std::vector<float> lp(M / LOSS_ROWS);
    CT2_HIP(hipMemcpyAsync(grad, g_mlp.grad, NPARAM * 4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipMemcpyAsync(lp.data(), g_mlp.lpart, lp.size() * 4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    double L = 0.0;
    for (float v : lp) L += v;
    if (loss) *loss = (float)(L / M);
    return CT2_OK;
}

// Features of sampled particles from the last PM evaluation (fp32, [m][128]).
extern "C" int ct2_mlp_features(const int64_t* idx, size_t m, float* out) {
    CT2_TRY(require_mlp());
    CT2_TRY(pm_require());
    Context& c = ctx();
    if (!idx || !out || m == 0 || m > (1u << 22)) return set_error(CT2_ERR_INVALID, "bad feature request");
    for (size_t s = 0; s < m; ++s)
        if (idx[s] < 0 || (size_t)idx[s] >= c.n) return set_error(CT2_ERR_INVALID, "index out of range");
    CT2_TRY(ensure_work((int)m, true));
    CT2_HIP(hipMemcpyAsync(g_mlp.idx, idx, m * 8, hipMemcpyHostToDevice, c.stream));
    features_kernel<<<((int)m + 255) / 256, 256, 0, c.stream>>>(c.x, c.y, c.z, g_mlp.idx, 0, (int)m, feat_grid(), g_mlp.X,
                                                                g_mlp.Xf);
    CT2_LAUNCH_CHECK();
    CT2_HIP(hipMemcpyAsync(out, g_mlp.Xf, m * KIN * 4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

extern "C" int ct2_mlp_data_reset(size_t capacity) {
    CT2_TRY(require_mlp());
    if (capacity == 0 || capacity > (1ul << 27)) return set_error(CT2_ERR_INVALID, "bad data capacity");
    dev_free(g_mlp.DX);
    dev_free(g_mlp.DT);
    g_mlp.DX = nullptr;
    g_mlp.DT = nullptr;
    g_mlp.dcap = g_mlp.dsize = 0;
    int rc;
    if ((rc = dev_alloc((void**)&g_mlp.DX, capacity * KIN * 2, "training features")) ||
        (rc = dev_alloc((void**)&g_mlp.DT, capacity * 16, "training targets")))
        return rc;
    g_mlp.dcap = (long)capacity;
    return CT2_OK;
}

extern "C" size_t ct2_mlp_data_size(void) { return (size_t)g_mlp.dsize; }

// Training samples for particles idx: features of the current PM state and
// targets (periodic FP64 direct sum - PM force) / F_unit.  Runs the PM
// force first.
extern "C" int ct2_mlp_data_append(const int64_t* idx, size_t m) {
    CT2_TRY(require_mlp());
    CT2_TRY(pm_require());
    Context& c = ctx();
    if (!g_mlp.DX) return set_error(CT2_ERR_STATE, "ct2_mlp_data_reset has not been called");
    if (!idx || m == 0 || m > (1u << 20)) return set_error(CT2_ERR_INVALID, "append 1..2^20 samples at a time");
    if (g_mlp.dsize + (long)m > g_mlp.dcap) return set_error(CT2_ERR_INVALID, "training set capacity exceeded");
    for (size_t s = 0; s < m; ++s)
        if (idx[s] < 0 || (size_t)idx[s] >= c.n) return set_error(CT2_ERR_INVALID, "index out of range");
    CT2_TRY(ensure_work((int)m, true));
    CT2_TRY(pm_forces());
    CT2_HIP(hipMemcpyAsync(g_mlp.idx, idx, m * 8, hipMemcpyHostToDevice, c.stream));
    CT2_TRY(ewald_judge(g_mlp.idx, (int)m, g_mlp.judge));
    const long off = g_mlp.dsize;
    features_kernel<<<((int)m + 255) / 256, 256, 0, c.stream>>>(c.x, c.y, c.z, g_mlp.idx, 0, (int)m, feat_grid(),
                                                                g_mlp.DX + (size_t)off * KIN, nullptr);
    CT2_LAUNCH_CHECK();
    targets_kernel<<<((int)m + 255) / 256, 256, 0, c.stream>>>(g_mlp.judge, g_mlp.idx, (int)m, c.fx, c.fy, c.fz,
                                                               1.0 / force_unit(), g_mlp.DT + (size_t)off * 4);
    CT2_LAUNCH_CHECK();
    CT2_HIP(hipStreamSynchronize(c.stream));
    g_mlp.dsize += (long)m;
    // the mesh the training set corrects (the same for every append, or unknown)
    int ng, order;
    double rs, rcut;
    pm_config(&ng, &rs, &order, &rcut);
    const double ppn = (double)c.n / ((double)ng * ng * ng);
    if (off == 0) {
        g_mlp.data_rs = rs, g_mlp.data_order = order, g_mlp.data_rcut = rcut, g_mlp.data_ppn = ppn;
    } else if (g_mlp.data_rs != rs || g_mlp.data_order != order || g_mlp.data_rcut != rcut || g_mlp.data_ppn != ppn) {
        g_mlp.data_order = 0;
    }
    return CT2_OK;
}

// `steps` Adam steps on batches of `batch` samples drawn from the training set.
extern "C" int ct2_mlp_train(int steps, int batch, float lr, uint64_t seed, float* losses) {
    CT2_TRY(require_mlp());
    if (g_mlp.dsize == 0) return set_error(CT2_ERR_STATE, "empty training set");
    if (steps <= 0 || batch <= 0 || batch % 128 || batch > (1 << 20) || !(lr > 0.0f))
        return set_error(CT2_ERR_INVALID, "bad training parameters");
    CT2_TRY(ensure_work(batch, true));
    Context& c = ctx();
    hipStream_t st = c.stream;
    const float b1 = 0.9f, b2 = 0.999f;
    // the step losses stay on the device until LOSS_CHUNK of them (or the
    // last) are done: no host round trip between steps
    auto step = [&](int s) -> int {
        const uint64_t key = (seed * 0xD1B54A32D192ED03ull) ^ ((uint64_t)s * 0x632BE59BD9B4E019ull);
        batch_kernel<<<batch / BATCH_ROWS, BATCH_THREADS, 0, st>>>(g_mlp.DX, g_mlp.DT, g_mlp.dsize, key, batch,
                                                                   g_mlp.X, g_mlp.XT, g_mlp.T);
        CT2_LAUNCH_CHECK();
        CT2_TRY(forward(st, batch, true));
        const int slot = s % LOSS_CHUNK;
        CT2_TRY(backward(st, batch, losses ? g_mlp.lsteps + slot : nullptr));
        ++g_mlp.adam_t;
        const float c1 = 1.0f - std::pow(b1, (float)g_mlp.adam_t), c2 = 1.0f - std::pow(b2, (float)g_mlp.adam_t);
        adam_step<<<(int)((NPARAM + 255) / 256), 256, 0, st>>>(g_mlp.w, g_mlp.m, g_mlp.v, g_mlp.grad, (int)NPARAM, lr, b1,
                                                               b2, c1, c2, 1e-8f, packed());
        CT2_LAUNCH_CHECK();
        if (losses && (slot == LOSS_CHUNK - 1 || s == steps - 1)) {
            CT2_HIP(hipMemcpyAsync(losses + (s - slot), g_mlp.lsteps, (size_t)(slot + 1) * 4, hipMemcpyDeviceToHost,
                                   st));
            CT2_HIP(hipStreamSynchronize(st));
        }
        return CT2_OK;
    };
    int rc = CT2_OK;
    for (int s = 0; s < steps && rc == CT2_OK; ++s) rc = step(s);
    // the inference layouts follow the weights even when a step failed (Adam
    // keeps the training layouts current itself)
    const int rp = repack(st);
    if (rc != CT2_OK) return rc;
    CT2_TRY(rp);
    CT2_HIP(hipStreamSynchronize(st));
    g_mlp.trained = true;
    g_mlp.mesh_known = g_mlp.data_order > 0;
    g_mlp.mesh_rs = g_mlp.data_rs, g_mlp.mesh_order = g_mlp.data_order;
    g_mlp.mesh_rcut = g_mlp.data_rcut, g_mlp.mesh_ppn = g_mlp.data_ppn;
    return CT2_OK;
}

// PM force plus the learned correction into the stored-force arrays.
extern "C" int ct2_pm_mlp_forces(void) {
    CT2_TRY(pm_forces());
    return mlp_correct_forces();
}

// Whether the learned correction can correct the current PM force: trained
// weights, and the mesh they were trained for (when known) is the current
// one: split radius in cells, assignment, near-field cutoff and particles per
// node (the grid side may differ: every feature is in grid units).
bool mlp_usable() {
    if (!g_mlp.ready || !g_mlp.trained) return false;
    int ng, order;
    double rs, rcut;
    pm_config(&ng, &rs, &order, &rcut);
    if (ng == 0) return false;
    if (!g_mlp.mesh_known) return true;
    const double ppn = (double)ctx().n / ((double)ng * ng * ng);
    auto close = [](double a, double b) { return std::fabs(a - b) <= 1e-9 * std::max(std::fabs(a), std::fabs(b)); };
    return order == g_mlp.mesh_order && close(rs, g_mlp.mesh_rs) &&
           close(rcut, g_mlp.mesh_rcut) && (g_mlp.mesh_ppn == 0.0 || close(ppn, g_mlp.mesh_ppn));
}

extern "C" int ct2_kick_pm_mlp(const ct2_kick_params* p, unsigned flags, int* max_rung) {
    CT2_TRY(require_particles());
    CT2_TRY(check_kick_params(p));
    if (flags & ~(unsigned)CT2_KICK_STORE) return set_error(CT2_ERR_INVALID, "unknown flags 0x%x", flags);
    // without usable weights the direct-sum kick is the fallback
    if (!mlp_usable()) return ct2_kick(p, flags, max_rung);
    CT2_TRY(pm_forces());
    CT2_TRY(mlp_correct_forces());
    ctx().last_path = CT2_PATH_PM_MLP;
    return run_epilogue(p, max_rung);
}
