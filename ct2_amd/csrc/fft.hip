// fft.hip -- 3-D complex FFT for gfx950: three passes over HBM, one per axis.
//
// Each pass is one kernel that moves a batch of lines through a Stockham
// autosort radix-8 (mixed 2/4/8) transform: the first stage reads HBM
// straight into registers, the middle stages exchange through LDS, the last
// stage writes HBM from registers.  Natural order in and out, so no
// bit-reversal pass exists (the reference's scramble() step is absorbed).
//   z pass (contiguous lines): `fft_rows`, P lanes per line, LDS rows
//     padded by one element in eight to keep LDS accesses conflict-free;
//   y and x passes (stride n and n^2): `fft_cols`, C adjacent columns per
//     workgroup so that every HBM access is a C*sizeof(complex) segment
//     (128 bytes for FP32 at C=16).
// Twiddles W_N^m are tabulated once per size on the host in FP64 libm
// (twiddles.cpp:6-54 computes the same table) and rounded to the working
// precision; each workgroup copies the table into LDS.
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "internal.h"

#ifndef COL_K512
#define COL_K512 1
#endif
#ifndef FFT_XCD_SWIZZLE
#define FFT_XCD_SWIZZLE 1
#endif
#ifndef FFT_X_SLAB_FAST
#define FFT_X_SLAB_FAST 1
#endif


namespace ct2 {
namespace {

template <typename T>
struct V2;
template <>
struct V2<float> {
    using t = float2;
};
template <>
struct V2<double> {
    using t = double2;
};

template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) {
    return T2{a.x + b.x, a.y + b.y};
}
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) {
    return T2{a.x - b.x, a.y - b.y};
}
template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    return T2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * (-i * DIR)
template <int DIR, typename T2>
__device__ __forceinline__ T2 mul_mi(T2 a) {
    if constexpr (DIR > 0)
        return T2{a.y, -a.x};
    else
        return T2{-a.y, a.x};
}

// In-register DFT of R points, X_k = sum_j v_j e^{-2 pi i DIR jk/R}.
template <int R, int DIR, typename T2>
__device__ __forceinline__ void dft(T2* v) {
    if constexpr (R == 2) {
        const T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const T2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        const T2 c = cadd(v[1], v[3]), d = mul_mi<DIR>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[1] = cadd(b, d);
        v[2] = csub(a, c);
        v[3] = csub(b, d);
    } else {
        static_assert(R == 8, "radix");
        using T = decltype(v[0].x);
        T2 e[4] = {v[0], v[2], v[4], v[6]};
        T2 o[4] = {v[1], v[3], v[5], v[7]};
        dft<4, DIR>(e);
        dft<4, DIR>(o);
        const T s = (T)0.70710678118654752440;
        // o1 *= (1 - i DIR)/sqrt2 ; o2 *= -i DIR ; o3 *= (-1 - i DIR)/sqrt2
        o[1] = T2{(o[1].x + DIR * o[1].y) * s, (o[1].y - DIR * o[1].x) * s};
        o[2] = mul_mi<DIR>(o[2]);
        o[3] = T2{(-o[3].x + DIR * o[3].y) * s, (-o[3].y - DIR * o[3].x) * s};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
}

// Radix plan: N = prod radix(N, s); the largest radix is 8 except N <= 4 and 16.
__host__ __device__ constexpr int rmax_of(int N) { return N <= 4 ? N : (N == 16 ? 4 : 8); }
__host__ __device__ constexpr int nstages_of(int N) {
    int s = 0, m = 1;
    while (m < N) {
        m *= rmax_of(N);
        ++s;
    }
    return s;
}
// stage 0 carries the remainder radix, the later stages rmax
__host__ __device__ constexpr int radix_of(int N, int s) {
    int rest = N;
    for (int k = 1; k < nstages_of(N); ++k) rest /= rmax_of(N);
    return s == 0 ? rest : rmax_of(N);
}
__host__ __device__ constexpr int ns_of(int N, int s) {
    int m = 1;
    for (int k = 0; k < s; ++k) m *= radix_of(N, k);
    return m;
}

// Butterflies per lane and stage of the column kernel (COL_K512 for the FP32
// 512-point columns, 1 elsewhere).  With 2, a workgroup takes 32 adjacent
// columns (256-byte segments, 128 KB of LDS: one workgroup per CU); with 1, 16
// columns (128-byte segments, 64 KB: two per CU), which with the XCD-contiguous
// block order measured 3% faster per 512^3 transform.
template <typename T, int N>
constexpr int col_k() {
    return (sizeof(T) == 4 && N == 512) ? COL_K512 : 1;
}

// Lines per workgroup for the row kernel and columns per workgroup for the
// column kernel, sized for <= 64 KB of LDS (128 KB for the 2-butterfly
// columns) and >= 256 lanes where possible.  K butterflies per lane and stage.
template <typename T, int N, int K = 1>
struct Cfg {
    static constexpr int RM = rmax_of(N);
    static constexpr int P = N / (RM * K);  // lanes per line
    static constexpr int ELEM = (int)sizeof(T) * 2;
    static constexpr int want_cols = (256 / P) > 16 * K ? (256 / P) : 16 * K;
    static constexpr int lds_cols = (65536 * K / (N * ELEM)) > 0 ? (65536 * K / (N * ELEM)) : 1;
    static constexpr int C0 = want_cols < lds_cols ? want_cols : lds_cols;
    static constexpr int COLS = C0 < N ? C0 : N;
    static constexpr int want_rows = (256 / P) > 1 ? (256 / P) : 1;
    static constexpr int ROWS = want_rows < N * N ? want_rows : N * N;
    static constexpr int ROWLEN = N + N / 8;  // padded LDS row
};

template <bool ROWS, typename T, int N>
struct Lanes {
    static constexpr int value =
        ROWS ? Cfg<T, N>::P * Cfg<T, N>::ROWS : Cfg<T, N, col_k<T, N>()>::P * Cfg<T, N, col_k<T, N>()>::COLS;
};

// Shared kernel body.  Element i of batch line c lives at gmem(i) in HBM
// and at lds(i) in LDS.
// U64: the first stage reads the real field W * 2^-32 from the uint64 array
// wsrc (same indexing) instead of complex data (the PM density transform).
template <typename T, int N, bool ROWS, int DIR, int S, int K = 1, bool U64 = false>
__device__ __forceinline__ void run_stages(typename V2<T>::t* __restrict__ data, typename V2<T>::t* sm,
                                           const typename V2<T>::t* tw, long base, long stride_i, int p,
                                           int c, const unsigned long long* __restrict__ wsrc = nullptr) {
    using T2 = typename V2<T>::t;
    using CF = Cfg<T, N, K>;
    constexpr int NS = nstages_of(N);
    if constexpr (S < NS) {
        constexpr int R = radix_of(N, S);
        constexpr int Ns = ns_of(N, S);
        constexpr int L = N / R;
        constexpr int NB = CF::RM * K / R;  // butterflies per lane in this stage
        constexpr bool FIRST = S == 0, LAST = S == NS - 1;
        auto lds = [&](int i) -> int {
            if constexpr (ROWS)
                return c * CF::ROWLEN + i + (i >> 3);
            else
                return i * CF::COLS + c;
        };
        T2 v[CF::RM * K];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int b = p + u * CF::P;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = b + r * L;
                if constexpr (FIRST && U64)
                    v[u * R + r] = T2{(T)((double)wsrc[base + (long)i * stride_i] * 0x1p-32), (T)0};
                else if constexpr (FIRST)
                    v[u * R + r] = data[base + (long)i * stride_i];
                else
                    v[u * R + r] = sm[lds(i)];
            }
            if constexpr (!FIRST) {
                const int k = b & (Ns - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[u * R + r] = cmul(v[u * R + r], tw[r * k * (N / (Ns * R))]);
            }
            dft<R, DIR>(&v[u * R]);
        }
        if constexpr (LAST) {
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int b = p + u * CF::P;
#pragma unroll
                for (int r = 0; r < R; ++r) data[base + (long)(b + r * L) * stride_i] = v[u * R + r];
            }
        } else {
            if constexpr (!FIRST) __syncthreads();
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int b = p + u * CF::P;
                const int k = b & (Ns - 1);
                const int o = (b / Ns) * Ns * R + k;
#pragma unroll
                for (int r = 0; r < R; ++r) sm[lds(o + r * Ns)] = v[u * R + r];
            }
            __syncthreads();
            run_stages<T, N, ROWS, DIR, S + 1, K, U64>(data, sm, tw, base, stride_i, p, c, wsrc);
        }
    }
}

template <typename T, int N, int DIR>
__device__ __forceinline__ void load_twiddles(typename V2<T>::t* tw, const typename V2<T>::t* __restrict__ twg,
                                              int lanes) {
    for (int m = threadIdx.x; m < N; m += lanes) {
        auto w = twg[m];
        if constexpr (DIR < 0) w.y = -w.y;
        tw[m] = w;
    }
}

// z pass: lines of N contiguous elements; workgroup = ROWS lines.  With U64
// the lines are read from the uint64 field wsrc (real, scaled by 2^-32) and
// the transform is written to data.
template <typename T, int N, int DIR, bool U64 = false>
__global__ __launch_bounds__((Lanes<true, T, N>::value)) void fft_rows(typename V2<T>::t* __restrict__ data,
                                                                     const typename V2<T>::t* __restrict__ twg,
                                                                     const unsigned long long* __restrict__ wsrc =
                                                                         nullptr) {
    using T2 = typename V2<T>::t;
    using CF = Cfg<T, N>;
    __shared__ T2 sm[CF::ROWS * CF::ROWLEN];
    __shared__ T2 tw[N];
    load_twiddles<T, N, DIR>(tw, twg, Lanes<true, T, N>::value);
    const int p = threadIdx.x % CF::P;
    const int c = threadIdx.x / CF::P;
    data += (size_t)blockIdx.y * ((size_t)N * N * N);  // cube blockIdx.y of a batch
    const long base = ((long)blockIdx.x * CF::ROWS + c) * N;
    if constexpr (nstages_of(N) == 1) __syncthreads();
    run_stages<T, N, true, DIR, 0, 1, U64>(data, sm, tw, base, 1, p, c, wsrc);
}

// y / x pass: COLS adjacent columns (unit stride) of lines with stride
// stride_i; block b covers columns (b % inner)*COLS.. of slab b / inner, or
// with slab_fast columns (b / N)*COLS.. of slab b % N.
template <typename T, int N, int DIR>
__global__ __launch_bounds__((Lanes<false, T, N>::value)) void fft_cols(typename V2<T>::t* __restrict__ data,
                                                                      const typename V2<T>::t* __restrict__ twg,
                                                                      long stride_i, long stride_outer, int inner,
                                                                      bool slab_fast) {
    using T2 = typename V2<T>::t;
    constexpr int K = col_k<T, N>();
    using CF = Cfg<T, N, K>;
    __shared__ T2 sm[N * CF::COLS];
    __shared__ T2 tw[N];
    load_twiddles<T, N, DIR>(tw, twg, Lanes<false, T, N>::value);
    const int c = threadIdx.x % CF::COLS;
    const int p = threadIdx.x / CF::COLS;
    data += (size_t)blockIdx.y * ((size_t)N * N * N);  // cube blockIdx.y of a batch
    unsigned b = blockIdx.x;
#if FFT_XCD_SWIZZLE
    // workgroups are dealt to the 8 XCDs in turn: give XCD x the contiguous
    // blocks [x G/8, (x+1) G/8)
    const unsigned G = gridDim.x;
    if (G % 8u == 0u) b = (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
#endif
    // y pass: the column groups of one slab side by side; x pass (lines at
    // stride n^2): slabs side by side, so that the workgroups in flight read
    // many slabs (512^3 forward transform 1.259 -> 1.242-1.257 ms; slab-fast
    // order in the y pass measured 1.311 ms)
    const long outer = slab_fast ? b % N : b / inner, in = slab_fast ? b / N : b % inner;
    const long base = outer * stride_outer + in * CF::COLS + c;
    if constexpr (nstages_of(N) == 1) __syncthreads();
    run_stages<T, N, false, DIR, 0, K>(data, sm, tw, base, stride_i, p, c);
}

// ---- twiddle tables (device, per precision and size) ----
std::map<std::pair<int, int>, void*> g_tw;  // (bytes per real, N) -> W_N^m, m < N

template <typename T>
int twiddle_table(int N, const typename V2<T>::t** out) {
    const auto key = std::make_pair((int)sizeof(T), N);
    auto it = g_tw.find(key);
    if (it == g_tw.end()) {
        std::vector<typename V2<T>::t> h(N);
        const double twopi = 2.0 * M_PI;
        for (int m = 0; m < N; ++m) {
            h[m].x = (T)std::cos(-twopi * m / N);
            h[m].y = (T)std::sin(-twopi * m / N);
        }
        void* d = nullptr;
        CT2_TRY(dev_alloc(&d, sizeof(h[0]) * N, "FFT twiddles"));
        CT2_HIP(hipMemcpy(d, h.data(), sizeof(h[0]) * N, hipMemcpyHostToDevice));
        it = g_tw.emplace(key, d).first;
    }
    *out = (const typename V2<T>::t*)it->second;
    return CT2_OK;
}

// batch cubes of N^3 one after the other (blockIdx.y); the uint64 source
// (wsrc) only with batch 1
template <typename T, int N, int DIR>
int fft3d_launch(typename V2<T>::t* d, hipStream_t st, const unsigned long long* wsrc = nullptr, int batch = 1) {
    using CF = Cfg<T, N>;
    const typename V2<T>::t* tw = nullptr;
    CT2_TRY(twiddle_table<T>(N, &tw));
    const long n = N, n2 = n * n;
    // z
    if (wsrc)
        fft_rows<T, N, DIR, true><<<(unsigned)(n2 / CF::ROWS), Lanes<true, T, N>::value, 0, st>>>(d, tw, wsrc);
    else
        fft_rows<T, N, DIR><<<dim3((unsigned)(n2 / CF::ROWS), batch), Lanes<true, T, N>::value, 0, st>>>(d, tw);
    CT2_LAUNCH_CHECK();
    const int inner = N / Cfg<T, N, col_k<T, N>()>::COLS;
    // y: columns z, slabs x
    fft_cols<T, N, DIR><<<dim3((unsigned)(n * inner), batch), Lanes<false, T, N>::value, 0, st>>>(d, tw, n, n2, inner,
                                                                                                false);
    CT2_LAUNCH_CHECK();
    // x: columns z, slabs y
    fft_cols<T, N, DIR><<<dim3((unsigned)(n * inner), batch), Lanes<false, T, N>::value, 0, st>>>(d, tw, n2, n, inner,
                                                                                                FFT_X_SLAB_FAST);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

template <typename T, int DIR>
int fft3d_dispatch(void* data, int n, hipStream_t st, int batch = 1) {
    using T2 = typename V2<T>::t;
    T2* d = (T2*)data;
    switch (n) {
        case 8: return fft3d_launch<T, 8, DIR>(d, st, nullptr, batch);
        case 16: return fft3d_launch<T, 16, DIR>(d, st, nullptr, batch);
        case 32: return fft3d_launch<T, 32, DIR>(d, st, nullptr, batch);
        case 64: return fft3d_launch<T, 64, DIR>(d, st, nullptr, batch);
        case 128: return fft3d_launch<T, 128, DIR>(d, st, nullptr, batch);
        case 256: return fft3d_launch<T, 256, DIR>(d, st, nullptr, batch);
        case 512: return fft3d_launch<T, 512, DIR>(d, st, nullptr, batch);
        case 1024:
            if constexpr (sizeof(T) == 4) return fft3d_launch<T, 1024, DIR>(d, st, nullptr, batch);
            [[fallthrough]];
        default: return set_error(CT2_ERR_INVALID, "unsupported FFT size %d", n);
    }
}

}  // namespace

// Forward FP32 transform of the real field W * 2^-32 (uint64 fixed point, n^3)
// into the complex grid out: the density transform of the PM force.
int fft3d_density(const unsigned long long* W, void* out, int n) {
    CT2_TRY(require_init());
    if (!W || !out) return set_error(CT2_ERR_INVALID, "null FFT buffer");
    hipStream_t st = ctx().stream;
    float2* d = (float2*)out;
    switch (n) {
        case 8: return fft3d_launch<float, 8, 1>(d, st, W);
        case 16: return fft3d_launch<float, 16, 1>(d, st, W);
        case 32: return fft3d_launch<float, 32, 1>(d, st, W);
        case 64: return fft3d_launch<float, 64, 1>(d, st, W);
        case 128: return fft3d_launch<float, 128, 1>(d, st, W);
        case 256: return fft3d_launch<float, 256, 1>(d, st, W);
        case 512: return fft3d_launch<float, 512, 1>(d, st, W);
        case 1024: return fft3d_launch<float, 1024, 1>(d, st, W);
        default: return set_error(CT2_ERR_INVALID, "unsupported FFT size %d", n);
    }
}

int fft3d_batch(void* data, int n, int dir, bool f64, int batch) {
    CT2_TRY(require_init());
    if (!data) return set_error(CT2_ERR_INVALID, "null FFT buffer");
    if (!is_pow2(n) || n < 8 || n > (f64 ? 512 : 1024))
        return set_error(CT2_ERR_INVALID, "FFT size %d is not a power of two in [8, %d]", n, f64 ? 512 : 1024);
    if (dir != 1 && dir != -1) return set_error(CT2_ERR_INVALID, "FFT direction must be +1 or -1");
    if (batch < 1 || batch > 65535) return set_error(CT2_ERR_INVALID, "FFT batch %d outside [1, 65535]", batch);
    hipStream_t st = ctx().stream;
    if (f64)
        return dir > 0 ? fft3d_dispatch<double, 1>(data, n, st, batch) : fft3d_dispatch<double, -1>(data, n, st, batch);
    return dir > 0 ? fft3d_dispatch<float, 1>(data, n, st, batch) : fft3d_dispatch<float, -1>(data, n, st, batch);
}

int fft3d(void* data, int n, int dir, bool f64) { return fft3d_batch(data, n, dir, f64, 1); }

void release_fft_plans() {
    for (auto& kv : g_tw) dev_free(kv.second);
    g_tw.clear();
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_fft3d_c2c(void* data, int n, int dir) { return fft3d(data, n, dir, false); }
extern "C" int ct2_fft3d_c2c_f64(void* data, int n, int dir) { return fft3d(data, n, dir, true); }
