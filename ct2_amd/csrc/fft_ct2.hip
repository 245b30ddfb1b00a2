// fft_ct2.hip -- ct2's own FFT entry points (fft.hpp:18-29) on the GPU, with
// ct2's data conventions, for callers that hold their data in device memory:
//
//   fft_complex (fft_complex_vector.cpp:9-65)  bit-reversed input, natural
//       output, FFTW_FORWARD sign for dir = +1 (dir = -1: the conjugate
//       transform, the evident intent of :57-58), unnormalised; `lanes`
//       interleaved transforms, element n lane l at re X[2 lanes n + l],
//       im X[2 lanes n + lanes + l] (complex<simd_f64>).
//   fft_real (fft_real_scalar.cpp:9-180, fft_real_vector.cpp:9-183)  real
//       bit-reversed input, half-complex output X[k] = Re F_k,
//       X[N-k] = Im F_k; lanes interleaved as [n][lane].
//   scramble (scramble.cpp:8-43), transpose_yz / transpose_r2c /
//       transpose_c2r (transpose.cpp:7-128, intended maps), and the 3-D
//       fft_r2c_3d / fft_c2r_3d (fft_r2c_3d.cpp, fft_c2r_3d.cpp: stubs in the
//       reference; here the planar layout out[(N/2+1)(N i + j) + k] = Re,
//       + N^2 (N/2+1) for Im, that transpose.cpp:36-56 prepares).
//
// Each 1-D transform is one workgroup: the whole sequence sits in LDS as FP64
// complex (N <= 4096: 64 KiB), log2 N radix-2 butterfly stages in place (the
// input already is in bit-reversed order, so the classic iterative DIT
// produces natural order), twiddles from an FP64 libm table, e^{-2 pi i m/N}.
// The 3-D transforms run the FP64 3-D transform of fft.hip on a complex copy.
#include <cmath>
#include <vector>

#include "internal.h"

namespace ct2 {
namespace {

constexpr int F1_THREADS = 256;
constexpr int F1_MAXN = 4096;

struct Twiddles {
    int n = 0;
    double2* d = nullptr;
};
std::vector<Twiddles> g_tw;
double2* g_cplx = nullptr;  // 3-D scratch
size_t g_cplx_bytes = 0;

// Table W[m] = e^{-2 pi i m / N}, m < N/2, FP64 from libm.
int twiddles(int N, const double2** out) {
    for (const Twiddles& t : g_tw)
        if (t.n == N) {
            *out = t.d;
            return CT2_OK;
        }
    std::vector<double2> h(std::max(1, N / 2));
    for (int m = 0; m < N / 2; ++m) h[m] = make_double2(std::cos(2.0 * M_PI * m / N), -std::sin(2.0 * M_PI * m / N));
    Twiddles t;
    t.n = N;
    CT2_TRY(dev_alloc((void**)&t.d, h.size() * sizeof(double2), "1-D FFT twiddles"));
    CT2_HIP(hipMemcpy(t.d, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice));
    g_tw.push_back(t);
    *out = t.d;
    return CT2_OK;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// One workgroup per (batch, lane) transform.
template <bool REAL>
__global__ __launch_bounds__(F1_THREADS) void fft1d_ct2(double* __restrict__ X, int N, int lanes, int dir,
                                                        const double2* __restrict__ W) {
    extern __shared__ double2 s[];
    const int t = blockIdx.x;
    const long b = t / lanes;
    const int l = t % lanes;
    const int tid = threadIdx.x;
    if (REAL) {
        const double* src = X + b * lanes * N;
        for (int n = tid; n < N; n += F1_THREADS) s[n] = make_double2(src[(long)lanes * n + l], 0.0);
    } else {
        const double* src = X + b * 2 * lanes * N;
        for (int n = tid; n < N; n += F1_THREADS)
            s[n] = make_double2(src[2L * lanes * n + l], src[2L * lanes * n + lanes + l]);
    }
    __syncthreads();
    for (int half = 1; half < N; half <<= 1) {
        const int stride = N / (2 * half);
        for (int q = tid; q < N / 2; q += F1_THREADS) {
            const int k = q & (half - 1);
            const int i0 = ((q - k) << 1) + k, i1 = i0 + half;
            double2 w = W[k * stride];
            if (dir < 0) w.y = -w.y;
            const double2 a = s[i0], c = cmul(s[i1], w);
            s[i0] = make_double2(a.x + c.x, a.y + c.y);
            s[i1] = make_double2(a.x - c.x, a.y - c.y);
        }
        __syncthreads();
    }
    if (REAL) {
        double* dst = X + b * lanes * N;
        for (int k = tid; k <= N / 2; k += F1_THREADS) {
            const double2 v = s[k];
            dst[(long)lanes * k + l] = v.x;
            if (k > 0 && k < N / 2) dst[(long)lanes * (N - k) + l] = v.y;
        }
    } else {
        double* dst = X + b * 2 * lanes * N;
        for (int n = tid; n < N; n += F1_THREADS) {
            dst[2L * lanes * n + l] = s[n].x;
            dst[2L * lanes * n + lanes + l] = s[n].y;
        }
    }
}

__device__ __forceinline__ int bitrev(int j, int bits) { return (int)(__brev((unsigned)j) >> (32 - bits)); }

// scramble: block j of N3 values <-> block bitrev(j), within each of N1 groups.
__global__ __launch_bounds__(256) void scramble_kernel(double* __restrict__ X, long total, int N2, int N3, int bits) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int t = (int)(e % N3);
    const long gj = e / N3;
    const int j = (int)(gj % N2);
    const long g = gj / N2;
    const int r = bits ? bitrev(j, bits) : 0;
    if (j < r) {
        double* base = X + g * (long)N2 * N3;
        const double a = base[(long)j * N3 + t];
        base[(long)j * N3 + t] = base[(long)r * N3 + t];
        base[(long)r * N3 + t] = a;
    }
}

__global__ __launch_bounds__(256) void transpose_yz_kernel(double* __restrict__ X, int N) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long n3 = (long)N * N * N;
    if (e >= n3) return;
    const int z = (int)(e % N), y = (int)((e / N) % N);
    if (y < z) {
        const long x = e / ((long)N * N);
        const long j = (x * N + z) * N + y;
        const double a = X[e];
        X[e] = X[j];
        X[j] = a;
    }
}

// Half-complex rows along y of in[x][y][z] -> planar out[x][z][ky].
__global__ __launch_bounds__(256) void transpose_r2c_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                            int N) {
    const int h = N / 2 + 1;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;  // over [x][iy < h][z]
    const long tot = (long)N * h * N;
    if (e >= tot) return;
    const int iz = (int)(e % N), iy = (int)((e / N) % h);
    const long ix = e / ((long)N * h);
    const long dimag = (long)N * N * h;
    const long jr = (ix * N + iz) * h + iy;
    out[jr] = in[(ix * N + iy) * N + iz];
    out[jr + dimag] = (iy == 0 || iy == N / 2) ? 0.0 : in[(ix * N + (N - iy)) * N + iz];
}

__global__ __launch_bounds__(256) void transpose_c2r_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                            int N) {
    const int h = N / 2 + 1;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long tot = (long)N * h * N;
    if (e >= tot) return;
    const int iz = (int)(e % N), iy = (int)((e / N) % h);
    const long ix = e / ((long)N * h);
    const long dimag = (long)N * N * h;
    const long jr = (ix * N + iz) * h + iy;
    out[(ix * N + iy) * N + iz] = in[jr];
    if (iy != 0 && iy != N / 2) out[(ix * N + (N - iy)) * N + iz] = in[jr + dimag];
}

__global__ __launch_bounds__(256) void real_to_cplx(const double* __restrict__ in, double2* __restrict__ c, long n3) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < n3) c[q] = make_double2(in[q], 0.0);
}

// Planar half spectrum from the full complex spectrum.
__global__ __launch_bounds__(256) void cplx_to_planar(const double2* __restrict__ c, double* __restrict__ out, int N) {
    const int h = N / 2 + 1;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long tot = (long)N * N * h;
    if (e >= tot) return;
    const int k = (int)(e % h);
    const long ij = e / h;
    const double2 v = c[ij * N + k];
    out[e] = v.x;
    out[tot + e] = v.y;
}

// Full Hermitian spectrum from the planar half spectrum.
__global__ __launch_bounds__(256) void planar_to_cplx(const double* __restrict__ in, double2* __restrict__ c, int N) {
    const int h = N / 2 + 1;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    const long n3 = (long)N * N * N;
    if (e >= n3) return;
    const long dimag = (long)N * N * h;
    const int k = (int)(e % N), j = (int)((e / N) % N), i = (int)(e / ((long)N * N));
    if (k < h) {
        const long q = ((long)i * N + j) * h + k;
        c[e] = make_double2(in[q], in[dimag + q]);
    } else {
        const int i2 = (N - i) % N, j2 = (N - j) % N, k2 = N - k;
        const long q = ((long)i2 * N + j2) * h + k2;
        c[e] = make_double2(in[q], -in[dimag + q]);
    }
}

__global__ __launch_bounds__(256) void cplx_to_real(const double2* __restrict__ c, double* __restrict__ out, long n3) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < n3) out[q] = c[q].x;
}

int check_1d(const void* X, int N, int dir, int lanes, size_t batch) {
    CT2_TRY(require_init());
    if (!X) return set_error(CT2_ERR_INVALID, "null data");
    if (!is_pow2(N) || N < 2 || N > F1_MAXN) return set_error(CT2_ERR_INVALID, "N = %d is not a power of two in [2, %d]", N, F1_MAXN);
    if (dir != 1 && dir != -1) return set_error(CT2_ERR_INVALID, "dir must be +1 or -1");
    if (lanes != 1 && lanes != 4) return set_error(CT2_ERR_INVALID, "lanes must be 1 or 4");
    if (batch == 0 || batch * (size_t)lanes > (size_t)1 << 30) return set_error(CT2_ERR_INVALID, "bad batch %zu", batch);
    return CT2_OK;
}

int grid_1d(long total) { return (int)((total + 255) / 256); }

int check_3d(const void* a, const void* b, int N) {
    CT2_TRY(require_init());
    if (!a || !b) return set_error(CT2_ERR_INVALID, "null data");
    if (!is_pow2(N) || N < 8 || N > 512) return set_error(CT2_ERR_INVALID, "N = %d is not a power of two in [8, 512]", N);
    return CT2_OK;
}

}  // namespace

void release_fft_ct2() {
    for (Twiddles& t : g_tw) dev_free(t.d);
    g_tw.clear();
    dev_free(g_cplx);
    g_cplx = nullptr;
    g_cplx_bytes = 0;
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_fft_complex_1d(double* X, int N, int dir, int lanes, size_t batch) {
    CT2_TRY(check_1d(X, N, dir, lanes, batch));
    const double2* W;
    CT2_TRY(twiddles(N, &W));
    Context& c = ctx();
    fft1d_ct2<false><<<(unsigned)(batch * lanes), F1_THREADS, N * sizeof(double2), c.stream>>>(X, N, lanes, dir, W);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_fft_real_1d(double* X, int N, int dir, int lanes, size_t batch) {
    CT2_TRY(check_1d(X, N, dir, lanes, batch));
    const double2* W;
    CT2_TRY(twiddles(N, &W));
    Context& c = ctx();
    fft1d_ct2<true><<<(unsigned)(batch * lanes), F1_THREADS, N * sizeof(double2), c.stream>>>(X, N, lanes, dir, W);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_scramble(double* X, int N1, int N2, int N3) {
    CT2_TRY(require_init());
    if (!X || N1 <= 0 || N3 <= 0 || !is_pow2(N2)) return set_error(CT2_ERR_INVALID, "bad scramble arguments");
    const long total = (long)N1 * N2 * N3;
    if (total > (1L << 33)) return set_error(CT2_ERR_INVALID, "scramble too large");
    scramble_kernel<<<grid_1d(total), 256, 0, ctx().stream>>>(X, total, N2, N3, ilog2(N2));
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_transpose_yz(double* X, int N) {
    CT2_TRY(require_init());
    if (!X || N <= 0 || N > 2048) return set_error(CT2_ERR_INVALID, "bad transpose arguments");
    transpose_yz_kernel<<<grid_1d((long)N * N * N), 256, 0, ctx().stream>>>(X, N);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_transpose_r2c(const double* in, double* out, int N) {
    CT2_TRY(require_init());
    if (!in || !out || N < 2 || N % 2 || N > 2048) return set_error(CT2_ERR_INVALID, "bad transpose arguments");
    transpose_r2c_kernel<<<grid_1d((long)N * (N / 2 + 1) * N), 256, 0, ctx().stream>>>(in, out, N);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_transpose_c2r(const double* in, double* out, int N) {
    CT2_TRY(require_init());
    if (!in || !out || N < 2 || N % 2 || N > 2048) return set_error(CT2_ERR_INVALID, "bad transpose arguments");
    transpose_c2r_kernel<<<grid_1d((long)N * (N / 2 + 1) * N), 256, 0, ctx().stream>>>(in, out, N);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_fft_r2c_3d(const double* in, double* out, int N) {
    CT2_TRY(check_3d(in, out, N));
    Context& c = ctx();
    const long n3 = (long)N * N * N;
    CT2_TRY(ensure_scratch((void**)&g_cplx, &g_cplx_bytes, n3 * sizeof(double2), "3-D r2c scratch"));
    real_to_cplx<<<grid_1d(n3), 256, 0, c.stream>>>(in, g_cplx, n3);
    CT2_LAUNCH_CHECK();
    CT2_TRY(fft3d(g_cplx, N, 1, true));
    cplx_to_planar<<<grid_1d((long)N * N * (N / 2 + 1)), 256, 0, c.stream>>>(g_cplx, out, N);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_fft_c2r_3d(const double* in, double* out, int N) {
    CT2_TRY(check_3d(in, out, N));
    Context& c = ctx();
    const long n3 = (long)N * N * N;
    CT2_TRY(ensure_scratch((void**)&g_cplx, &g_cplx_bytes, n3 * sizeof(double2), "3-D c2r scratch"));
    planar_to_cplx<<<grid_1d(n3), 256, 0, c.stream>>>(in, g_cplx, N);
    CT2_LAUNCH_CHECK();
    CT2_TRY(fft3d(g_cplx, N, -1, true));
    cplx_to_real<<<grid_1d(n3), 256, 0, c.stream>>>(g_cplx, out, n3);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

extern "C" int ct2_twiddles(int N, double* cosw, double* sinw, double* invsinw) {
    if (N <= 0 || !cosw || !sinw || !invsinw) return set_error(CT2_ERR_INVALID, "bad twiddle arguments");
    for (int n = 0; n < N; ++n) {
        const double a = 2.0 * M_PI * n / N;
        cosw[n] = std::cos(-a);
        sinw[n] = std::sin(-a);
        invsinw[n] = std::sin(a);
    }
    return CT2_OK;
}
