// ic_grf.hip -- lattice initial conditions displaced by a Gaussian random
// field (Zel'dovich approximation), built on the device with the FP64 3-D FFT.
//
//   white noise w(q) ~ N(0,1) per lattice site (counter-based, seeded)
//   delta_hat(k) = w_hat(k) k^{n_s/2}, |m| < n/2 (spherical cut at Nyquist)
//   psi_hat(k)   = i k delta_hat(k) / k^2       (delta = -div psi)
//   psi scaled so that its rms per component is rms_cells lattice cells
//   x = q + psi (fixed32, RNE), v = vfac * psi
// The two inverse transforms carry psi_x + i psi_y and psi_z.
#include <cmath>
#include <vector>

#include "internal.h"

namespace ct2 {
namespace {

constexpr int GB = 256;

__device__ __forceinline__ uint64_t mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(GB) void grf_noise(double2* __restrict__ A, long n3, uint64_t seed) {
    const long p = (long)blockIdx.x * GB + threadIdx.x;
    if (p >= n3) return;
    const uint64_t key = mix(seed * 0xD1B54A32D192ED03ull ^ mix((uint64_t)p));
    const double u1 = ((double)(mix(key + 1) >> 11) + 0.5) * 0x1p-53;
    const double u2 = ((double)(mix(key + 2) >> 11) + 0.5) * 0x1p-53;
    A[p] = make_double2(sqrt(-2.0 * log(u1)) * cospi(2.0 * u2), 0.0);
}

__global__ __launch_bounds__(GB) void grf_filter(double2* __restrict__ A, double2* __restrict__ B, int g, double ns) {
    const long c = (long)blockIdx.x * GB + threadIdx.x;
    const int n = 1 << g;
    if (c >= (long)n * n * n) return;
    const int l = (int)(c & (n - 1)), j = (int)((c >> g) & (n - 1)), i = (int)(c >> (2 * g));
    const int mi = i < n / 2 ? i : i - n, mj = j < n / 2 ? j : j - n, ml = l < n / 2 ? l : l - n;
    const double m2 = (double)mi * mi + (double)mj * mj + (double)ml * ml;
    const double2 w = A[c];
    if (m2 == 0.0 || m2 >= 0.25 * n * n) {
        A[c] = make_double2(0.0, 0.0);
        B[c] = make_double2(0.0, 0.0);
        return;
    }
    const double twopi = 6.283185307179586476925;
    const double k2 = twopi * twopi * m2;
    const double amp = pow(k2, 0.25 * ns) / k2;  // |delta| ~ k^{ns/2}, then 1/k^2
    const double kx = twopi * mi, ky = twopi * mj, kz = twopi * ml;
    // psi_hat_d = i k_d amp w ;  A = psi_x + i psi_y, B = psi_z
    const double2 iw = make_double2(-w.y * amp, w.x * amp);  // i * amp * w
    A[c] = make_double2(kx * iw.x - ky * iw.y, kx * iw.y + ky * iw.x);
    B[c] = make_double2(kz * iw.x, kz * iw.y);
}

__global__ __launch_bounds__(GB) void grf_sumsq(const double2* __restrict__ A, const double2* __restrict__ B, long n3,
                                                double* __restrict__ part) {
    __shared__ double red[GB / 64];
    double s = 0.0;
    for (long p = (long)blockIdx.x * GB + threadIdx.x; p < n3; p += (long)gridDim.x * GB) {
        const double2 a = A[p], b = B[p];
        s += a.x * a.x + a.y * a.y + b.x * b.x;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(GB) void grf_place(const double2* __restrict__ A, const double2* __restrict__ B, int g,
                                                double scale, double vfac, uint32_t* __restrict__ x,
                                                uint32_t* __restrict__ y, uint32_t* __restrict__ z,
                                                float* __restrict__ vx, float* __restrict__ vy, float* __restrict__ vz) {
    const long p = (long)blockIdx.x * GB + threadIdx.x;
    const int n = 1 << g;
    if (p >= (long)n * n * n) return;
    const int k = (int)(p & (n - 1)), j = (int)((p >> g) & (n - 1)), i = (int)(p >> (2 * g));
    const double h = 1.0 / n;
    const double d[3] = {scale * A[p].x, scale * A[p].y, scale * B[p].x};
    const double q[3] = {(i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h};
    uint32_t* out[3] = {x, y, z};
    float* vo[3] = {vx, vy, vz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double u = q[a] + d[a];
        u -= floor(u);
        out[a][p] = (uint32_t)(uint64_t)(int64_t)rint(u * 4294967296.0);
        vo[a][p] = (float)(vfac * d[a]);
    }
}

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" int ct2_ic_grf(int n, double rms_cells, double spectral_index, uint64_t seed, double vfac) {
    CT2_TRY(require_particles());
    Context& c = ctx();
    if (!is_pow2(n) || n < 8 || n > 512) return set_error(CT2_ERR_INVALID, "lattice side %d is not a power of two in [8, 512]", n);
    if ((size_t)n * n * n != c.n) return set_error(CT2_ERR_INVALID, "container holds %zu particles, lattice needs %d^3", c.n, n);
    if (!(rms_cells > 0.0)) return set_error(CT2_ERR_INVALID, "rms displacement must be positive");
    const long n3 = (long)n * n * n;
    const int g = ilog2(n);
    double2 *A = nullptr, *B = nullptr;
    double* part = nullptr;
    const int nb = 1024;
    int rc;
    if ((rc = dev_alloc((void**)&A, n3 * sizeof(double2), "GRF grid A")) ||
        (rc = dev_alloc((void**)&B, n3 * sizeof(double2), "GRF grid B")) ||
        (rc = dev_alloc((void**)&part, nb * sizeof(double), "GRF reduction"))) {
        dev_free(A);
        dev_free(B);
        return rc;
    }
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(c.stream);
        dev_free(A);
        dev_free(B);
        dev_free(part);
        return code;
    };
    const unsigned blocks = (unsigned)((n3 + GB - 1) / GB);
    grf_noise<<<blocks, GB, 0, c.stream>>>(A, n3, seed);
    if (hipGetLastError() != hipSuccess) return fail(set_error(CT2_ERR_HIP, "grf_noise launch failed"));
    if ((rc = fft3d(A, n, 1, true))) return fail(rc);
    grf_filter<<<blocks, GB, 0, c.stream>>>(A, B, g, spectral_index);
    if (hipGetLastError() != hipSuccess) return fail(set_error(CT2_ERR_HIP, "grf_filter launch failed"));
    if ((rc = fft3d(A, n, -1, true)) || (rc = fft3d(B, n, -1, true))) return fail(rc);
    grf_sumsq<<<nb, GB, 0, c.stream>>>(A, B, n3, part);
    std::vector<double> hp(nb);
    if (hipMemcpyAsync(hp.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return fail(set_error(CT2_ERR_HIP, "GRF reduction failed"));
    double s = 0.0;
    for (double v : hp) s += v;
    const double rms = std::sqrt(s / (3.0 * (double)n3));
    if (!(rms > 0.0)) return fail(set_error(CT2_ERR_INVALID, "degenerate displacement field"));
    const double scale = rms_cells / n / rms;
    grf_place<<<blocks, GB, 0, c.stream>>>(A, B, g, scale, vfac, c.x, c.y, c.z, c.vx, c.vy, c.vz);
    if (hipGetLastError() != hipSuccess) return fail(set_error(CT2_ERR_HIP, "grf_place launch failed"));
    return fail(CT2_OK);
}
