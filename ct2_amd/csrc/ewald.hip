// ewald.hip -- the periodic FP64 judge: direct sum over every source with all
// periodic images and a uniform neutralising background (Ewald summation),
// the force ct2's Green's function intends (particle.cpp:141-144,
// greens + greens_ewald), with Plummer softening of the nearest image.
//
// With alpha = 10 (box units) the real-space term of every image but the
// minimum one is below erfc(5) = 1.5e-12 of its Newtonian value, so:
//   real space   one FP64 pass over all sources per sink (workgroup per sink,
//                wavefront-shuffle + LDS reduction, fixed order):
//                F += d [ (s^2+eps^2)^-3/2 - (erf(a s) - 2 a s/sqrt(pi) e^{-a^2 s^2}) / s^3 ]
//                P += (s^2+eps^2)^-1/2 - erf(a s)/s
//   k space      structure factors T(k) = sum_j GM e^{2 pi i k.x_j} over the
//                half-space |k|^2 <= 330 (e^{-pi^2 k^2/a^2} < 1e-14 beyond),
//                phases reduced exactly in 32-bit integer arithmetic; then per
//                sink F += 4 sum g_k k Im[e^{-2 pi i k.x_i} T(k)],
//                P += (2/pi) sum g_k Re[...], g_k = e^{-pi^2 k^2/a^2}/k^2.
//   constant     P -= GM N pi/a^2.
#include <cmath>
#include <vector>

#include "internal.h"

namespace ct2 {
namespace {

constexpr double EW_ALPHA = 10.0;
constexpr int EW_K2MAX = 330;
constexpr int KZB = 8;          // kz values per structure-factor task
constexpr int SF_CHUNK = 65536;  // particles per structure-factor chunk

struct KTask {
    int kx, ky, kz0, nkz, first;  // first: index of (kx,ky,kz0) in the flat k list
};

struct EwaldState {
    int nk = 0, ntask = 0;
    int4* klist = nullptr;      // (kx, ky, kz, -) flat, half-space
    double* gk = nullptr;       // g_k per flat k
    KTask* tasks = nullptr;
    double2* partial = nullptr;  // [chunk][nk]
    size_t partial_bytes = 0;
    double2* T = nullptr;        // [nk]
} g_ew;

__device__ __forceinline__ void sincos2pi(uint32_t theta, double& s, double& c) {
    // angle = 2 pi theta / 2^32
    sincospi((double)theta * 0x1p-31, &s, &c);
}

// grid: (ntask, nchunk); block sums particles of one chunk for one task.
__global__ __launch_bounds__(256) void ewald_structure(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                       const uint32_t* __restrict__ Z, int n, const KTask* __restrict__ tasks,
                                                       int nk, double gm, double2* __restrict__ partial) {
    __shared__ double red[4][KZB][2];
    const KTask t = tasks[blockIdx.x];
    const int chunk = blockIdx.y;
    const int p0 = chunk * SF_CHUNK, p1 = min(n, p0 + SF_CHUNK);
    double re[KZB], im[KZB];
#pragma unroll
    for (int q = 0; q < KZB; ++q) re[q] = im[q] = 0.0;
    for (int p = p0 + threadIdx.x; p < p1; p += 256) {
        const uint32_t x = X[p], y = Y[p], z = Z[p];
        const uint32_t th0 = (uint32_t)t.kx * x + (uint32_t)t.ky * y + (uint32_t)t.kz0 * z;
        double s, c, ss, cs;
        sincos2pi(th0, s, c);
        sincos2pi(z, ss, cs);
#pragma unroll
        for (int q = 0; q < KZB; ++q) {
            if (q < t.nkz) {
                re[q] += c;
                im[q] += s;
                const double c2 = c * cs - s * ss, s2 = s * cs + c * ss;
                c = c2;
                s = s2;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < KZB; ++q) {
        double a = re[q], b = im[q];
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off);
            b += __shfl_xor(b, off);
        }
        if (lane == 0) {
            red[wave][q][0] = a;
            red[wave][q][1] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x < t.nkz) {
        const int q = threadIdx.x;
        const double a = (red[0][q][0] + red[1][q][0]) + (red[2][q][0] + red[3][q][0]);
        const double b = (red[0][q][1] + red[1][q][1]) + (red[2][q][1] + red[3][q][1]);
        partial[(size_t)chunk * nk + t.first + q] = make_double2(gm * a, gm * b);
    }
}

__global__ __launch_bounds__(256) void ewald_reduce(const double2* __restrict__ partial, int nchunk, int nk,
                                                    double2* __restrict__ T) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= nk) return;
    double a = 0.0, b = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        const double2 v = partial[(size_t)c * nk + k];
        a += v.x;
        b += v.y;
    }
    T[k] = make_double2(a, b);
}

template <typename T>
__device__ __forceinline__ T block_sum256(T v, T* red /* [4] */) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Real-space minimum-image part plus the k-space sum, one workgroup per sink.
__global__ __launch_bounds__(256) void ewald_sink(const uint32_t* __restrict__ X, const uint32_t* __restrict__ Y,
                                                  const uint32_t* __restrict__ Z, int n, const long* __restrict__ idx,
                                                  int m, double soft, double gm, const int4* __restrict__ klist,
                                                  const double* __restrict__ gk, const double2* __restrict__ T, int nk,
                                                  double* __restrict__ out) {
    __shared__ double red[4];
    const int s = blockIdx.x;
    const long i = idx[s];
    const uint32_t xi = X[i], yi = Y[i], zi = Z[i];
    const double a = EW_ALPHA, e2 = soft * soft, tsp = 2.0 / sqrt(M_PI);
    double fx = 0.0, fy = 0.0, fz = 0.0, fp = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        const double dx = (double)(int32_t)(X[j] - xi) * 0x1p-32;
        const double dy = (double)(int32_t)(Y[j] - yi) * 0x1p-32;
        const double dz = (double)(int32_t)(Z[j] - zi) * 0x1p-32;
        const double s2 = dx * dx + dy * dy + dz * dz;
        const double w = 1.0 / sqrt(s2 + e2);
        const double sr = sqrt(s2), u = a * sr, u2 = u * u;
        double b, q;
        if (u < 1e-2) {  // series of the two regular combinations
            b = tsp * a * a * a * (2.0 / 3.0 - 0.4 * u2 + u2 * u2 / 7.0);
            q = tsp * a * (1.0 - u2 / 3.0 + u2 * u2 / 10.0);
        } else {
            const double ef = erf(u);
            b = (ef - tsp * u * exp(-u2)) / (s2 * sr);
            q = ef / sr;
        }
        const double k = w * w * w - b;
        fx += dx * k;
        fy += dy * k;
        fz += dz * k;
        fp += w - q;
    }
    // k space: 2 pi k.x_i reduced exactly mod 1 in uint32
    double kx = 0.0, ky = 0.0, kz = 0.0, kp = 0.0;
    for (int t = threadIdx.x; t < nk; t += 256) {
        const int4 kv = klist[t];
        const uint32_t th = (uint32_t)kv.x * xi + (uint32_t)kv.y * yi + (uint32_t)kv.z * zi;
        double sn, cs;
        sincospi((double)(0u - th) * 0x1p-31, &sn, &cs);  // e^{-2 pi i k.x_i}
        const double2 tk = T[t];
        const double im = sn * tk.x + cs * tk.y;  // Im[(cs + i sn)(T.x + i T.y)]
        const double re = cs * tk.x - sn * tk.y;
        const double g = gk[t];
        kx += g * kv.x * im;
        ky += g * kv.y * im;
        kz += g * kv.z * im;
        kp += g * re;
    }
    const double FX = block_sum256(fx, red), FY = block_sum256(fy, red), FZ = block_sum256(fz, red),
                 FP = block_sum256(fp, red);
    const double KX = block_sum256(kx, red), KY = block_sum256(ky, red), KZ = block_sum256(kz, red),
                 KP = block_sum256(kp, red);
    if (threadIdx.x == 0) {
        out[0 * (long)m + s] = gm * FX + 4.0 * KX;
        out[1 * (long)m + s] = gm * FY + 4.0 * KY;
        out[2 * (long)m + s] = gm * FZ + 4.0 * KZ;
        out[3 * (long)m + s] = gm * (FP - (double)n * M_PI / (a * a)) + (2.0 / M_PI) * KP;
    }
}

int ewald_setup() {
    if (g_ew.nk) return CT2_OK;
    std::vector<int4> kl;
    std::vector<double> gv;
    std::vector<KTask> tasks;
    const int kmax = (int)std::floor(std::sqrt((double)EW_K2MAX));
    for (int kx = 0; kx <= kmax; ++kx)
        for (int ky = -kmax; ky <= kmax; ++ky) {
            if (kx == 0 && ky < 0) continue;
            const int rem = EW_K2MAX - kx * kx - ky * ky;
            if (rem < 0) continue;
            const int kzm = (int)std::floor(std::sqrt((double)rem));
            const int lo = (kx == 0 && ky == 0) ? 1 : -kzm;
            for (int kz0 = lo; kz0 <= kzm; kz0 += KZB) {
                KTask t{kx, ky, kz0, std::min(KZB, kzm - kz0 + 1), (int)kl.size()};
                for (int q = 0; q < t.nkz; ++q) {
                    const int kz = kz0 + q;
                    const double k2 = (double)(kx * kx + ky * ky + kz * kz);
                    kl.push_back(make_int4(kx, ky, kz, 0));
                    gv.push_back(std::exp(-M_PI * M_PI * k2 / (EW_ALPHA * EW_ALPHA)) / k2);
                }
                tasks.push_back(t);
            }
        }
    g_ew.nk = (int)kl.size();
    g_ew.ntask = (int)tasks.size();
    hipStream_t st = ctx().stream;
    int rc;
    if ((rc = dev_alloc((void**)&g_ew.klist, kl.size() * sizeof(int4), "Ewald k list")) ||
        (rc = dev_alloc((void**)&g_ew.gk, gv.size() * sizeof(double), "Ewald weights")) ||
        (rc = dev_alloc((void**)&g_ew.tasks, tasks.size() * sizeof(KTask), "Ewald tasks")) ||
        (rc = dev_alloc((void**)&g_ew.T, kl.size() * sizeof(double2), "Ewald structure factors"))) {
        g_ew.nk = 0;
        return rc;
    }
    CT2_HIP(hipMemcpyAsync(g_ew.klist, kl.data(), kl.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    CT2_HIP(hipMemcpyAsync(g_ew.gk, gv.data(), gv.size() * sizeof(double), hipMemcpyHostToDevice, st));
    CT2_HIP(hipMemcpyAsync(g_ew.tasks, tasks.data(), tasks.size() * sizeof(KTask), hipMemcpyHostToDevice, st));
    CT2_HIP(hipStreamSynchronize(st));
    return CT2_OK;
}

}  // namespace

// Periodic FP64 judge of m sinks into out[4][m] (device).
int ewald_judge(const long* d_idx, int m, double* d_out) {
    CT2_TRY(ewald_setup());
    Context& c = ctx();
    const int n = (int)c.n;
    const int nchunk = (n + SF_CHUNK - 1) / SF_CHUNK;
    CT2_TRY(ensure_scratch((void**)&g_ew.partial, &g_ew.partial_bytes, sizeof(double2) * (size_t)nchunk * g_ew.nk,
                           "Ewald partial sums"));
    const double gm = c.opts.GM, soft = c.opts.hsoft;
    ewald_structure<<<dim3(g_ew.ntask, nchunk), 256, 0, c.stream>>>(c.x, c.y, c.z, n, g_ew.tasks, g_ew.nk, gm,
                                                                     g_ew.partial);
    CT2_LAUNCH_CHECK();
    ewald_reduce<<<(g_ew.nk + 255) / 256, 256, 0, c.stream>>>(g_ew.partial, nchunk, g_ew.nk, g_ew.T);
    CT2_LAUNCH_CHECK();
    ewald_sink<<<m, 256, 0, c.stream>>>(c.x, c.y, c.z, n, d_idx, m, soft, gm, g_ew.klist, g_ew.gk, g_ew.T, g_ew.nk,
                                        d_out);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}

void release_ewald() {
    dev_free(g_ew.klist);
    dev_free(g_ew.gk);
    dev_free(g_ew.tasks);
    dev_free(g_ew.partial);
    dev_free(g_ew.T);
    g_ew = EwaldState{};
}

}  // namespace ct2
