/*
 * kick.c -- FP64 oracle for the direct-sum kick, its epilogue and the drift.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* Minimum-image separation of two fixed32 coordinates, in box units.
 * sfmm's distance(x1, x0) as used at kick.cpp:44-46: the wrapped integer
 * difference read as signed.  Exact in double. */
static inline double fx_dist(uint32_t a, uint32_t b) { return (double)(int32_t)(a - b) * 0x1p-32; }

/*
 * All-pairs kick, FP64 (kick.cpp:36-74 with the bugs at :38-41 and :72
 * not reproduced: each source is counted once and the pair sum is kept).
 *
 * Plummer (the default force of the direct-sum configurations):
 *   F_i = GM sum_j dx (r^2+eps^2)^-3/2,   pot_i = GM sum_j (r^2+eps^2)^-1/2.
 * ct2 kernel (kick.cpp:47-66, uniform-density sphere of radius h):
 *   r < h: pot += 1.5/h - 0.5 r^2/h^3, F += dx/h^3;  else pot += 1/r, F += dx/r^3.
 * The self pair is included (it adds only to the potential).
 */
void orc_kick(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
              const int64_t* sinks, double gm, double soft, int kernel, double* ax, double* ay,
              double* az, double* pot) {
    const double e2 = soft * soft;
    const double h = soft, h2 = soft * soft, hinv = 1.0 / soft, hinv3 = hinv * hinv * hinv;
#pragma omp parallel for schedule(dynamic, 4)
    for (ptrdiff_t s = 0; s < (ptrdiff_t)m; s++) {
        const size_t i = sinks ? (size_t)sinks[s] : (size_t)s;
        const uint32_t xi = x[i], yi = y[i], zi = z[i];
        double sx = 0.0, sy = 0.0, sz = 0.0, sp = 0.0;
        if (kernel == ORC_KERNEL_PLUMMER) {
#pragma omp simd reduction(+ : sx, sy, sz, sp)
            for (size_t j = 0; j < n; j++) {
                const double dx = fx_dist(x[j], xi), dy = fx_dist(y[j], yi), dz = fx_dist(z[j], zi);
                const double w = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + e2);
                const double w3 = w * w * w;
                sx += dx * w3;
                sy += dy * w3;
                sz += dz * w3;
                sp += w;
            }
        } else {
#pragma omp simd reduction(+ : sx, sy, sz, sp)
            for (size_t j = 0; j < n; j++) {
                const double dx = fx_dist(x[j], xi), dy = fx_dist(y[j], yi), dz = fx_dist(z[j], zi);
                const double r2 = dx * dx + dy * dy + dz * dz;
                double k, p;
                if (r2 < h2) {
                    k = hinv3;
                    p = 1.5 * hinv - 0.5 * r2 * hinv3;
                } else {
                    const double w = 1.0 / sqrt(r2);
                    k = w * w * w;
                    p = w;
                }
                sx += dx * k;
                sy += dy * k;
                sz += dz * k;
                sp += p;
            }
        }
        (void)h;
        ax[s] = gm * sx;
        ay[s] = gm * sy;
        az[s] = gm * sz;
        pot[s] = gm * sp;
    }
}

/* The same sum in long double with a plain serial loop; used to pin the
 * vectorised FP64 sum above on tiny inputs. */
void orc_kick_ld(size_t n, const uint32_t* x, const uint32_t* y, const uint32_t* z, size_t m,
                 const int64_t* sinks, double gm, double soft, int kernel, double* ax,
                 double* ay, double* az, double* pot) {
    const long double h = soft;
    for (size_t s = 0; s < m; s++) {
        const size_t i = sinks ? (size_t)sinks[s] : s;
        long double sx = 0, sy = 0, sz = 0, sp = 0;
        for (size_t j = 0; j < n; j++) {
            const long double dx = (long double)(int32_t)(x[j] - x[i]) / 4294967296.0L;
            const long double dy = (long double)(int32_t)(y[j] - y[i]) / 4294967296.0L;
            const long double dz = (long double)(int32_t)(z[j] - z[i]) / 4294967296.0L;
            const long double r2 = dx * dx + dy * dy + dz * dz;
            long double k, p;
            if (kernel == ORC_KERNEL_PLUMMER) {
                const long double d = sqrtl(r2 + h * h);
                k = 1.0L / (d * d * d);
                p = 1.0L / d;
            } else if (r2 < h * h) {
                k = 1.0L / (h * h * h);
                p = 1.5L / h - 0.5L * r2 / (h * h * h);
            } else {
                const long double r = sqrtl(r2);
                k = 1.0L / (r * r * r);
                p = 1.0L / r;
            }
            sx += dx * k;
            sy += dy * k;
            sz += dz * k;
            sp += p;
        }
        ax[s] = (double)(gm * sx);
        ay[s] = (double)(gm * sy);
        az[s] = (double)(gm * sz);
        pot[s] = (double)(gm * sp);
    }
}

/*
 * Velocity half-kick and rung assignment, kick.cpp:75-99, in FP32 as the
 * reference computes it (this file is built with -ffp-contract=off so the
 * operation order below is the rounding order).
 *
 * Deviations, each following the evident intent:
 *  - the force of sink i is fx[i] (kick.cpp:82-98 indexes a SIMD lane with
 *    the particle index);
 *  - ceil(log2(t0/dt)) > rung is evaluated as the equivalent exact test
 *    t0/dt > 2^rung, so no libm rounding decides a rung.
 * Returns the maximum rung assigned (0 when descending is false).
 */
int orc_kick_epilogue(size_t n, const float* fx, const float* fy, const float* fz, float* vx,
                      float* vy, float* vz, int8_t* rung, const orc_kick_params* p, float hsoft) {
    int max_rung = 0;
    const float sgn = p->top ? 1.f : -1.f;
    const float dtk = 0.5f * p->t0 / (float)(1 << p->rung);
    const float two_r = ldexpf(1.0f, p->rung);
    for (size_t i = 0; i < n; i++) {
        const float f0 = fx[i], f1 = fy[i], f2 = fz[i];
        if (p->ascending && !p->first) {
            vx[i] = vx[i] + sgn * f0 * dtk;
            vy[i] = vy[i] + sgn * f1 * dtk;
            vz[i] = vz[i] + sgn * f2 * dtk;
        }
        if (p->descending) {
            const float g2 = f0 * f0 + f1 * f1 + f2 * f2;
            float dt = fminf(p->eta * sqrtf(p->scale * hsoft / sqrtf(g2)), p->t0);
            dt = fminf(p->max_dt, dt);
            const int up = (p->t0 / dt) > two_r;
            int r = rung[i];
            r = (p->rung + up) > (r - 1) ? (p->rung + up) : (r - 1);
            rung[i] = (int8_t)r;
            if (r > max_rung) max_rung = r;
            vx[i] = vx[i] + sgn * f0 * dtk;
            vy[i] = vy[i] + sgn * f1 * dtk;
            vz[i] = vz[i] + sgn * f2 * dtk;
        }
    }
    return max_rung;
}

/*
 * float -> fixed32 conversion used by the drift (drift.cpp:12-14,
 * simd_fixed32(float)).  sfmm's rounding is not in the reference tree, so it
 * is defined here: RNE(s * 2^32) mod 2^32 for finite s, and 0 for NaN/inf.
 */
uint32_t orc_fixed32_from_float(float s) {
    if (!isfinite(s)) return 0u;
    if (fabsf(s) >= 4294967296.0f) return 0u; /* s*2^32 is a multiple of 2^41 */
    const float t = rintf(s * 4294967296.0f);
    if (fabsf(t) < 2147483648.0f) return (uint32_t)(int32_t)t;
    /* |t| >= 2^31: t is a multiple of 2^8, so the reduction below is exact */
    const float q = floorf(t * 0x1p-32f);
    const float r = t - q * 4294967296.0f;
    return (uint32_t)r;
}

/* drift.cpp:4-17: lambda = dt/scale; x += fixed32(v*lambda) per axis. */
void orc_drift(size_t n, uint32_t* x, uint32_t* y, uint32_t* z, const float* vx,
               const float* vy, const float* vz, float dt, float scale) {
    const float lambda = dt / scale;
    for (size_t i = 0; i < n; i++) {
        x[i] += orc_fixed32_from_float(vx[i] * lambda);
        y[i] += orc_fixed32_from_float(vy[i] * lambda);
        z[i] += orc_fixed32_from_float(vz[i] * lambda);
    }
}

/* Thread count of the OpenMP loops (for timing the oracle on 1 or all cores). */
void orc_set_threads(int t) {
#ifdef _OPENMP
    omp_set_num_threads(t > 0 ? t : 1);
#else
    (void)t;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
