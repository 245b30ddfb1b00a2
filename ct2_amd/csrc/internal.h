// internal.h -- shared state and helpers of libct2hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "ct2hip.h"

namespace ct2 {

// Records the last error text and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Process-wide library state: one device, one stream, the particle
// container and the buffers the kernels share.
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    ct2_force_options opts{1.0f / 50.0f, 1.0f, CT2_KERNEL_PLUMMER};

    // particle container (SoA)
    size_t n = 0;
    uint32_t *x = nullptr, *y = nullptr, *z = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int8_t* rung = nullptr;

    // per-particle force/potential of the last stored kick
    float *fx = nullptr, *fy = nullptr, *fz = nullptr, *pot = nullptr;
    bool forces_valid = false;

    // kick scratch: split partial sums and the max-rung cell
    double* partial = nullptr;
    size_t partial_bytes = 0;
    int* maxrung = nullptr;

    // the force path the last kick took (CT2_PATH_*)
    int last_path = 0;

    // rung bookkeeping (particle.cpp:20-58): kicks act on the sinks
    // [rung_begin, n); push/pop save and restore rung_begin
    size_t rung_begin = 0;
    std::vector<size_t> rung_stack;
};

Context& ctx();

// Fails with CT2_ERR_STATE unless ct2_init succeeded.
int require_init();
// Fails with CT2_ERR_STATE unless particles are allocated.
int require_particles();

// hipMalloc with CT2 error mapping.
int dev_alloc(void** p, size_t bytes, const char* what);
void dev_free(void* p);

// Grows a scratch buffer to at least `bytes`.
int ensure_scratch(void** p, size_t* have, size_t bytes, const char* what);

// In-place 3-D complex FFT on the library stream (fft.hip).
int fft3d(void* data, int n, int dir, bool f64);
// `batch` such cubes stored one after the other, in one set of launches.
int fft3d_batch(void* data, int n, int dir, bool f64, int batch);
int fft3d_density(const unsigned long long* W, void* out, int n);

// PM force of the container into the context force arrays (pm.hip).
int pm_forces();
int pm_require();
void pm_grids(int* g, const unsigned long long** W, const float2** A, const float2** B);
// The mesh of the PM force: grid side (0 before ct2_pm_init), split radius in
// cells, assignment order, near-field cutoff in split radii (0 = off).
void pm_config(int* ng, double* rs_cells, int* order, double* rcut_rs);

// Near field of the PM force (near.hip): direct pairs closer than
// rcut_rs * rs with the softened kernel minus the mesh's filtered one, added
// to the stored force.  rs in box units.
int near_set(double rcut_rs);
bool near_enabled();
double near_rcut_rs();
int near_forces(double rs);
int near_cells();

// Periodic FP64 judge of m sinks (device index list) into d_out[4][m] (ewald.hip).
int ewald_judge(const long* d_idx, int m, double* d_out);

// Called by ct2_finalize to release module-owned buffers.
void release_kick_buffers();
// Looks up this device's copy of the no-allocation kick's static scratch
// (the device must be current); ct2_init calls it once per device.
int kick_noalloc_setup(int device);
void release_fft_plans();
void release_pm();
void release_ewald();
void release_mlp();
void release_near();
void release_fft_ct2();
void release_rungs();
void release_grid();
void release_mm();

inline bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }
inline int ilog2(long v) {
    int l = 0;
    while ((1L << l) < v) ++l;
    return l;
}

}  // namespace ct2

#define CT2_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::ct2::set_error(CT2_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                                    hipGetErrorString(e_));                                  \
    } while (0)

#define CT2_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != CT2_OK) return rc_; \
    } while (0)

// Checks the launch that was just enqueued.
#define CT2_LAUNCH_CHECK() CT2_HIP(hipGetLastError())
