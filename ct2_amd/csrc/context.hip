// context.hip -- device binding, stream, options and the SoA particle
// container (particle.hpp:22-58, particle.cpp:17-29, 273-288).
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "internal.h"

namespace ct2 {

static char g_err[1024] = "";
static Context g_ctx;

Context& ctx() { return g_ctx; }

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int require_init() {
    if (g_ctx.device < 0) return set_error(CT2_ERR_STATE, "ct2_init has not been called");
    return CT2_OK;
}

int require_particles() {
    CT2_TRY(require_init());
    if (g_ctx.n == 0) return set_error(CT2_ERR_STATE, "no particles allocated");
    return CT2_OK;
}

static int g_fail_after = 0, g_fail_allocs = 0;  // ct2_debug_fail_allocations

int dev_alloc(void** p, size_t bytes, const char* what) {
    *p = nullptr;
    if (bytes == 0) return CT2_OK;
    if (g_fail_allocs > 0 && g_fail_after > 0) {
        --g_fail_after;
    } else if (g_fail_allocs > 0) {
        --g_fail_allocs;
        return set_error(CT2_ERR_NOMEM, "device allocation of %zu bytes for %s failed: injected by the test hook",
                         bytes, what);
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return set_error(CT2_ERR_NOMEM, "device allocation of %zu bytes for %s failed: %s", bytes,
                         what, hipGetErrorString(e));
    }
    return CT2_OK;
}

void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

int ensure_scratch(void** p, size_t* have, size_t bytes, const char* what) {
    if (*have >= bytes && *p) return CT2_OK;
    dev_free(*p);
    *p = nullptr;
    *have = 0;
    CT2_TRY(dev_alloc(p, bytes, what));
    *have = bytes;
    return CT2_OK;
}

static void free_particles() {
    Context& c = g_ctx;
    for (void* p : {(void*)c.x, (void*)c.y, (void*)c.z, (void*)c.vx, (void*)c.vy, (void*)c.vz,
                    (void*)c.rung, (void*)c.fx, (void*)c.fy, (void*)c.fz, (void*)c.pot})
        dev_free(p);
    c.x = c.y = c.z = nullptr;
    c.vx = c.vy = c.vz = nullptr;
    c.rung = nullptr;
    c.fx = c.fy = c.fz = c.pot = nullptr;
    c.n = 0;
    c.forces_valid = false;
}

}  // namespace ct2

using namespace ct2;

extern "C" {

const char* ct2_last_error(void) { return g_err; }

int ct2_init(int device) {
    Context& c = g_ctx;
    if (c.device >= 0) {
        if (c.device == device) return CT2_OK;
        return set_error(CT2_ERR_STATE, "already initialised on device %d", c.device);
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return set_error(CT2_ERR_NODEVICE, "no HIP device available (%s)",
                         e == hipSuccess ? "count 0" : hipGetErrorString(e));
    }
    if (device < 0 || device >= count)
        return set_error(CT2_ERR_INVALID, "device %d out of range [0,%d)", device, count);
    hipDeviceProp_t prop;
    CT2_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(CT2_ERR_NODEVICE, "device %d is %s; this library is built for gfx950",
                         device, prop.gcnArchName);
    CT2_HIP(hipSetDevice(device));
    CT2_TRY(kick_noalloc_setup(device));
    CT2_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.device = device;
    return CT2_OK;
}

int ct2_finalize(void) {
    Context& c = g_ctx;
    if (c.device < 0) return CT2_OK;
    (void)hipStreamSynchronize(c.stream);
    free_particles();
    release_kick_buffers();
    release_fft_plans();
    release_pm();
    release_ewald();
    release_mlp();
    release_near();
    release_fft_ct2();
    release_rungs();
    release_grid();
    release_mm();
    (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    c.device = -1;
    return CT2_OK;
}

int ct2_get_stream(void** stream) {
    CT2_TRY(require_init());
    if (!stream) return set_error(CT2_ERR_INVALID, "null stream pointer");
    *stream = (void*)g_ctx.stream;
    return CT2_OK;
}

int ct2_synchronize(void) {
    CT2_TRY(require_init());
    CT2_HIP(hipStreamSynchronize(g_ctx.stream));
    return CT2_OK;
}

int ct2_set_force_options(const ct2_force_options* opts) {
    if (!opts) return set_error(CT2_ERR_INVALID, "null options");
    if (!(opts->hsoft > 0.0f) || !(opts->hsoft < 0.5f))
        return set_error(CT2_ERR_INVALID, "hsoft must be in (0, 0.5), got %g", (double)opts->hsoft);
    if (!(opts->GM > 0.0f)) return set_error(CT2_ERR_INVALID, "GM must be positive");
    if (opts->kernel != CT2_KERNEL_PLUMMER && opts->kernel != CT2_KERNEL_CT2)
        return set_error(CT2_ERR_INVALID, "unknown kernel %d", opts->kernel);
    g_ctx.opts = *opts;
    return CT2_OK;
}

int ct2_particles_alloc(size_t n) {
    CT2_TRY(require_init());
    if (n == 0) return set_error(CT2_ERR_INVALID, "particle count must be positive");
    if (n > (size_t)INT32_MAX) return set_error(CT2_ERR_INVALID, "particle count %zu exceeds 2^31-1", n);
    free_particles();
    Context& c = g_ctx;
    const size_t b4 = n * 4;
    int rc = CT2_OK;
    if ((rc = dev_alloc((void**)&c.x, b4, "x")) || (rc = dev_alloc((void**)&c.y, b4, "y")) ||
        (rc = dev_alloc((void**)&c.z, b4, "z")) || (rc = dev_alloc((void**)&c.vx, b4, "vx")) ||
        (rc = dev_alloc((void**)&c.vy, b4, "vy")) || (rc = dev_alloc((void**)&c.vz, b4, "vz")) ||
        (rc = dev_alloc((void**)&c.rung, n, "rung")) || (rc = dev_alloc((void**)&c.fx, b4, "fx")) ||
        (rc = dev_alloc((void**)&c.fy, b4, "fy")) || (rc = dev_alloc((void**)&c.fz, b4, "fz")) ||
        (rc = dev_alloc((void**)&c.pot, b4, "pot"))) {
        free_particles();
        return rc;
    }
    c.n = n;
    c.rung_begin = 0;
    c.rung_stack.clear();
    for (void* p : {(void*)c.x, (void*)c.y, (void*)c.z, (void*)c.vx, (void*)c.vy, (void*)c.vz,
                    (void*)c.fx, (void*)c.fy, (void*)c.fz, (void*)c.pot})
        CT2_HIP(hipMemsetAsync(p, 0, b4, c.stream));
    CT2_HIP(hipMemsetAsync(c.rung, 0, n, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

size_t ct2_particle_count(void) { return g_ctx.n; }

int ct2_particles_upload(const uint32_t* x, const uint32_t* y, const uint32_t* z, const float* vx,
                         const float* vy, const float* vz, const int8_t* rung) {
    CT2_TRY(require_particles());
    Context& c = g_ctx;
    const size_t b4 = c.n * 4;
    const void* src[7] = {x, y, z, vx, vy, vz, rung};
    void* dst[7] = {c.x, c.y, c.z, c.vx, c.vy, c.vz, c.rung};
    for (int a = 0; a < 7; ++a)
        if (src[a]) CT2_HIP(hipMemcpyAsync(dst[a], src[a], a == 6 ? c.n : b4, hipMemcpyHostToDevice, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

int ct2_particles_download(uint32_t* x, uint32_t* y, uint32_t* z, float* vx, float* vy, float* vz,
                           int8_t* rung) {
    CT2_TRY(require_particles());
    Context& c = g_ctx;
    const size_t b4 = c.n * 4;
    void* dst[7] = {x, y, z, vx, vy, vz, rung};
    const void* src[7] = {c.x, c.y, c.z, c.vx, c.vy, c.vz, c.rung};
    for (int a = 0; a < 7; ++a)
        if (dst[a]) CT2_HIP(hipMemcpyAsync(dst[a], src[a], a == 6 ? c.n : b4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

int ct2_host_alloc(size_t bytes, void** out) {
    if (!out) return set_error(CT2_ERR_INVALID, "null output");
    *out = nullptr;
    if (bytes == 0) return CT2_OK;
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();
        return set_error(CT2_ERR_NOMEM, "pinned host allocation of %zu bytes failed", bytes);
    }
    *out = p;
    return CT2_OK;
}

int ct2_host_free(void* p) {
    if (p) CT2_HIP(hipHostFree(p));
    return CT2_OK;
}

int ct2_particles_device_ptrs(ct2_particle_ptrs* out) {
    CT2_TRY(require_particles());
    if (!out) return set_error(CT2_ERR_INVALID, "null output");
    Context& c = g_ctx;
    *out = ct2_particle_ptrs{c.x, c.y, c.z, c.vx, c.vy, c.vz, c.rung, c.n};
    return CT2_OK;
}

int ct2_forces_download(float* fx, float* fy, float* fz, float* pot) {
    CT2_TRY(require_particles());
    Context& c = g_ctx;
    if (!c.forces_valid) return set_error(CT2_ERR_STATE, "no stored forces: run ct2_kick with CT2_KICK_STORE");
    const size_t b4 = c.n * 4;
    float* dst[4] = {fx, fy, fz, pot};
    const float* src[4] = {c.fx, c.fy, c.fz, c.pot};
    for (int a = 0; a < 4; ++a)
        if (dst[a]) CT2_HIP(hipMemcpyAsync(dst[a], src[a], b4, hipMemcpyDeviceToHost, c.stream));
    CT2_HIP(hipStreamSynchronize(c.stream));
    return CT2_OK;
}

int ct2_last_force_path(void) { return g_ctx.last_path; }

int ct2_debug_fail_allocations(int after, int count) {
    if (after < 0 || count < 0) return set_error(CT2_ERR_INVALID, "negative count");
    g_fail_after = after;
    g_fail_allocs = count;
    return CT2_OK;
}

int ct2_dev_alloc(size_t bytes, void** ptr) {
    CT2_TRY(require_init());
    if (!ptr || bytes == 0) return set_error(CT2_ERR_INVALID, "bad allocation request");
    return dev_alloc(ptr, bytes, "caller buffer");
}

int ct2_dev_free(void* ptr) {
    CT2_TRY(require_init());
    CT2_HIP(hipStreamSynchronize(g_ctx.stream));
    dev_free(ptr);
    return CT2_OK;
}

int ct2_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    CT2_TRY(require_init());
    if (!dst || !src) return set_error(CT2_ERR_INVALID, "null pointer");
    CT2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_ctx.stream));
    CT2_HIP(hipStreamSynchronize(g_ctx.stream));
    return CT2_OK;
}

int ct2_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    CT2_TRY(require_init());
    if (!dst || !src) return set_error(CT2_ERR_INVALID, "null pointer");
    CT2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
    CT2_HIP(hipStreamSynchronize(g_ctx.stream));
    return CT2_OK;
}

}  // extern "C"
