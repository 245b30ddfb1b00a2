// ct2_gpu_binding.cpp -- see ct2_gpu_binding.hpp.
#include "ct2_gpu_binding.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ct2hip.h"

ct2_pos_fixed* particle_x_ptr = nullptr;
ct2_pos_fixed* particle_y_ptr = nullptr;
ct2_pos_fixed* particle_z_ptr = nullptr;
float* particle_vx_ptr = nullptr;
float* particle_vy_ptr = nullptr;
float* particle_vz_ptr = nullptr;
char* particle_rung_ptr = nullptr;

namespace {
size_t g_count = 0;
std::vector<char> g_store;

// ct2 aborts on allocation failure (memory.cpp:7-10); the binding does the
// same on any library error, with the library's message.
void check(int rc, const char* what) {
    if (rc != CT2_OK) {
        std::fprintf(stderr, "ct2 GPU path: %s failed: %s\n", what, ct2_last_error());
        std::abort();
    }
}

void ensure_init() {
    static bool done = false;
    if (!done) {
        check(ct2_init(0), "ct2_init");
        done = true;
    }
}
}  // namespace

void ct2_gpu_set_options(float hsoft, float GM) {
    ensure_init();
    ct2_force_options o{hsoft, GM, CT2_KERNEL_PLUMMER};
    check(ct2_set_force_options(&o), "ct2_set_force_options");
}

void particle_allocate(size_t count) {
    ensure_init();
    g_count = count;
    // one block, sizeof(T) * count per array (particle.cpp:275-283 intent)
    g_store.assign(count * (3 * sizeof(ct2_pos_fixed) + 3 * sizeof(float) + 1), 0);
    char* p = g_store.data();
    particle_x_ptr = (ct2_pos_fixed*)p;
    particle_y_ptr = particle_x_ptr + count;
    particle_z_ptr = particle_y_ptr + count;
    particle_vx_ptr = (float*)(particle_z_ptr + count);
    particle_vy_ptr = particle_vx_ptr + count;
    particle_vz_ptr = particle_vy_ptr + count;
    particle_rung_ptr = (char*)(particle_vz_ptr + count);
    check(ct2_particles_alloc(count), "ct2_particles_alloc");
}

size_t particle_count() { return g_count; }

int kick(const kick_params& p) {
    check(ct2_particles_upload(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                               particle_vz_ptr, (const int8_t*)particle_rung_ptr),
          "upload");
    const ct2_kick_params q{p.ascending, p.descending, p.top, p.first, p.rung, p.t0, p.eta, p.max_dt, p.scale};
    int max_rung = 0;
    check(ct2_kick(&q, CT2_KICK_STORE, &max_rung), "ct2_kick");
    check(ct2_particles_download(nullptr, nullptr, nullptr, particle_vx_ptr, particle_vy_ptr, particle_vz_ptr,
                                 (int8_t*)particle_rung_ptr),
          "download");
    return max_rung;
}

void drift(float dt, float scale) {
    check(ct2_particles_upload(particle_x_ptr, particle_y_ptr, particle_z_ptr, particle_vx_ptr, particle_vy_ptr,
                               particle_vz_ptr, nullptr),
          "upload");
    check(ct2_drift(dt, scale), "ct2_drift");
    check(ct2_particles_download(particle_x_ptr, particle_y_ptr, particle_z_ptr, nullptr, nullptr, nullptr, nullptr),
          "download");
}
