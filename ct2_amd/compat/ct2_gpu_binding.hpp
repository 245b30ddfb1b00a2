// ct2_gpu_binding.hpp -- the reference-side binding: ct2's own free-function
// signatures (kick.hpp:10-24, drift.hpp:12, particle.hpp:22-58,
// particle.cpp:273) implemented over the C ABI of libct2hip.so.
//
// A ct2 build links this file instead of src/kick.cpp and src/drift.cpp.
// The particle globals stay host arrays, as ct2's accessors expect; each call
// uploads the arrays it reads and downloads the ones it writes.  Code that
// keeps particles resident on the device calls the ct2_* entry points directly.
#pragma once

#include <cstddef>
#include <cstdint>

// sfmm::fixed32 is a 32-bit unsigned fixed-point position (cosmictiger.hpp:22).
using ct2_pos_fixed = uint32_t;

struct kick_params {  // kick.hpp:10-20
    bool ascending;
    bool descending;
    bool top;
    bool first;
    int rung;
    float t0;
    float eta;
    float max_dt;
    float scale;
};

// particle.hpp:22-28
extern ct2_pos_fixed* particle_x_ptr;
extern ct2_pos_fixed* particle_y_ptr;
extern ct2_pos_fixed* particle_z_ptr;
extern float* particle_vx_ptr;
extern float* particle_vy_ptr;
extern float* particle_vz_ptr;
extern char* particle_rung_ptr;

void particle_allocate(size_t count);  // particle.cpp:273
size_t particle_count();               // particle.cpp:27
int kick(const kick_params& params);   // kick.hpp:24
void drift(float dt, float scale);     // drift.hpp:12

// Options the force path reads from get_options() (options.cpp:155,190).
void ct2_gpu_set_options(float hsoft, float GM);
