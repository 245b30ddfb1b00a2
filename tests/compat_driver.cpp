// Drives the ct2-signature binding (kick(), drift(), particle globals) the way
// ct2's kick_step would (main.cpp:62-81) and prints the resulting state, for
// tests/test_gpu_compat.py to compare with the CPU oracle.
#include <cstdio>
#include <cstdlib>

#include "ct2_gpu_binding.hpp"

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 512;
    particle_allocate(n);
    ct2_gpu_set_options(0.01f, 1.0f / n);
    unsigned s = 12345u;
    for (int i = 0; i < n; i++) {
        s = s * 1664525u + 1013904223u;
        particle_x_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_y_ptr[i] = s;
        s = s * 1664525u + 1013904223u;
        particle_z_ptr[i] = s;
        particle_vx_ptr[i] = particle_vy_ptr[i] = particle_vz_ptr[i] = 0.0f;
        particle_rung_ptr[i] = 0;
    }
    kick_params kp{true, true, true, false, 0, 0.01f, 0.15f, 0.01f, 1.0f};
    const int max_rung = kick(kp);
    drift(0.01f, 1.0f);
    std::printf("max_rung %d\n", max_rung);
    for (int i = 0; i < n; i++)
        std::printf("%u %u %u %.9e %.9e %.9e %d\n", particle_x_ptr[i], particle_y_ptr[i], particle_z_ptr[i],
                    particle_vx_ptr[i], particle_vy_ptr[i], particle_vz_ptr[i], (int)particle_rung_ptr[i]);
    return 0;
}
