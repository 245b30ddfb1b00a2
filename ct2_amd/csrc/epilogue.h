// epilogue.h -- the velocity half-kick and rung assignment of kick.cpp:75-99,
// shared by every force method (direct sum, PM, PM + learned correction).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "ct2hip.h"

namespace ct2 {

// Constants of the epilogue, computed once on the host in FP32.
struct EpiConst {
    int kick_ascending;  // ascending && !first
    int descending;
    float sgn;      // top ? +1 : -1
    float dtk;      // 0.5 t0 / 2^rung
    float eta;
    float scale_h;  // scale * hsoft
    float t0;
    float max_dt;
    float two_r;    // 2^rung
    int prung;
};

// True when the parameters ask for any velocity or rung update.
inline bool epilogue_active(const ct2_kick_params* p) { return (p->ascending && !p->first) || p->descending; }

inline EpiConst make_epilogue(const ct2_kick_params* p, float hsoft) {
#pragma clang fp contract(off)
    EpiConst e;
    const bool epi = epilogue_active(p);
    e.kick_ascending = p->ascending && !p->first;
    e.descending = p->descending;
    e.sgn = p->top ? 1.0f : -1.0f;
    e.dtk = epi ? 0.5f * p->t0 / (float)(1 << p->rung) : 0.0f;
    e.eta = p->eta;
    e.scale_h = p->scale * hsoft;
    e.t0 = p->t0;
    e.max_dt = p->max_dt;
    e.two_r = epi ? std::ldexp(1.0f, p->rung) : 1.0f;
    e.prung = p->rung;
    return e;
}

// kick.cpp:75-99 for one sink in the reference's FP32 operation order (no
// contraction).  Deviations, each following the evident intent: the force of
// sink i is used (kick.cpp:82-98 indexes a SIMD lane with the particle index),
// and ceil(log2(t0/dt)) > rung is evaluated as the equivalent exact test
// t0/dt > 2^rung.
__device__ __forceinline__ void epilogue(float f0, float f1, float f2, float& vx, float& vy, float& vz,
                                         int8_t& rung, const EpiConst& e, int& rmax) {
#pragma clang fp contract(off)
    if (e.kick_ascending) {
        vx = vx + e.sgn * f0 * e.dtk;
        vy = vy + e.sgn * f1 * e.dtk;
        vz = vz + e.sgn * f2 * e.dtk;
    }
    if (e.descending) {
        const float g2 = f0 * f0 + f1 * f1 + f2 * f2;
        float dt = fminf(e.eta * sqrtf(e.scale_h / sqrtf(g2)), e.t0);
        dt = fminf(e.max_dt, dt);
        const int up = (e.t0 / dt) > e.two_r;
        int r = rung;
        r = max(e.prung + up, r - 1);
        rung = (int8_t)r;
        rmax = max(rmax, r);
        vx = vx + e.sgn * f0 * e.dtk;
        vy = vy + e.sgn * f1 * e.dtk;
        vz = vz + e.sgn * f2 * e.dtk;
    }
}

// Validates the kick parameters the epilogue will use.
int check_kick_params(const ct2_kick_params* p);

// Launches the epilogue over the stored forces of the context (used by the
// mesh-based kicks; the direct sum fuses it into its finalize kernel).
int run_epilogue(const ct2_kick_params* p, int* max_rung);

}  // namespace ct2
