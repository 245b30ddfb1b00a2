// drift.hip -- the position drift of drift.cpp:4-18 as one HBM-streaming
// kernel: 36 bytes per particle (read x,y,z and vx,vy,vz, write x,y,z).
#include "fixed32.h"
#include "internal.h"

namespace ct2 {
namespace {

constexpr int DRIFT_BLOCK = 256;

// Four particles per thread through 16-byte loads and stores; the tail
// (n % 4) is done by the first threads with scalar accesses.
__global__ __launch_bounds__(DRIFT_BLOCK) void drift_kernel(uint32_t* __restrict__ x, uint32_t* __restrict__ y,
                                                           uint32_t* __restrict__ z, const float* __restrict__ vx,
                                                           const float* __restrict__ vy, const float* __restrict__ vz,
                                                           long n, float lambda) {
#pragma clang fp contract(off)
    const long q = (long)blockIdx.x * DRIFT_BLOCK + threadIdx.x;
    const long n4 = n >> 2;
    if (q < n4) {
        uint4 X = reinterpret_cast<uint4*>(x)[q];
        uint4 Y = reinterpret_cast<uint4*>(y)[q];
        uint4 Z = reinterpret_cast<uint4*>(z)[q];
        const float4 U = reinterpret_cast<const float4*>(vx)[q];
        const float4 V = reinterpret_cast<const float4*>(vy)[q];
        const float4 W = reinterpret_cast<const float4*>(vz)[q];
        X.x += fixed32_from_float(U.x * lambda);
        X.y += fixed32_from_float(U.y * lambda);
        X.z += fixed32_from_float(U.z * lambda);
        X.w += fixed32_from_float(U.w * lambda);
        Y.x += fixed32_from_float(V.x * lambda);
        Y.y += fixed32_from_float(V.y * lambda);
        Y.z += fixed32_from_float(V.z * lambda);
        Y.w += fixed32_from_float(V.w * lambda);
        Z.x += fixed32_from_float(W.x * lambda);
        Z.y += fixed32_from_float(W.y * lambda);
        Z.z += fixed32_from_float(W.z * lambda);
        Z.w += fixed32_from_float(W.w * lambda);
        reinterpret_cast<uint4*>(x)[q] = X;
        reinterpret_cast<uint4*>(y)[q] = Y;
        reinterpret_cast<uint4*>(z)[q] = Z;
    }
    const long tail = n - (n4 << 2);
    if (q < tail) {
        const long i = (n4 << 2) + q;
        x[i] += fixed32_from_float(vx[i] * lambda);
        y[i] += fixed32_from_float(vy[i] * lambda);
        z[i] += fixed32_from_float(vz[i] * lambda);
    }
}

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" int ct2_drift(float dt, float scale) {
#pragma clang fp contract(off)
    CT2_TRY(require_particles());
    if (!(scale != 0.0f)) return set_error(CT2_ERR_INVALID, "drift scale must be non-zero");
    Context& c = ctx();
    const float lambda = dt / scale;  // drift.cpp:6, computed once on the host
    const long n = (long)c.n;
    const long work = std::max<long>(n >> 2, n & 3);
    const unsigned blocks = (unsigned)((work + DRIFT_BLOCK - 1) / DRIFT_BLOCK);
    drift_kernel<<<blocks, DRIFT_BLOCK, 0, c.stream>>>(c.x, c.y, c.z, c.vx, c.vy, c.vz, n, lambda);
    CT2_LAUNCH_CHECK();
    return CT2_OK;
}
