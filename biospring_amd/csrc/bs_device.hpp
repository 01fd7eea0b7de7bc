// bs_device.hpp -- device helpers shared by the engine's kernels.
#pragma once

#include "bs_engine.hpp"

namespace bs
{

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }

// The status words a call's kernels share: [0] rebuild due, [1] smallest
// particle id with a non-finite position (INT_MAX: none), [2] a neighbour list
// overflowed, [3] longest list seen, [5] a non-finite position was seen before
// this step began (k_prestep's snapshot of [1]), [6] steps completed in this
// call.  Once a list overflowed or a
// step produced a non-finite position, every later kernel of the call leaves
// the state as it is: the call returns the state of that step.
__device__ inline bool halted(const int32_t *status) { return status[2] != 0 || status[5] != 0; }

__device__ inline bool gated_off(const int32_t *gate) { return gate != nullptr && (*gate == 0 || halted(gate)); }

// Makes one wave's LDS writes visible to the other lanes of the same wave.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum, valid in thread 0.  Fixed order: deterministic.
__device__ inline double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        sh[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            r += sh[w];
    __syncthreads();
    return r;
}

// The reference's two-stage cutoff test, exactly (nsearch.hpp:230 then
// Particle.cpp:400-402 through Vector3f::norm, measure.hpp:163-196): an FP64
// distance of exact float differences, then an FP64 norm of float-rounded
// differences rounded to float; both must be below the float cutoff.
// Products and sums are kept unfused, as in the reference's x86-64 build.
__device__ __noinline__ bool exact_two_stage(float4 pi, float4 pj, float rc)
{
#pragma clang fp contract(off)
    const double ax = (double)pi.x - (double)pj.x;
    const double ay = (double)pi.y - (double)pj.y;
    const double az = (double)pi.z - (double)pj.z;
    const double s1 = ax * ax + ay * ay + az * az;
    if (!(sqrt(s1) < (double)rc))
        return false;
    const double bx = (double)(pj.x - pi.x);
    const double by = (double)(pj.y - pi.y);
    const double bz = (double)(pj.z - pi.z);
    float d = 0.0f;
    if (!(fabs(bx) < 1e-40 && fabs(by) < 1e-40 && fabs(bz) < 1e-40))
        d = (float)sqrt(bx * bx + by * by + bz * bz);
    return d < rc;
}

// Fast path of the test: float r^2 outside a relative guard band of 4e-6
// decides; inside it, the exact test runs.  The band is wider than the
// float error of r^2 (about 3e-7 relative), so the decision is exact.
__device__ inline bool in_cutoff(float4 pi, float4 pj, float rc, float rc2lo, float rc2hi, float r2)
{
    if (r2 < rc2lo)
        return true;
    if (r2 >= rc2hi)
        return false;
    return exact_two_stage(pi, pj, rc);
}

// 1/sqrt(r2) to about 1e-14 relative: the FP32 hardware seed and one FP64
// Newton step (the FP64 square root and divide are multi-instruction
// sequences on CDNA).  For r2 > 0; below the smallest normal float the
// seed is clamped (finite, inexact).
__device__ inline double rsqrt_d(double r2)
{
    const double y = (double)__builtin_amdgcn_rsqf((float)fmax(r2, 1.1754944e-38));
    return y * fma(-0.5 * r2, y * y, 1.5);
}

// Cell coordinate on the engine's binning grid: edge >= min_edge per axis and
// the whole box covered by `dim` cells.
__device__ inline int32_t internal_coord(float x, double mn, double len, int32_t dim, float min_edge)
{
    const double e = fmax((double)min_edge, len / (double)dim);
    double q = ((double)x - mn) / e;
    q = q > 0.0 ? q : 0.0; // also maps NaN to 0
    int32_t c = q < (double)dim ? (int32_t)q : dim - 1;
    return c < dim - 1 ? c : dim - 1;
}

} // namespace bs
