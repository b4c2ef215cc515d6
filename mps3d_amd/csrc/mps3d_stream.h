// Streaming kernels that measure what HBM delivers to a hand-written kernel
// on this device, for the access patterns of the solver's kernels: read
// only, write only, copy, triad, and the red-black sweep's pattern (three
// streams read, one written at the same index). Every access is 16 bytes per
// lane (global_load_dwordx4 / global_store_dwordx4), each lane handles
// kStreamUnroll such vectors spaced one grid-row apart, so a wave keeps that
// many independent loads in flight. No reference counterpart: the figures
// are the ceiling against which the stencil kernels' bandwidth is read.
#pragma once

#include "mps3d_kernels.h"

namespace mps {

constexpr int kStreamBlock = 256;
constexpr int kStreamUnroll = 4;

// KIND 0: partial sums of a (read)    1: o = s (write)    2: o = a (copy)
//      3: o = a + s b (triad)         4: o = a + s (b + c) (sweep pattern)
template <int KIND>
__global__ void __launch_bounds__(kStreamBlock) k_stream(const float4 *__restrict__ a, const float4 *__restrict__ b,
                                                         const float4 *__restrict__ c, float4 *__restrict__ o,
                                                         long long n4, float s, double *partials)
{
    const long long base = (long long)blockIdx.x * kStreamBlock * kStreamUnroll + threadIdx.x;
    float4 va[kStreamUnroll], vb[kStreamUnroll], vc[kStreamUnroll];
#pragma unroll
    for (int q = 0; q < kStreamUnroll; ++q) {
        const long long t = base + (long long)q * kStreamBlock;
        const bool ok = t < n4;
        if (KIND != 1) va[q] = ok ? a[t] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (KIND >= 3) vb[q] = ok ? b[t] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (KIND == 4) vc[q] = ok ? c[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (KIND == 0) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kStreamUnroll; ++q) acc += (double)((va[q].x + va[q].y) + (va[q].z + va[q].w));
        const double t = block_sum(acc);
        if (threadIdx.x == 0) partials[blockIdx.x] = t;
        return;
    }
#pragma unroll
    for (int q = 0; q < kStreamUnroll; ++q) {
        const long long t = base + (long long)q * kStreamBlock;
        if (t >= n4) continue;
        float4 r;
        if (KIND == 1) {
            r = make_float4(s, s, s, s);
        } else if (KIND == 2) {
            r = va[q];
        } else if (KIND == 3) {
            r = make_float4(va[q].x + s * vb[q].x, va[q].y + s * vb[q].y, va[q].z + s * vb[q].z,
                            va[q].w + s * vb[q].w);
        } else {
            r = make_float4(va[q].x + s * (vb[q].x + vc[q].x), va[q].y + s * (vb[q].y + vc[q].y),
                            va[q].z + s * (vb[q].z + vc[q].z), va[q].w + s * (vb[q].w + vc[q].w));
        }
        o[t] = r;
    }
}

}  // namespace mps
