// image.hip -- image egress on the device: the histogram of
// Tracer.histogram (fresnel/tracer.py:79-108) over the resident linear output.
#pragma once

#include "rt/scene_dev.h"

namespace fh
{
constexpr uint32_t HIST_BINS = 512;

// Bin of v**(1/2.2) in 512 equal bins over [0, 1], or -1 outside [0, 1] (NaN
// included).  With float32 input numpy.histogram(x, bins=512, range=[0, 1])
// puts x in bin floor(512 x) and x = 1 in the last bin: 512 x is exact, and
// the bin edges k / 512 are exact float32 values.  The exponent is the
// float32 value of 1 / 2.2 (numpy raises a float32 array to a Python float in
// float32).
__device__ inline int gamma_bin(float v)
    {
    const float p = powf(v, 0.454545468f);
    if (!(p >= 0.0f && p <= 1.0f))
        return -1;
    const int b = (int)(p * 512.0f);
    return b < (int)HIST_BINS - 1 ? b : (int)HIST_BINS - 1;
    }

// Histograms of the pixels with alpha > 0: lightness 0.21 R + 0.72 G + 0.07 B
// (float32, left to right, no contraction), R, G, B, each to the power 1/2.2.
// counts[bin * 4 + column] (numpy's (512, 4) layout) must be zero on entry.
// Each block bins a grid-stride share of the pixels in LDS and adds its
// non-zero bins to the global counts.
__global__ __launch_bounds__(256) void k_histogram(const float4* __restrict__ linear, uint32_t n,
                                                   uint32_t* __restrict__ counts)
    {
    __shared__ uint32_t h[4 * HIST_BINS];
    for (uint32_t k = threadIdx.x; k < 4 * HIST_BINS; k += blockDim.x)
        h[k] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        {
        const float4 c = linear[i];
        if (!(c.w > 0.0f))
            continue;
        const float l = 0.21f * c.x + 0.72f * c.y + 0.07f * c.z;
        const float v[4] = {l, c.x, c.y, c.z};
#pragma unroll
        for (int k = 0; k < 4; k++)
            {
            const int b = gamma_bin(v[k]);
            if (b >= 0)
                atomicAdd(&h[b * 4 + k], 1u);
            }
        }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 4 * HIST_BINS; k += blockDim.x)
        if (h[k])
            atomicAdd(&counts[k], h[k]);
    }
    } // namespace fh
