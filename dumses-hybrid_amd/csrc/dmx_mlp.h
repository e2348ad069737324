// dmx_mlp.h -- the learned slope limiter: network layout and the per-cell
// arithmetic shared by the inference kernel of the step and the training
// kernels.
//
// Network: 40 -> 64 (GELU) -> 64 (GELU) -> 8.  Inputs are the scale-free
// features of the 5-cell stencil of the 8 primitive variables along one
// direction; outputs are the 8 slopes in units of the per-variable scale.
// For variable v with centre q0, differences d_m = q_m - q0 (m = -2,-1,+1,+2)
// and scale s = max |d_m|:
//     x[5v + 0..4] = d_-2/s, d_-1/s, q0/(|q0| + s), d_+1/s, d_+2/s
// (zero when s = 0), slope_v = y_v * s.  GELU is the tanh approximation of
// Hendrycks and Gimpel (2016).  Parameters are one float vector:
// W1[40][64], b1[64], W2[64][64], b2[64], W3[64][8], b3[8] (row = input).
#pragma once

#include <hip/hip_runtime.h>

namespace dmx {
namespace mlp {

constexpr int NIN = 40, NH = 64, NOUT = 8;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NOUT, NPARAM = OFF_B3 + NOUT;
static_assert(NPARAM == 7304, "parameter count");

__device__ __forceinline__ float gelu(float x)
{
    const float c = 0.7978845608028654f;   // sqrt(2/pi)
    float t = tanhf(c * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.0f + t);
}

__device__ __forceinline__ float gelu_grad(float x)
{
    const float c = 0.7978845608028654f;
    float t = tanhf(c * (x + 0.044715f * x * x * x));
    float du = c * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

// the 5 features of one variable and its scale
__device__ __forceinline__ float features5(const float q[5], float f[5])
{
    const float q0 = q[2];
    float d0 = q[0] - q0, d1 = q[1] - q0, d3 = q[3] - q0, d4 = q[4] - q0;
    float s = fmaxf(fmaxf(fabsf(d0), fabsf(d1)), fmaxf(fabsf(d3), fabsf(d4)));
    if (s > 0.0f) {
        float r = 1.0f / s;
        f[0] = d0 * r; f[1] = d1 * r; f[2] = q0 / (fabsf(q0) + s); f[3] = d3 * r; f[4] = d4 * r;
    } else {
        f[0] = f[1] = f[2] = f[3] = f[4] = 0.0f;
    }
    return s;
}

__device__ __forceinline__ float van_leer(float dl, float dr)
{
    float p = dl * dr;
    return p > 0.0f ? 2.0f * p / (dl + dr) : 0.0f;
}

// the MC slope (scheme.f90:452-477 with slope_type 2) from the one-sided
// differences: sign(dl + dr) min(2|dl|, 2|dr|, |dl + dr| / 2), 0 at an extremum
__device__ __forceinline__ float mc_slope(float dl, float dr)
{
    const float c = 0.5f * (dl + dr);
    const float m = fminf(fminf(2.0f * fabsf(dl), 2.0f * fabsf(dr)), fabsf(c));
    return dl * dr > 0.0f ? copysignf(m, c) : 0.0f;
}

// ---------------------------------------------------------------------------
// bf16 MFMA (gfx950 v_mfma_f32_16x16x32_bf16): D[16x16] += A[16x32] B[32x16].
// Lane l holds A row (l % 16), k = 8 (l / 16) + 0..7; B column (l % 16),
// k = 8 (l / 16) + 0..7; D column (l % 16), rows 4 (l / 16) + 0..3.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16x16x32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

} // namespace mlp
} // namespace dmx
