// Per-pixel forward colour transforms (T.800 G.2 RCT, G.3 ICT) after the DC
// level shift, shared by the colour kernels and the fused colour + first DWT
// level kernel so both compute identical values.  FP32 operation order is
// the CPU oracle's (fmaf chains).
#pragma once
#include <hip/hip_runtime.h>

namespace rg_px {

__device__ __forceinline__ float ict_y(float r, float g, float b)
{
    return __builtin_fmaf((float)0.114, b, __builtin_fmaf((float)0.587, g, (float)0.299 * r));
}
__device__ __forceinline__ float ict_cb(float r, float g, float b)
{
    return __builtin_fmaf((float)0.5, b, __builtin_fmaf((float)-0.33126, g, (float)-0.16875 * r));
}
__device__ __forceinline__ float ict_cr(float r, float g, float b)
{
    return __builtin_fmaf((float)-0.08131, b, __builtin_fmaf((float)-0.41869, g, (float)0.5 * r));
}

// MODE: 0 shift only, 1 RCT, 2 ICT.  Forward transform of one level-shifted
// pixel v[0..NC) into o[0..NC) (T: float for 9/7 planes, int32 for 5/3).
template <int NC, int MODE, bool FLT, class T>
__device__ __forceinline__ void px_fwd(const int (&v)[NC], T (&o)[NC])
{
    if constexpr (MODE == 2) {
        const float r = (float)v[0], g = (float)v[1], b = (float)v[2];
        o[0] = ict_y(r, g, b);
        o[1] = ict_cb(r, g, b);
        o[2] = ict_cr(r, g, b);
    } else if constexpr (MODE == 1) {
        o[0] = (T)((v[0] + 2 * v[1] + v[2]) >> 2);
        o[1] = (T)(v[2] - v[1]);
        o[2] = (T)(v[0] - v[1]);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) o[c] = (T)v[c];
    }
}

// The FP32 modes from the level-shifted samples already as floats (exact:
// integers below 2^24), the same arithmetic as px_fwd.
template <int NC, int MODE>
__device__ __forceinline__ void px_fwd_f(const float (&v)[NC], float (&o)[NC])
{
    static_assert(MODE == 0 || MODE == 2, "FP32 modes");
    if constexpr (MODE == 2) {
        o[0] = ict_y(v[0], v[1], v[2]);
        o[1] = ict_cb(v[0], v[1], v[2]);
        o[2] = ict_cr(v[0], v[1], v[2]);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) o[c] = v[c];
    }
}

__device__ __forceinline__ uint8_t clamp_u8(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }
// A reconstructed FP32 sample (level shifted) to uint8: + 128, rounded half
// up, clamped.
__device__ __forceinline__ uint8_t out_u8f(float x)
{
    float s = x + 128.0f;
    return clamp_u8((int)floorf(s + 0.5f));
}

// Inverse transform of one pixel: the planes' values v[0..NC) -> NC bytes
// (T.800 G.2 / G.3 inverse, then the inverse level shift).
template <int NC, int MODE, bool FLT, class T>
__device__ __forceinline__ void px_inv(const T (&v)[NC], uint8_t (&o)[NC])
{
    if constexpr (FLT && MODE == 2) {
        const float y = v[0], cb = v[1], cr = v[2];
        o[0] = out_u8f(__builtin_fmaf((float)1.402, cr, y));
        o[1] = out_u8f(__builtin_fmaf((float)-0.71414, cr, __builtin_fmaf((float)-0.34413, cb, y)));
        o[2] = out_u8f(__builtin_fmaf((float)1.772, cb, y));
    } else if constexpr (!FLT && MODE == 1) {
        const int g = v[0] - ((v[1] + v[2]) >> 2);
        o[0] = clamp_u8(v[2] + g + 128);
        o[1] = clamp_u8(g + 128);
        o[2] = clamp_u8(v[1] + g + 128);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if constexpr (FLT) o[c] = out_u8f(v[c]);
            else o[c] = clamp_u8(v[c] + 128);
        }
    }
}

}  // namespace rg_px
