// Colour stage: interleaved uint8 <-> level-shifted planes with RCT / ICT.
//
// Replaces the host-side level shift and interleave of the reference
// (src/OCLTest.cpp:81-90, src/OCLMemoryManager.cpp:41-54) and its post-DWT
// de-interleave kernel (src/CL/oclplanar.cl:23-39).  T.800 G.2 (RCT) and
// G.3 (ICT); FP32 operation order identical to the CPU oracle (fmaf chains).
//
// Streaming design, both directions: a thread handles 4-pixel quads strided
// by the workgroup, so lane k of a wavefront reads bytes [12k, 12k+12) (RGB)
// of one contiguous span and writes bytes [16k, 16k+16) of each plane's span
// (the inverse the other way round): every memory instruction is a fully
// coalesced, contiguous wavefront access.  No LDS.
#include <algorithm>
#include <type_traits>
#include "colour_px.h"
#include "rg_internal.h"

namespace {

constexpr int QPT = 4;               // forward: 4-pixel quads per thread
constexpr int SPAN = 256 * QPT * 4;  // forward: pixels per 256-thread workgroup

// Streaming (non-temporal) 16-byte and 4-byte stores: the planes and images
// these kernels write are read by the next kernel, long after they would
// have left the L2.
typedef float rg_f4 __attribute__((ext_vector_type(4)));
typedef int rg_i4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void st_nt4(float *p, float a, float b, float c, float d)
{
    rg_f4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<rg_f4 *>(p));
}
__device__ __forceinline__ void st_nt4(int32_t *p, int32_t a, int32_t b, int32_t c, int32_t d)
{
    rg_i4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<rg_i4 *>(p));
}

using rg_px::px_fwd;

// One quad: pixels [p, p+4) (p < npix, p a multiple of 4).  A full quad reads
// NC aligned dwords (4 pixels x NC bytes) and writes one 16-byte vector per
// plane; a quad cut by the image end goes pixel by pixel.
template <int NC, int MODE, bool FLT>
__device__ __forceinline__ void colour_fwd_quad(const uint8_t *__restrict__ src, long npix, void *__restrict__ planes,
                                                long p)
{
    using T = typename std::conditional<FLT, float, int32_t>::type;
    T *out = reinterpret_cast<T *>(planes);
    if (p + 4 <= npix) {
        const uint32_t *s32 = reinterpret_cast<const uint32_t *>(src + p * NC);
        uint32_t w[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) w[k] = s32[k];
        T o[NC][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = (int)((w[(NC * i + c) >> 2] >> (((NC * i + c) & 3) * 8)) & 0xFFu) - 128;
            T t[NC];
            px_fwd<NC, MODE, FLT>(v, t);
#pragma unroll
            for (int c = 0; c < NC; ++c) o[c][i] = t[c];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            st_nt4(out + (long)c * npix + p, o[c][0], o[c][1], o[c][2], o[c][3]);
        }
    } else {
        for (long q = p; q < npix; ++q) {
            int v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = (int)src[q * NC + c] - 128;
            T t[NC];
            px_fwd<NC, MODE, FLT>(v, t);
#pragma unroll
            for (int c = 0; c < NC; ++c) out[(long)c * npix + q] = t[c];
        }
    }
}

// A 256-thread workgroup covers SPAN pixels from p_wg: thread t takes quads
// t, t+256, ... so each load / store instruction of the wavefront touches one
// contiguous range (12 B per lane in, 16 B per lane per plane out for RGB).
template <int NC, int MODE, bool FLT>
__device__ __forceinline__ void colour_fwd_span(const uint8_t *__restrict__ src, long npix, void *__restrict__ planes,
                                                long p_wg)
{
#pragma unroll
    for (int u = 0; u < QPT; ++u) {
        const long p = p_wg + 4 * ((long)u * 256 + threadIdx.x);
        if (p < npix) colour_fwd_quad<NC, MODE, FLT>(src, npix, planes, p);
    }
}

template <int NC, int MODE, bool FLT>
__global__ void __launch_bounds__(256) colour_fwd_kernel(const uint8_t *__restrict__ src, long npix,
                                                          void *__restrict__ planes)
{
    colour_fwd_span<NC, MODE, FLT>(src, npix, planes, (long)blockIdx.x * SPAN);
}

using rg_px::clamp_u8;
using rg_px::out_u8f;
using rg_px::px_inv;

// One quad of the inverse: pixels [p, p+4).  A full quad reads one 16-byte
// vector per plane and writes NC dwords (4 pixels x NC bytes, p*NC is a
// multiple of 4); a quad cut by the image end goes pixel by pixel.
template <int NC, int MODE, bool FLT>
__device__ __forceinline__ void colour_inv_quad(const void *__restrict__ planes, long npix, uint8_t *__restrict__ dst,
                                                long p)
{
    using T = typename std::conditional<FLT, float, int32_t>::type;
    using V = typename std::conditional<FLT, float4, int4>::type;
    const T *in = reinterpret_cast<const T *>(planes);
    if (p + 4 <= npix) {
        V vv[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) vv[c] = *reinterpret_cast<const V *>(in + (long)c * npix + p);
        uint32_t w[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) w[k] = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            T v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = i == 0 ? vv[c].x : (i == 1 ? vv[c].y : (i == 2 ? vv[c].z : vv[c].w));
            uint8_t o[NC];
            px_inv<NC, MODE, FLT>(v, o);
#pragma unroll
            for (int c = 0; c < NC; ++c) w[(NC * i + c) >> 2] |= (uint32_t)o[c] << (((NC * i + c) & 3) * 8);
        }
        uint32_t *d32 = reinterpret_cast<uint32_t *>(dst + p * NC);
#pragma unroll
        for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(w[k], d32 + k);
    } else {
        for (long q = p; q < npix; ++q) {
            T v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = in[(long)c * npix + q];
            uint8_t o[NC];
            px_inv<NC, MODE, FLT>(v, o);
#pragma unroll
            for (int c = 0; c < NC; ++c) dst[q * NC + c] = o[c];
        }
    }
}

// Same span layout as the forward: quads t, t+256, ... of the workgroup's
// SPAN pixels, so each wavefront load is 1 KB contiguous per plane and each
// store NC x 256 B contiguous.
template <int NC, int MODE, bool FLT>
__global__ void __launch_bounds__(256) colour_inv_kernel(const void *__restrict__ planes, long npix,
                                                          uint8_t *__restrict__ dst)
{
    const long p_wg = (long)blockIdx.x * SPAN;
#pragma unroll
    for (int u = 0; u < QPT; ++u) {
        const long p = p_wg + 4 * ((long)u * 256 + threadIdx.x);
        if (p < npix) colour_inv_quad<NC, MODE, FLT>(planes, npix, dst, p);
    }
}

// Generic per-pixel path for other channel counts (shift only).
template <bool FLT>
__global__ void colour_fwd_generic(const uint8_t *__restrict__ src, long npix, int nc, void *__restrict__ planes)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    for (int c = 0; c < nc; ++c) {
        int v = (int)src[i * nc + c] - 128;
        if (FLT) reinterpret_cast<float *>(planes)[(long)c * npix + i] = (float)v;
        else reinterpret_cast<int32_t *>(planes)[(long)c * npix + i] = v;
    }
}

template <bool FLT>
__global__ void colour_inv_generic(const void *__restrict__ planes, long npix, int nc, uint8_t *__restrict__ dst)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    for (int c = 0; c < nc; ++c) {
        if (FLT) dst[i * nc + c] = out_u8f(reinterpret_cast<const float *>(planes)[(long)c * npix + i]);
        else dst[i * nc + c] = clamp_u8(reinterpret_cast<const int32_t *>(planes)[(long)c * npix + i] + 128);
    }
}

__global__ void planar_split_i16_kernel(const int16_t *__restrict__ src, long npix, int nc, int16_t *__restrict__ dst)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    for (int c = 0; c < nc; ++c) dst[(long)c * npix + i] = src[i * nc + c];
}

}  // namespace

int rg_launch_colour_fwd(hipStream_t s, const uint8_t *src, int w, int h, int ncomp, int colour,
                         int out_float, void *planes)
{
    long npix = (long)w * h;
    if (npix == 0) return RG_SUCCESS;
    RG_CHECK(((uintptr_t)src & 3) == 0, "colour: source must be 4-byte aligned");
    dim3 grid((unsigned)((npix + SPAN - 1) / SPAN)), block(256);
    if (ncomp == 3 && colour == RG_COLOUR_RCT)
        colour_fwd_kernel<3, 1, false><<<grid, block, 0, s>>>(src, npix, planes);
    else if (ncomp == 3 && colour == RG_COLOUR_ICT)
        colour_fwd_kernel<3, 2, true><<<grid, block, 0, s>>>(src, npix, planes);
    else if (colour != RG_COLOUR_NONE)
        return rg_fail(RG_ERR_INVALID, "colour transform %d needs 3 components (got %d)", colour, ncomp);
    else if (ncomp == 1)
        out_float ? colour_fwd_kernel<1, 0, true><<<grid, block, 0, s>>>(src, npix, planes)
                  : colour_fwd_kernel<1, 0, false><<<grid, block, 0, s>>>(src, npix, planes);
    else if (ncomp == 3)
        out_float ? colour_fwd_kernel<3, 0, true><<<grid, block, 0, s>>>(src, npix, planes)
                  : colour_fwd_kernel<3, 0, false><<<grid, block, 0, s>>>(src, npix, planes);
    else {
        dim3 g2((unsigned)((npix + 255) / 256));
        out_float ? colour_fwd_generic<true><<<g2, block, 0, s>>>(src, npix, ncomp, planes)
                  : colour_fwd_generic<false><<<g2, block, 0, s>>>(src, npix, ncomp, planes);
    }
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_colour_inv(hipStream_t s, const void *planes, int w, int h, int ncomp, int colour,
                         int in_float, uint8_t *dst)
{
    long npix = (long)w * h;
    if (npix == 0) return RG_SUCCESS;
    RG_CHECK(((uintptr_t)dst & 3) == 0 && ((uintptr_t)planes & 15) == 0,
             "colour: destination must be 4-byte and planes 16-byte aligned");
    dim3 grid((unsigned)((npix + SPAN - 1) / SPAN)), block(256);
    if (ncomp == 3 && colour == RG_COLOUR_RCT)
        colour_inv_kernel<3, 1, false><<<grid, block, 0, s>>>(planes, npix, dst);
    else if (ncomp == 3 && colour == RG_COLOUR_ICT)
        colour_inv_kernel<3, 2, true><<<grid, block, 0, s>>>(planes, npix, dst);
    else if (colour != RG_COLOUR_NONE)
        return rg_fail(RG_ERR_INVALID, "colour transform %d needs 3 components (got %d)", colour, ncomp);
    else if (ncomp == 1)
        in_float ? colour_inv_kernel<1, 0, true><<<grid, block, 0, s>>>(planes, npix, dst)
                 : colour_inv_kernel<1, 0, false><<<grid, block, 0, s>>>(planes, npix, dst);
    else if (ncomp == 3)
        in_float ? colour_inv_kernel<3, 0, true><<<grid, block, 0, s>>>(planes, npix, dst)
                 : colour_inv_kernel<3, 0, false><<<grid, block, 0, s>>>(planes, npix, dst);
    else {
        dim3 g2((unsigned)((npix + 255) / 256));
        in_float ? colour_inv_generic<true><<<g2, block, 0, s>>>(planes, npix, ncomp, dst)
                 : colour_inv_generic<false><<<g2, block, 0, s>>>(planes, npix, ncomp, dst);
    }
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_planar_split_i16(hipStream_t s, const int16_t *src, int w, int h, int ncomp, int16_t *dst)
{
    long npix = (long)w * h;
    if (npix == 0) return RG_SUCCESS;
    planar_split_i16_kernel<<<dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s>>>(src, npix, ncomp, dst);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// int32 coefficient planes -> saturated int16 planes (the per-channel int16
// planes the reference's planar stage leaves for its bit-plane coder).
namespace {
__global__ void sat_i16_kernel(const int32_t *__restrict__ src, long n, int16_t *__restrict__ dst)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = src[i];
    dst[i] = (int16_t)(v < -32768 ? -32768 : (v > 32767 ? 32767 : v));
}
}  // namespace

int rg_launch_sat_i16(hipStream_t s, const int32_t *src, long n, int16_t *dst)
{
    if (n == 0) return RG_SUCCESS;
    sat_i16_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(src, n, dst);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// ---- batched colour: one workgroup per job {image, first 16-px chunk} ----
namespace {
__global__ void __launch_bounds__(256) colour_fwd_batch_kernel(const rg_colour_job_params *__restrict__ P,
                                                               const int2 *__restrict__ jobs)
{
    const int2 j = jobs[blockIdx.x];
    const rg_colour_job_params p = P[j.x];
    const long p0 = (long)j.y * 16;   // jobs count 16-pixel chunks; a workgroup spans 256 of them
    if (p.ncomp == 3 && p.colour == RG_COLOUR_ICT) colour_fwd_span<3, 2, true>(p.src, p.npix, p.planes, p0);
    else if (p.ncomp == 3 && p.colour == RG_COLOUR_RCT) colour_fwd_span<3, 1, false>(p.src, p.npix, p.planes, p0);
    else if (p.ncomp == 3) {
        if (p.out_float) colour_fwd_span<3, 0, true>(p.src, p.npix, p.planes, p0);
        else colour_fwd_span<3, 0, false>(p.src, p.npix, p.planes, p0);
    } else {
        if (p.out_float) colour_fwd_span<1, 0, true>(p.src, p.npix, p.planes, p0);
        else colour_fwd_span<1, 0, false>(p.src, p.npix, p.planes, p0);
    }
}
}  // namespace

int rg_launch_colour_fwd_batch(hipStream_t s, const rg_colour_job_params *params, const int2 *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    colour_fwd_batch_kernel<<<dim3((unsigned)njobs), dim3(256), 0, s>>>(params, jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// ---- synthetic test images on the device ----
// Same recipe and FP64 operation order as roger_amd/synth.py: 128 + sum of 4
// sinusoids (amplitude 24) + integer noise -8..8 (+ a luminance walk for RGB),
// rounded half up and clamped.  Per channel: 12 params (period, theta, phi) x4
// already expanded to (cos(theta), sin(theta), period, phi) by the host.
namespace {
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void synth_kernel(uint8_t *__restrict__ dst, int w, int h, int ncomp, const double *__restrict__ par,
                             const uint64_t *__restrict__ noise_base, const double *__restrict__ walk_x,
                             const double *__restrict__ walk_y)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const double twopi = 2.0 * 3.141592653589793;
    for (int c = 0; c < ncomp; ++c) {
        const double *p = par + 16 * c;
        double f = 128.0;
        for (int k = 0; k < 4; ++k) {
            double t = (double)x * p[4 * k] + (double)y * p[4 * k + 1];
            double arg = twopi * t / p[4 * k + 2] + p[4 * k + 3];
            f += 24.0 * sin(arg);
        }
        uint64_t z = mix64(noise_base[c] * (1ull << 20) + (uint64_t)i);
        double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        f = f + (floor(u * 17.0) - 8.0);
        if (walk_x) f = f + (walk_x[x] + walk_y[y]);
        double r = floor(f + 0.5);
        dst[i * ncomp + c] = (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
    }
}
}  // namespace

extern "C" int rg_synth_u8(rg_context *ctx, uint8_t *d_dst, int w, int h, int ncomp, const double *params,
                           const uint64_t *noise_base, const double *walk_x, const double *walk_y)
{
    RG_CHECK(ctx && d_dst && params && noise_base, "rg_synth_u8: null argument");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "rg_synth_u8: bad shape");
    RG_CHECK((walk_x == nullptr) == (walk_y == nullptr), "rg_synth_u8: give both walks or neither");
    // small host arrays -> device (stream-ordered, freed after the launch completes)
    size_t npar = 16 * (size_t)ncomp, bytes = npar * 8 + ncomp * 8 + (walk_x ? ((size_t)w + h) * 8 : 0);
    char *tmp = nullptr;
    RG_HIP(hipMallocAsync((void **)&tmp, bytes, ctx->stream));
    RG_HIP(hipMemcpyAsync(tmp, params, npar * 8, hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(hipMemcpyAsync(tmp + npar * 8, noise_base, ncomp * 8, hipMemcpyHostToDevice, ctx->stream));
    double *wx = nullptr, *wy = nullptr;
    if (walk_x) {
        wx = (double *)(tmp + npar * 8 + ncomp * 8);
        wy = wx + w;
        RG_HIP(hipMemcpyAsync(wx, walk_x, (size_t)w * 8, hipMemcpyHostToDevice, ctx->stream));
        RG_HIP(hipMemcpyAsync(wy, walk_y, (size_t)h * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    long n = (long)w * h;
    synth_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream>>>(
        d_dst, w, h, ncomp, (const double *)tmp, (const uint64_t *)(tmp + npar * 8), wx, wy);
    RG_LAUNCH_CHECK();
    RG_HIP(hipFreeAsync(tmp, ctx->stream));
    RG_HIP(hipStreamSynchronize(ctx->stream));
    return RG_SUCCESS;
}

// ---- streaming copy: the measured HBM ceiling the bandwidth-bound kernels are
//      compared with.  One 16-byte load and one non-temporal 16-byte store per
//      thread, one pass over the buffer: on MI355X this sustains ~7.3 TB/s
//      (read + write), against ~6.0 TB/s with ordinary stores and ~5.2 TB/s
//      for a grid-stride loop of 16 workgroups per CU. ----
namespace {
__global__ void __launch_bounds__(256) stream_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 v = src[i];
    rg_i4 w = {(int)v.x, (int)v.y, (int)v.z, (int)v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<rg_i4 *>(dst + i));
}
}  // namespace

extern "C" int rg_stream_copy(rg_context *ctx, void *d_dst, const void *d_src, size_t bytes)
{
    RG_CHECK(ctx && d_dst && d_src, "rg_stream_copy: null argument");
    RG_CHECK(bytes % 16 == 0 && ((uintptr_t)d_dst & 15) == 0 && ((uintptr_t)d_src & 15) == 0,
             "rg_stream_copy: 16-byte aligned sizes only");
    const long n = (long)(bytes / 16);
    if (n == 0) return RG_SUCCESS;
    RG_CHECK((n + 255) / 256 <= 0x7fffffffL, "rg_stream_copy: %zu bytes is too large for one launch", bytes);
    stream_copy_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream>>>((const uint4 *)d_src,
                                                                                       (uint4 *)d_dst, n);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}
