// EBCOT tier-1 decoder kernel: one 64-thread workgroup (one wavefront) per
// code-block running t1_dec_block (t1_dec_core.h).  The reference has no
// decoder at all (its OCLDecoder only runs the DWT stage,
// src/OCLDecoder.cpp:35-38, and even that is never called).
#include "rg_internal.h"
#include "wave_hip.h"
#include "t1_dec_core.h"

using namespace rg_t1;

namespace {

// Six waves per SIMD (80 VGPRs, 136 bytes per lane spilled, against 121
// VGPRs and four waves): the scalar-bound decode interleaves more blocks per
// SIMD; 4096x4096 RGB tier-1 decode 32.1 -> 31.2 ms (bypass style).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8))) t1_decode_kernel(const uint8_t *__restrict__ data, const long *__restrict__ offs,
                                                       const rg_cb_result *__restrict__ res,
                                                       const rg_cb_desc *__restrict__ desc, int32_t *__restrict__ q,
                                                       int pitch, size_t plane_stride, int style,
                                                       const int32_t *__restrict__ seglens, int maxp)
{
    __shared__ uint8_t lut[LUT_BYTES];
    __shared__ uint32_t rawbuf[DEC_RAW_WORDS];
    const rg_cb_desc d = desc[blockIdx.x];
    const rg_cb_result r = res[blockIdx.x];
    DecJob job;
    job.data = data + offs[blockIdx.x];
    job.len = r.len; job.npasses = r.npasses; job.nbp = r.nbp;
    job.w = d.w; job.h = d.h; job.orient = d.orient; job.pitch = pitch;
    job.dst = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
    job.style = style;
    job.seglen = seglens ? seglens + (size_t)blockIdx.x * maxp : nullptr;
    t1_dec_block(job, lut, rawbuf);
}

}  // namespace

int rg_launch_t1_decode(hipStream_t s, const uint8_t *data, const long *offs, const rg_cb_result *res,
                        const rg_cb_desc *desc, int ncb, int32_t *q, int pitch, size_t plane_stride, int style,
                        const int32_t *seglens)
{
    if (ncb <= 0) return RG_SUCCESS;
    t1_decode_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(data, offs, res, desc, q, pitch, plane_stride, style,
                                                              seglens, RG_MAX_PASSES);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}
