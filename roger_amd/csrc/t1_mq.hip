// Standard code-block style, second half of the split tier-1: one MQ coder
// per lane (t1_mq_lanes in t1_core.h), 64 code-blocks per two-wavefront
// workgroup (interval side in wavefront 0, code side in wavefront 1, the same
// block in lane l of both), from the symbol streams the context-modelling
// kernels wrote.  Replaces, with
// t1_encode.hip, the reference's bit-plane coder launch src/OCLBPC.cpp:38-51.
#include "rg_internal.h"
#include "wave_hip.h"
#include "t1_core.h"

using namespace rg_t1;

namespace {

__global__ void __launch_bounds__(128) t1_mq_kernel(const uint8_t *__restrict__ syms, int symcap,
                                                    const int4 *__restrict__ cm, int ncb, uint8_t *__restrict__ slots,
                                                    int slot, rg_cb_result *__restrict__ res)
{
    __shared__ uint32_t tab[MQ_TT_WORDS];
    __shared__ __attribute__((aligned(16))) uint8_t ring[MQ2_LDS_BYTES];
    t1_mq_lanes<false, true>(tab, ring, nullptr, blockIdx.x * 64 + wv_lane(), ncb, syms, symcap, cm, slots, slot,
                             res);
}

__global__ void __launch_bounds__(128) t1_mq_jobs_kernel(const T1Job *__restrict__ jobs, int njobs)
{
    __shared__ uint32_t tab[MQ_TT_WORDS];
    __shared__ __attribute__((aligned(16))) uint8_t ring[MQ2_LDS_BYTES];
    t1_mq_lanes<true, true>(tab, ring, jobs, blockIdx.x * 64 + wv_lane(), njobs, nullptr, 0, nullptr, nullptr, 0,
                            nullptr);
}

}  // namespace

int rg_launch_t1_mq(hipStream_t s, const uint8_t *syms, int symcap, const int4 *cm, int ncb, uint8_t *out,
                    int slot_bytes, rg_cb_result *res)
{
    if (ncb <= 0) return RG_SUCCESS;
    t1_mq_kernel<<<dim3((unsigned)((ncb + 63) / 64)), dim3(128), 0, s>>>(syms, symcap, cm, ncb, out, slot_bytes, res);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_t1_mq_jobs(hipStream_t s, const void *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    t1_mq_jobs_kernel<<<dim3((unsigned)((njobs + 63) / 64)), dim3(128), 0, s>>>((const T1Job *)jobs, njobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}
