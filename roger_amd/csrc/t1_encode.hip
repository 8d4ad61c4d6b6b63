// EBCOT tier-1 encode kernels and launchers over the cores in t1_core.h.
//
// Replaces the reference's bit-plane coder launch src/OCLBPC.cpp:38-51 (one
// kernel per channel over src/CL/oclbpc.cl:279-746).
//
// Kernels (all 64-thread workgroups = one wavefront):
//   t1_encode_kernel       inline path, one workgroup per descriptor
//   t1_encode_jobs_kernel  inline path, one workgroup per job-list entry
//   t1_cm_kernel           split path, context modelling per descriptor
//   t1_cm_jobs_kernel      split path, context modelling per job-list entry
//   t1_pack_kernel         parallel styles, pass segments into the slots
// The lane coders live in t1_mq.hip (one MQ chain per code-block) and
// t1_chain.hip (the sorted pass chains of the parallel styles); the three
// files compile in parallel.
#include <cstdio>
#include <cstdlib>
#include "rg_internal.h"
#include "wave_hip.h"
#include "t1_core.h"

using namespace rg_t1;

namespace {

// Plain path: one workgroup (one wavefront) per descriptor.
__global__ void __launch_bounds__(64) t1_encode_kernel(const int32_t *__restrict__ q, int pitch, size_t plane_stride,
                                                       const rg_cb_desc *__restrict__ desc, uint8_t *__restrict__ out,
                                                       int slot_bytes, rg_cb_result *__restrict__ res)
{
    const rg_cb_desc d = desc[blockIdx.x];
    T1Job job;
    job.src = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
    job.pitch = pitch; job.w = d.w; job.h = d.h; job.orient = d.orient; job.cap = slot_bytes;
    job.out = out + (size_t)blockIdx.x * slot_bytes;
    job.res = res + blockIdx.x;
    job.style = 0;
    __shared__ __align__(16) uint8_t lut[CM_LDS_BYTES];
    t1_block<false>(job, lut);
}

// Batched path: one workgroup per entry of a job list sorted largest block
// first; the dispatcher hands workgroups out in index order as compute units
// free up, which gives the longest-first schedule without a device counter.
//
// (A persistent grid pulling indices from an atomic counter is NOT used: its
// loop needs every lane converged at the broadcast of lane 0's index, and the
// compiler may legally split the lanes across iterations once the block body
// ends in a lane-0-only store -- lanes 1..63 then re-enter with index 0 and
// never leave the loop.)
__global__ void __launch_bounds__(64) t1_encode_jobs_kernel(const T1Job *__restrict__ jobs)
{
    __shared__ __align__(16) uint8_t lut[CM_LDS_BYTES];
    t1_block<false>(jobs[blockIdx.x], lut);
}

// Persistent form of the batched tier-1: a grid sized to the machine pulls job
// indices from a device counter (jobs sorted largest block first) until the
// list is exhausted.  One atomic per job: lane 0 alone takes the next index
// and leaves it in workgroup memory; after the barrier every lane reads the
// same word, so all lanes see the same index and leave the loop in the same
// iteration (no lane ever depends on a broadcast that assumes convergence).
// The second barrier keeps lane 0's next write behind every lane's read.
template <bool SPLIT>
__device__ __forceinline__ void t1_queue_loop(const T1Job *__restrict__ jobs, int njobs, int *__restrict__ counter)
{
    __shared__ __align__(16) uint8_t lut[CM_LDS_BYTES];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int i = next;
        __syncthreads();
        if (i >= njobs) break;
        t1_block<SPLIT>(jobs[i], lut);
    }
}

__global__ void __launch_bounds__(64) t1_encode_queue_kernel(const T1Job *__restrict__ jobs, int njobs,
                                                             int *__restrict__ counter)
{
    t1_queue_loop<false>(jobs, njobs, counter);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) t1_cm_queue_kernel(const T1Job *__restrict__ jobs, int njobs,
                                                         int *__restrict__ counter)
{
    t1_queue_loop<true>(jobs, njobs, counter);
}

// Queue counter reset by a kernel (not a memset node: graph replays stay exact).
__global__ void reset_counter_kernel(int *counter) { *counter = 0; }

// ------------------------------------------------------------------------
// Split path.  Context modelling runs one wavefront per code-block and writes
// the block's (context, bit) stream; the MQ coder then runs one LANE per
// code-block, so 64 blocks advance their serial coder chains together in
// vector registers instead of one block per wavefront in scalar registers.
// ------------------------------------------------------------------------
// Four waves per SIMD: the register budget drops from 146 to 128 VGPRs at
// the price of 60 bytes per lane of spills, and the context modelling (issue
// bound, one wave per block) gains more from the extra wave than the spills
// cost: the 4096x4096 RGB encode went from 2.83 to 2.66 ms.  The batched
// forms below take the same budget (256 mixed images: 43.8 -> 39.0 ms).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) t1_cm_kernel(const int32_t *__restrict__ q, int pitch, size_t plane_stride,
                                                   const rg_cb_desc *__restrict__ desc, uint8_t *__restrict__ syms,
                                                   int symcap, int4 *__restrict__ cm, int32_t *__restrict__ poffs,
                                                   int style, int qbeg, int qend, uint4 *__restrict__ cms,
                                                   uint64_t *__restrict__ sppm)
{
    const rg_cb_desc d = desc[blockIdx.x];
    T1Job job;
    job.src = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
    job.pitch = pitch; job.w = d.w; job.h = d.h; job.orient = d.orient;
    job.sym = syms + (size_t)blockIdx.x * symcap;
    job.symcap = symcap;
    job.cm = cm + blockIdx.x;
    job.poff = poffs ? poffs + (size_t)blockIdx.x * MAXP : nullptr;
    job.style = style;
    job.cms = cms ? cms + (size_t)blockIdx.x * 64 : nullptr;
    job.sppm = sppm ? sppm + (size_t)blockIdx.x * RD_PLANES * 64 : nullptr;
    __shared__ __align__(16) uint8_t lut[CM_LDS_BYTES];
    CMPhase ph;
    ph.qbeg = qbeg;
    ph.qend = qend;
    t1_block<true>(job, lut, ph);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) t1_cm_jobs_kernel(const T1Job *__restrict__ jobs)
{
    __shared__ __align__(16) uint8_t lut[CM_LDS_BYTES];
    t1_block<true>(jobs[blockIdx.x], lut);
}

__global__ void __launch_bounds__(64) t1_pack_kernel(const int4 *__restrict__ cm, const int32_t *__restrict__ poffs,
                                                     const uint8_t *__restrict__ arena, size_t arena_stride,
                                                     const int32_t *__restrict__ seglens, uint8_t *__restrict__ slots,
                                                     int slot, rg_cb_result *__restrict__ res)
{
    __shared__ int32_t sh[PACK_WORDS];
    const size_t b = blockIdx.x;
    t1_pack_block(cm[b], poffs + b * MAXP, arena + b * arena_stride, seglens + b * MAXP, slots + b * slot, slot,
                  res + b, sh);
}

__global__ void __launch_bounds__(64) t1_pack_jobs_kernel(const T1Job *__restrict__ jobs)
{
    __shared__ int32_t sh[PACK_WORDS];
    const T1Job &j = jobs[blockIdx.x];
    t1_pack_block(*j.cm, j.poff, j.seg, j.seglen, j.out, j.cap, j.res, sh);
}

}  // namespace

int rg_t1_maxp() { return MAXP; }
size_t rg_t1_seg_bytes(int symcap) { return (size_t)seg_bytes(symcap); }

int rg_launch_t1_parallel(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc,
                          int ncb, uint8_t *syms, int symcap, int4 *cm, int32_t *poffs, uint8_t *arena,
                          int32_t *seglens, uint8_t *out, int slot_bytes, rg_cb_result *res, int style,
                          int32_t *chain_ws, const rg_t1_overlap *ov, uint64_t *sppm)
{
    if (ncb <= 0) return RG_SUCCESS;
    if (!chain_ws) return rg_fail(RG_ERR_INVALID, "parallel code-block styles need the chain workspace");
    const size_t astride = rg_t1_seg_bytes(symcap);
    const dim3 g((unsigned)ncb), b(64);
    const rg_t1_flat fj{syms, symcap, cm, poffs, arena, astride, seglens};
    if (ov) {
        // Phases overlapped: the first passes (the first four bit-planes, where
        // every code-block's longest MQ chains are) are modelled and listed
        // first, in one or two phases, and each early phase's chains are coded
        // on its own stream while the first stream models the next passes; the
        // last phase codes its (raw or short) chains on the first stream and
        // packing joins them all.  The lists are built on the first stream
        // before the next modelling launch: built beside it, their 256-thread
        // workgroups waited for SIMD slots until its whole grid was dispatched.
        int sp1, sp2;
        rg_t1_splits(&sp1, &sp2);
        int bounds[RG_T1_MAX_EARLY + 1] = {0, sp2, sp2, sp2};
        if (ov->nearly == 2) bounds[1] = sp1;
        if (ov->nearly == 3) {
            // [0, 4), [4, sp1), [sp1, sp2); with sp1 <= 4 the third early phase is empty
            bounds[1] = sp1 > RG_T1_PHASE_SPLIT0 ? RG_T1_PHASE_SPLIT0 : sp1;
            bounds[2] = sp1 > RG_T1_PHASE_SPLIT0 ? sp1 : sp2;
        }
        for (int k = 0; k < ov->nearly; ++k) {
            const int qlo = bounds[k], qhi = bounds[k + 1];
            t1_cm_kernel<<<g, b, 0, s>>>(q, pitch, plane_stride, desc, syms, symcap, cm, poffs, style, qlo, qhi,
                                         (uint4 *)ov->cms, sppm);
            RG_LAUNCH_CHECK();
            if (int rc = rg_launch_chain_lists_flat(s, fj, ncb, style, ov->ws[k], qlo, qhi)) return rc;
            RG_HIP(hipEventRecord(ov->ev_in[k], s));
            RG_HIP(hipStreamWaitEvent(ov->se[k], ov->ev_in[k], 0));
            if (int rc = rg_launch_chain_code_flat(ov->se[k], fj, ncb, ov->ws[k])) return rc;
            RG_HIP(hipEventRecord(ov->ev_out[k], ov->se[k]));
        }
        t1_cm_kernel<<<g, b, 0, s>>>(q, pitch, plane_stride, desc, syms, symcap, cm, poffs, style,
                                     sp2, 1 << 30, (uint4 *)ov->cms, sppm);
        RG_LAUNCH_CHECK();
        if (int rc = rg_launch_chains_flat(s, fj, ncb, style, chain_ws, sp2, MAXP)) return rc;
        for (int k = 0; k < ov->nearly; ++k) RG_HIP(hipStreamWaitEvent(s, ov->ev_out[k], 0));
        t1_pack_kernel<<<g, b, 0, s>>>(cm, poffs, arena, astride, seglens, out, slot_bytes, res);
        RG_LAUNCH_CHECK();
        return RG_SUCCESS;
    }
    t1_cm_kernel<<<g, b, 0, s>>>(q, pitch, plane_stride, desc, syms, symcap, cm, poffs, style, 0, 1 << 30,
                                 nullptr, sppm);
    RG_LAUNCH_CHECK();
    if (int rc = rg_launch_chains_flat(s, fj, ncb, style, chain_ws, 0, MAXP)) return rc;
    t1_pack_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(cm, poffs, arena, astride, seglens, out, slot_bytes, res);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_t1_jobs_parallel(hipStream_t s, const void *jobs, int njobs, int style, int32_t *chain_ws)
{
    if (njobs <= 0) return RG_SUCCESS;
    if (!chain_ws) return rg_fail(RG_ERR_INVALID, "parallel code-block styles need the chain workspace");
    t1_cm_jobs_kernel<<<dim3((unsigned)njobs), dim3(64), 0, s>>>((const T1Job *)jobs);
    RG_LAUNCH_CHECK();
    if (int rc = rg_launch_chains_jobs(s, jobs, njobs, style, chain_ws, 0, MAXP)) return rc;
    t1_pack_jobs_kernel<<<dim3((unsigned)njobs), dim3(64), 0, s>>>((const T1Job *)jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// Symbol-stream capacity per code-block sample for the split path.
int rg_t1_symcap(int cbw, int cbh) { return ((cbw * cbh * 24) + 15) / 16 * 16; }

int rg_launch_t1_split(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc, int ncb,
                       uint8_t *syms, int symcap, int4 *cm, uint8_t *out, int slot_bytes, rg_cb_result *res)
{
    if (ncb <= 0) return RG_SUCCESS;
    t1_cm_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(q, pitch, plane_stride, desc, syms, symcap, cm, nullptr,
                                                          0, 0, 1 << 30, nullptr, nullptr);
    RG_LAUNCH_CHECK();
    return rg_launch_t1_mq(s, syms, symcap, cm, ncb, out, slot_bytes, res);
}

int rg_launch_t1_jobs_split(hipStream_t s, const void *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    t1_cm_jobs_kernel<<<dim3((unsigned)njobs), dim3(64), 0, s>>>((const T1Job *)jobs);
    RG_LAUNCH_CHECK();
    return rg_launch_t1_mq_jobs(s, jobs, njobs);
}

int rg_t1_job_bytes() { return (int)sizeof(T1Job); }

// Persistent forms: reset the queue counter, then `nwaves` workgroups pull
// code-blocks until the job list is exhausted.
int rg_launch_t1_queue(hipStream_t s, const void *jobs, int njobs, int style, int *counter, int nwaves,
                       int32_t *chain_ws)
{
    if (njobs <= 0) return RG_SUCCESS;
    if ((style & RG_STYLE_TERMALL) && !chain_ws)
        return rg_fail(RG_ERR_INVALID, "parallel code-block styles need the chain workspace");
    reset_counter_kernel<<<1, 1, 0, s>>>(counter);
    RG_LAUNCH_CHECK();
    const dim3 g((unsigned)nwaves), b(64);
    const T1Job *J = (const T1Job *)jobs;
    if (style & RG_STYLE_TERMALL) {
        t1_cm_queue_kernel<<<g, b, 0, s>>>(J, njobs, counter);
        RG_LAUNCH_CHECK();
        if (int rc = rg_launch_chains_jobs(s, jobs, njobs, style, chain_ws, 0, MAXP)) return rc;
        t1_pack_jobs_kernel<<<dim3((unsigned)njobs), b, 0, s>>>(J);
    } else if (rg_t1_inline()) {
        t1_encode_queue_kernel<<<g, b, 0, s>>>(J, njobs, counter);
    } else {
        t1_cm_queue_kernel<<<g, b, 0, s>>>(J, njobs, counter);
        RG_LAUNCH_CHECK();
        return rg_launch_t1_mq_jobs(s, jobs, njobs);
    }
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_t1_jobs(hipStream_t s, const void *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    t1_encode_jobs_kernel<<<dim3((unsigned)njobs), dim3(64), 0, s>>>((const T1Job *)jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// Host-side fill of one job record (the struct layout stays private to the tier-1 sources).
void rg_t1_job_fill(void *dst, const int32_t *src, int pitch, int w, int h, int orient, uint8_t *out, int cap,
                    rg_cb_result *res, uint8_t *sym, int symcap, int4 *cm, int32_t *poff, uint8_t *seg,
                    int32_t *seglen, int style)
{
    T1Job j;
    j.src = src; j.pitch = pitch; j.w = w; j.h = h; j.orient = orient; j.cap = cap; j.out = out; j.res = res;
    j.sym = sym; j.symcap = symcap; j.style = style; j.cm = cm; j.poff = poff; j.seg = seg; j.seglen = seglen;
    *reinterpret_cast<T1Job *>(dst) = j;
}

// Phases of the overlapped tier-1 of the parallel styles.  Measured on the
// 4096x4096 RGB encode with the two-wavefront chain coder: RESET|TERMALL|
// BYPASS 2.71 ms in one phase, 2.49 in two (split before pass 13; 2.52 at 10,
// 2.51 at 16), 2.67 in three (7, 10); RESET|TERMALL 3.02 / 3.15 / 3.36 ms --
// with every pass an MQ chain, the early chains' coding does not hide behind
// the rest of the modelling.  With the faster coders since then (best of 4 x
// 10 encodes, three runs each): two phases split before pass 10 1.83 ms, 13
// 1.79-1.80, 16 1.78, 19 1.81, 22 1.82-1.83; one phase 1.82; three 2.02-2.07.
// RG_T1_PHASES=1..4 overrides (RG_T1_TWO_PHASE=0: one).
int rg_t1_phases(int style)
{
    static int v = -1;
    if (v < 0) {
        const char *two = getenv("RG_T1_TWO_PHASE"), *n = getenv("RG_T1_PHASES");
        v = 0;
        if (n && n[0] >= '1' && n[0] <= '4') v = n[0] - '0';
        if (two && two[0] == '0') v = 1;
    }
    if (v) return v;
    return (style & STYLE_BYPASS) ? 2 : 1;
}

// Phase boundaries (coding-pass indices, each 1 mod 3 so that a phase ends
// before a significance-propagation pass): RG_T1_PHASE_SPLIT1 and
// RG_T1_PHASE_SPLIT unless RG_T1_SPLITS="a,b" names others (4 <= a < b <= 94).
void rg_t1_splits(int *s1, int *s2)
{
    static int a = -1, b = -1;
    if (a < 0) {
        a = RG_T1_PHASE_SPLIT1;
        b = RG_T1_PHASE_SPLIT;
        if (const char *e = getenv("RG_T1_SPLITS")) {
            int x = 0, y = 0;
            if (sscanf(e, "%d,%d", &x, &y) == 2 && x >= 4 && x < y && y <= 94 && x % 3 == 1 && y % 3 == 1) {
                a = x;
                b = y;
            }
        }
    }
    *s1 = a;
    *s2 = b;
}

// Tier-1 variant: the split context-model + lane-MQ path unless RG_T1_INLINE=1
// selects the single-kernel path (one wavefront per block, scalar MQ coder).
bool rg_t1_inline()
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("RG_T1_INLINE");
        v = (e && e[0] == '1') ? 1 : 0;
    }
    return v == 1;
}

int rg_launch_t1_encode(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride,
                        const rg_cb_desc *desc, int ncb, uint8_t *out, int slot_bytes, rg_cb_result *res)
{
    if (ncb <= 0) return RG_SUCCESS;
    t1_encode_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(q, pitch, plane_stride, desc, out, slot_bytes, res);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}
