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
//   t1_mq_kernel           split path, MQ coder, one lane per code-block
//   t1_mq_jobs_kernel      same over a job list
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
    __shared__ uint8_t lut[LUT_BYTES];
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
    __shared__ uint8_t lut[LUT_BYTES];
    t1_block<false>(jobs[blockIdx.x], lut);
}

// Persistent form of the batched tier-1: a grid sized to the machine pulls job
// indices from a device counter (jobs sorted largest block first) until the
// list is exhausted.  Every lane executes the atomic (lane 0 adds 1, the
// others 0) and takes lane 0's result with readlane, and the block body has no
// lane-dependent statement, so the loop's exit is wave-uniform by
// construction: the index never comes from a broadcast that assumes the lanes
// are converged.
template <bool SPLIT>
__device__ __forceinline__ void t1_queue_loop(const T1Job *__restrict__ jobs, int njobs, int *__restrict__ counter)
{
    __shared__ uint8_t lut[LUT_BYTES];
    for (;;) {
        const int got = atomicAdd(counter, threadIdx.x == 0 ? 1 : 0);
        const int i = __builtin_amdgcn_readlane(got, 0);
        if (i >= njobs) break;
        t1_block<SPLIT>(jobs[i], lut);
    }
}

__global__ void __launch_bounds__(64) t1_encode_queue_kernel(const T1Job *__restrict__ jobs, int njobs,
                                                             int *__restrict__ counter)
{
    t1_queue_loop<false>(jobs, njobs, counter);
}

__global__ void __launch_bounds__(64) t1_cm_queue_kernel(const T1Job *__restrict__ jobs, int njobs,
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
__global__ void __launch_bounds__(64) t1_cm_kernel(const int32_t *__restrict__ q, int pitch, size_t plane_stride,
                                                   const rg_cb_desc *__restrict__ desc, uint8_t *__restrict__ syms,
                                                   int symcap, int4 *__restrict__ cm, int32_t *__restrict__ poffs,
                                                   int style)
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
    __shared__ uint8_t lut[LUT_BYTES];
    t1_block<true>(job, lut);
}

__global__ void __launch_bounds__(64) t1_cm_jobs_kernel(const T1Job *__restrict__ jobs)
{
    __shared__ uint8_t lut[LUT_BYTES];
    t1_block<true>(jobs[blockIdx.x], lut);
}

__global__ void __launch_bounds__(64) t1_mq_kernel(const uint8_t *__restrict__ syms, int symcap,
                                                   const int4 *__restrict__ cm, int ncb, uint8_t *__restrict__ slots,
                                                   int slot, rg_cb_result *__restrict__ res)
{
    __shared__ uint32_t tab[48];
    __shared__ uint8_t ring[RING_BYTES];
    t1_mq_lanes<false>(tab, ring, nullptr, blockIdx.x * 64 + threadIdx.x, ncb, syms, symcap, cm, slots, slot, res);
}

__global__ void __launch_bounds__(64) t1_mq_jobs_kernel(const T1Job *__restrict__ jobs, int njobs)
{
    __shared__ uint32_t tab[48];
    __shared__ uint8_t ring[RING_BYTES];
    t1_mq_lanes<true>(tab, ring, jobs, blockIdx.x * 64 + threadIdx.x, njobs, nullptr, 0, nullptr, nullptr, 0,
                      nullptr);
}

// ---- parallel style: grid (block groups of 64) x (pass index) ----
// Row q codes pass q of 64 blocks; under BYPASS a raw row (pass_is_raw, the
// same for the whole workgroup) packs bits instead of running MQ coders.
__global__ void __launch_bounds__(64) t1_mqp_kernel(const uint8_t *__restrict__ syms, int symcap,
                                                    const int4 *__restrict__ cm, const int32_t *__restrict__ poffs,
                                                    int ncb, uint8_t *__restrict__ arena, size_t arena_stride,
                                                    int32_t *__restrict__ seglens, int style)
{
    __shared__ uint32_t tab[48];
    __shared__ uint8_t ring[RING_BYTES];
    const int b = blockIdx.x * 64 + threadIdx.x;
    bool active = b < ncb;
    int p = 0, s0 = 0, s1 = 0, at = 0, cap = 0;
    const int q = (int)blockIdx.y;
    if (active) active = pass_range(cm[b], poffs + (size_t)b * MAXP, q, p, s0, s1, at, cap);
    const size_t bb = active ? (size_t)b : 0;
    if (pass_is_raw(style, q))
        t1_raw_pass_lane(active, syms + bb * symcap, s0, s1, arena + bb * arena_stride + at, cap,
                         seglens + bb * MAXP + p);
    else
        t1_mq_pass_lane(tab, ring, active, syms + bb * symcap, s0, s1, arena + bb * arena_stride + at, cap,
                        seglens + bb * MAXP + p);
}

__global__ void __launch_bounds__(64) t1_mqp_jobs_kernel(const T1Job *__restrict__ jobs, int njobs, int style)
{
    __shared__ uint32_t tab[48];
    __shared__ uint8_t ring[RING_BYTES];
    const int b = blockIdx.x * 64 + threadIdx.x;
    bool active = b < njobs;
    int p = 0, s0 = 0, s1 = 0, at = 0, cap = 0;
    const T1Job &j = jobs[active ? b : 0];
    const int q = (int)blockIdx.y;
    if (active) active = pass_range(*j.cm, j.poff, q, p, s0, s1, at, cap);
    if (pass_is_raw(style, q))
        t1_raw_pass_lane(active, j.sym, s0, s1, j.seg + at, cap, j.seglen + p);
    else
        t1_mq_pass_lane(tab, ring, active, j.sym, s0, s1, j.seg + at, cap, j.seglen + p);
}

// ---- parallel styles, sorted form (see t1_chain_lane) ----
// Block b's record from the encoder's flat per-block arrays.
struct FlatJobs {
    const uint8_t *syms;
    int symcap;
    const int4 *cm;
    const int32_t *poffs;
    uint8_t *arena;
    size_t astride;
    int32_t *seglens;
    __device__ T1Job operator()(int b) const
    {
        T1Job j{};
        j.sym = const_cast<uint8_t *>(syms) + (size_t)b * symcap;
        j.symcap = symcap;
        j.cm = const_cast<i32x4 *>(cm + b);
        j.poff = const_cast<int32_t *>(poffs) + (size_t)b * MAXP;
        j.seg = arena + (size_t)b * astride;
        j.seglen = seglens + (size_t)b * MAXP;
        return j;
    }
};
// Job b of a job list (batched path).
struct ListJobs {
    const T1Job *jobs;
    __device__ const T1Job &operator()(int b) const { return jobs[b]; }
};

__global__ void __launch_bounds__(256) chain_zero_kernel(int *ws)
{
    for (int i = threadIdx.x; i < CH_LIST; i += 256) ws[i] = 0;
}

// Histogram of chain keys, privatised per workgroup in LDS.
template <class J>
__global__ void __launch_bounds__(256) chain_hist_kernel(J job, int nblocks, int style, int *ws)
{
    __shared__ int h[CH_HIST];
    for (int k = threadIdx.x; k < CH_HIST; k += 256) h[k] = 0;
    __syncthreads();
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < (long)nblocks * MAXP) {
        const T1Job &jb = job((int)(i / MAXP));
        const int key = chain_key(*jb.cm, jb.poff, (int)(i % MAXP), style);
        if (key >= 0) atomicAdd(&h[key], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CH_HIST; k += 256)
        if (h[k]) atomicAdd(&ws[k], h[k]);
}

// Exclusive prefix sum of the histogram into the cursors; list lengths.
__global__ void __launch_bounds__(CH_HIST) chain_scan_kernel(int *ws)
{
    __shared__ int a[CH_HIST];
    const int t = threadIdx.x;
    const int v = ws[t];
    a[t] = v;
    __syncthreads();
    for (int o = 1; o < CH_HIST; o <<= 1) {
        const int x = t >= o ? a[t - o] : 0;
        __syncthreads();
        a[t] += x;
        __syncthreads();
    }
    ws[CH_HIST + t] = a[t] - v;
    if (t == CH_BUCKETS - 1) ws[CH_COUNTS] = a[t];
    if (t == CH_HIST - 1) ws[CH_COUNTS + 1] = a[t] - a[CH_BUCKETS - 1];
}

// Places every chain id at its bucket's cursor: one global atomic per
// (workgroup, bucket) reserves a range, LDS atomics number the entries in it.
template <class J>
__global__ void __launch_bounds__(256) chain_scatter_kernel(J job, int nblocks, int style, int *ws)
{
    __shared__ int h[CH_HIST];
    for (int k = threadIdx.x; k < CH_HIST; k += 256) h[k] = 0;
    __syncthreads();
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    int key = -1, local = 0;
    if (i < (long)nblocks * MAXP) {
        const T1Job &jb = job((int)(i / MAXP));
        key = chain_key(*jb.cm, jb.poff, (int)(i % MAXP), style);
        if (key >= 0) local = atomicAdd(&h[key], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CH_HIST; k += 256)
        if (h[k]) h[k] = atomicAdd(&ws[CH_HIST + k], h[k]);
    __syncthreads();
    if (key >= 0) ws[CH_LIST + h[key] + local] = (int)i;
}

// Codes the sorted lists: MQ wavefronts first (longest chains first), then
// raw ones; workgroups past the end of the lists exit.
template <class J>
__global__ void __launch_bounds__(64) chain_code_kernel(J job, const int *ws)
{
    __shared__ uint32_t tab[48];
    __shared__ uint8_t ring[RING_BYTES];
    const int nmq = ws[CH_COUNTS], ntot = nmq + ws[CH_COUNTS + 1];
    const int mqw = (nmq + 63) >> 6;
    const int w = (int)blockIdx.x;
    const bool raw = w >= mqw;
    const int base = raw ? nmq + (w - mqw) * 64 : w * 64;
    const int end = raw ? ntot : nmq;
    if (base >= end) return;
    // The kernel lasts as long as its longest chain, whose wavefront shares a
    // SIMD with shorter ones: issue priority by the length of the wavefront's
    // first (longest) chain lets the long chains run at their own latency
    // while short wavefronts fill their stalls.
    {
        const int id = ws[CH_LIST + base];
        const T1Job &jb = job(id / MAXP);
        int p, s0, s1, at, cap;
        const int len = pass_range(*jb.cm, jb.poff, id % MAXP, p, s0, s1, at, cap) ? s1 - s0 : 0;
        if (len >= 2048) __builtin_amdgcn_s_setprio(3);
        else if (len >= 1024) __builtin_amdgcn_s_setprio(2);
        else if (len >= 256) __builtin_amdgcn_s_setprio(1);
    }
    t1_chain_lane(tab, ring, raw, ws + CH_LIST, base, end, job);
}

template <class J>
int launch_chains(hipStream_t s, const J &job, int nblocks, int style, int *ws)
{
    const long n = (long)nblocks * MAXP;
    const unsigned g = (unsigned)((n + 255) / 256);
    chain_zero_kernel<<<1, 256, 0, s>>>(ws);
    RG_LAUNCH_CHECK();
    chain_hist_kernel<<<g, 256, 0, s>>>(job, nblocks, style, ws);
    RG_LAUNCH_CHECK();
    chain_scan_kernel<<<1, CH_HIST, 0, s>>>(ws);
    RG_LAUNCH_CHECK();
    chain_scatter_kernel<<<g, 256, 0, s>>>(job, nblocks, style, ws);
    RG_LAUNCH_CHECK();
    // at most nblocks * (MAXP - 2) chains; +1 wavefront for the raw list's
    // unaligned start
    chain_code_kernel<<<(unsigned)(((long)nblocks * (MAXP - 2) + 63) / 64 + 2), 64, 0, s>>>(job, ws);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

__global__ void __launch_bounds__(64) t1_pack_kernel(const int4 *__restrict__ cm, const int32_t *__restrict__ poffs,
                                                     const uint8_t *__restrict__ arena, size_t arena_stride,
                                                     const int32_t *__restrict__ seglens, uint8_t *__restrict__ slots,
                                                     int slot, rg_cb_result *__restrict__ res)
{
    const size_t b = blockIdx.x;
    t1_pack_block(cm[b], poffs + b * MAXP, arena + b * arena_stride, seglens + b * MAXP, slots + b * slot, slot,
                  res + b);
}

__global__ void __launch_bounds__(64) t1_pack_jobs_kernel(const T1Job *__restrict__ jobs)
{
    const T1Job &j = jobs[blockIdx.x];
    t1_pack_block(*j.cm, j.poff, j.seg, j.seglen, j.out, j.cap, j.res);
}

}  // namespace

int rg_t1_maxp() { return MAXP; }
size_t rg_t1_seg_bytes(int symcap) { return (size_t)seg_bytes(symcap); }
size_t rg_t1_chain_ws_bytes(int nblocks) { return sizeof(int) * ((size_t)CH_LIST + (size_t)nblocks * MAXP); }

int rg_launch_t1_parallel(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc,
                          int ncb, uint8_t *syms, int symcap, int4 *cm, int32_t *poffs, uint8_t *arena,
                          int32_t *seglens, uint8_t *out, int slot_bytes, rg_cb_result *res, int style,
                          int32_t *chain_ws)
{
    if (ncb <= 0) return RG_SUCCESS;
    const size_t astride = rg_t1_seg_bytes(symcap);
    t1_cm_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(q, pitch, plane_stride, desc, syms, symcap, cm, poffs,
                                                          style);
    RG_LAUNCH_CHECK();
    if (chain_ws) {
        const FlatJobs fj{syms, symcap, cm, poffs, arena, astride, seglens};
        if (int rc = launch_chains(s, fj, ncb, style, chain_ws)) return rc;
    } else {
        t1_mqp_kernel<<<dim3((unsigned)((ncb + 63) / 64), (unsigned)(MAXP - 2)), dim3(64), 0, s>>>(
            syms, symcap, cm, poffs, ncb, arena, astride, seglens, style);
        RG_LAUNCH_CHECK();
    }
    t1_pack_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(cm, poffs, arena, astride, seglens, out, slot_bytes, res);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_t1_jobs_parallel(hipStream_t s, const void *jobs, int njobs, int style, int32_t *chain_ws)
{
    if (njobs <= 0) return RG_SUCCESS;
    t1_cm_jobs_kernel<<<dim3((unsigned)njobs), dim3(64), 0, s>>>((const T1Job *)jobs);
    RG_LAUNCH_CHECK();
    if (chain_ws) {
        if (int rc = launch_chains(s, ListJobs{(const T1Job *)jobs}, njobs, style, chain_ws)) return rc;
    } else {
        t1_mqp_jobs_kernel<<<dim3((unsigned)((njobs + 63) / 64), (unsigned)(MAXP - 2)), dim3(64), 0, s>>>(
            (const T1Job *)jobs, njobs, style);
        RG_LAUNCH_CHECK();
    }
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
                                                          0);
    RG_LAUNCH_CHECK();
    t1_mq_kernel<<<dim3((unsigned)((ncb + 63) / 64)), dim3(64), 0, s>>>(syms, symcap, cm, ncb, out, slot_bytes, res);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_t1_jobs_split(hipStream_t s, const void *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    t1_cm_jobs_kernel<<<dim3((unsigned)njobs), dim3(64), 0, s>>>((const T1Job *)jobs);
    RG_LAUNCH_CHECK();
    t1_mq_jobs_kernel<<<dim3((unsigned)((njobs + 63) / 64)), dim3(64), 0, s>>>((const T1Job *)jobs, njobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_t1_job_bytes() { return (int)sizeof(T1Job); }

// Persistent forms: reset the queue counter, then `nwaves` workgroups pull
// code-blocks until the job list is exhausted.
int rg_launch_t1_queue(hipStream_t s, const void *jobs, int njobs, int style, int *counter, int nwaves,
                       int32_t *chain_ws)
{
    if (njobs <= 0) return RG_SUCCESS;
    reset_counter_kernel<<<1, 1, 0, s>>>(counter);
    RG_LAUNCH_CHECK();
    const dim3 g((unsigned)nwaves), b(64);
    const T1Job *J = (const T1Job *)jobs;
    if (style & RG_STYLE_TERMALL) {
        t1_cm_queue_kernel<<<g, b, 0, s>>>(J, njobs, counter);
        RG_LAUNCH_CHECK();
        if (chain_ws) {
            if (int rc = launch_chains(s, ListJobs{J}, njobs, style, chain_ws)) return rc;
        } else {
            t1_mqp_jobs_kernel<<<dim3((unsigned)((njobs + 63) / 64), (unsigned)(MAXP - 2)), b, 0, s>>>(J, njobs,
                                                                                                      style);
            RG_LAUNCH_CHECK();
        }
        RG_LAUNCH_CHECK();
        t1_pack_jobs_kernel<<<dim3((unsigned)njobs), b, 0, s>>>(J);
    } else if (rg_t1_inline()) {
        t1_encode_queue_kernel<<<g, b, 0, s>>>(J, njobs, counter);
    } else {
        t1_cm_queue_kernel<<<g, b, 0, s>>>(J, njobs, counter);
        RG_LAUNCH_CHECK();
        t1_mq_jobs_kernel<<<dim3((unsigned)((njobs + 63) / 64)), b, 0, s>>>(J, njobs);
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
