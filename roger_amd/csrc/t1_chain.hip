// Parallel code-block styles (RESET | TERMALL, optionally | BYPASS), second
// half of the split tier-1: every (code-block, pass) chain is listed longest
// first by a device counting sort and coded one per lane, 64 per two-wavefront
// workgroup (t1_chain_lane in t1_core.h).  Replaces, with t1_encode.hip, the
// reference's bit-plane coder launch src/OCLBPC.cpp:38-51.
#include "rg_internal.h"
#include "wave_hip.h"
#include "t1_core.h"

using namespace rg_t1;

namespace {

// Block b's record from the encoder's flat per-block arrays.
struct FlatJobs {
    rg_t1_flat f;
    __device__ T1Job operator()(int b) const
    {
        T1Job j{};
        j.sym = const_cast<uint8_t *>(f.syms) + (size_t)b * f.symcap;
        j.symcap = f.symcap;
        j.cm = const_cast<i32x4 *>(f.cm + b);
        j.poff = const_cast<int32_t *>(f.poffs) + (size_t)b * MAXP;
        j.seg = f.arena + (size_t)b * f.astride;
        j.seglen = f.seglens + (size_t)b * MAXP;
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

// (block, pass) items per workgroup of the list kernels: 256 threads x
// CH_ITEMS, so that each workgroup's LDS histogram covers enough items to keep
// the global atomics per bucket few.
constexpr int CH_ITEMS = 16;

// Key of item t of this thread (items of a workgroup are consecutive, the
// threads interleaved for coalesced pass-offset loads); -1 if none.
template <class J>
__device__ __forceinline__ int chain_item_key(const J &job, int nblocks, int style, int qlo, int qhi, int t, long &i)
{
    i = ((long)blockIdx.x * CH_ITEMS + t) * 256 + threadIdx.x;
    if (i >= (long)nblocks * MAXP) return -1;
    const int q = (int)(i % MAXP);
    if (q < qlo || q >= qhi) return -1;
    const T1Job &jb = job((int)(i / MAXP));
    return chain_key(*jb.cm, jb.poff, q, style, qlo, qhi);
}

// Histogram of chain keys, privatised per workgroup in LDS.
template <class J>
__global__ void __launch_bounds__(256) chain_hist_kernel(J job, int nblocks, int style, int qlo, int qhi, int *ws)
{
    __shared__ int h[CH_HIST];
    for (int k = threadIdx.x; k < CH_HIST; k += 256) h[k] = 0;
    __syncthreads();
#pragma unroll 4
    for (int t = 0; t < CH_ITEMS; ++t) {
        long i;
        const int key = chain_item_key(job, nblocks, style, qlo, qhi, t, i);
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
__global__ void __launch_bounds__(256) chain_scatter_kernel(J job, int nblocks, int style, int qlo, int qhi, int *ws)
{
    __shared__ int h[CH_HIST];
    for (int k = threadIdx.x; k < CH_HIST; k += 256) h[k] = 0;
    __syncthreads();
    int key[CH_ITEMS], local[CH_ITEMS];
    long id[CH_ITEMS];
#pragma unroll
    for (int t = 0; t < CH_ITEMS; ++t) {
        key[t] = chain_item_key(job, nblocks, style, qlo, qhi, t, id[t]);
        local[t] = key[t] >= 0 ? atomicAdd(&h[key[t]], 1) : 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CH_HIST; k += 256)
        if (h[k]) h[k] = atomicAdd(&ws[CH_HIST + k], h[k]);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < CH_ITEMS; ++t)
        if (key[t] >= 0) ws[CH_LIST + h[key[t]] + local[t]] = (int)id[t];
}

// Codes the sorted lists with two-wavefront workgroups: MQ workgroups first
// (longest chains first), 64 chains each, every chain coded by lane l of both
// wavefronts (interval side and code side, ring_chain_wg); then raw ones, 64
// chains per wavefront; workgroups and raw wavefronts past the end of the
// lists exit (a raw wavefront has no barrier to hold the other up).
template <class J>
__global__ void __launch_bounds__(128) chain_code_kernel(J job, const int *ws)
{
    __shared__ uint32_t tab[MQ_TT_WORDS];
    __shared__ __attribute__((aligned(16))) uint8_t ring[MQ2_LDS_BYTES];
    static_assert(MQ2_LDS_BYTES >= 2 * RING_BYTES, "one ring per raw wavefront");
    const int nmq = ws[CH_COUNTS], ntot = nmq + ws[CH_COUNTS + 1];
    const int mqw = (nmq + 63) >> 6;
    const int w = (int)blockIdx.x;
    const bool raw = w >= mqw;
    const int base = raw ? nmq + (w - mqw) * 128 + 64 * wv_wave() : w * 64;
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
    if (raw) t1_chain_lane<J, false>(tab, ring + wv_wave() * RING_BYTES, true, ws + CH_LIST, base, end, job);
    else t1_chain_lane<J, true>(tab, ring, false, ws + CH_LIST, base, end, job);
}

// Chain lists of the passes [qlo, qhi) of every block (counting sort).
template <class J>
int launch_lists(hipStream_t s, const J &job, int nblocks, int style, int *ws, int qlo, int qhi)
{
    const long n = (long)nblocks * MAXP;
    const unsigned g = (unsigned)((n + 256 * CH_ITEMS - 1) / (256 * CH_ITEMS));
    chain_zero_kernel<<<1, 256, 0, s>>>(ws);
    RG_LAUNCH_CHECK();
    chain_hist_kernel<<<g, 256, 0, s>>>(job, nblocks, style, qlo, qhi, ws);
    RG_LAUNCH_CHECK();
    chain_scan_kernel<<<1, CH_HIST, 0, s>>>(ws);
    RG_LAUNCH_CHECK();
    chain_scatter_kernel<<<g, 256, 0, s>>>(job, nblocks, style, qlo, qhi, ws);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// Coding of the listed chains: at most nblocks * (MAXP - 2) of them; +1
// wavefront for the raw list's unaligned start.
template <class J>
int launch_code(hipStream_t s, const J &job, int nblocks, const int *ws)
{
    chain_code_kernel<<<(unsigned)(((long)nblocks * (MAXP - 2) + 63) / 64 + 2), 128, 0, s>>>(job, ws);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

template <class J>
int launch_chains(hipStream_t s, const J &job, int nblocks, int style, int *ws, int qlo = 0, int qhi = MAXP)
{
    if (int rc = launch_lists(s, job, nblocks, style, ws, qlo, qhi)) return rc;
    return launch_code(s, job, nblocks, ws);
}

}  // namespace

size_t rg_t1_chain_ws_bytes(int nblocks) { return sizeof(int) * ((size_t)CH_LIST + (size_t)nblocks * MAXP); }

int rg_launch_chains_flat(hipStream_t s, const rg_t1_flat &f, int nblocks, int style, int32_t *ws, int qlo, int qhi)
{
    if (nblocks <= 0) return RG_SUCCESS;
    return launch_chains(s, FlatJobs{f}, nblocks, style, (int *)ws, qlo, qhi);
}

int rg_launch_chain_lists_flat(hipStream_t s, const rg_t1_flat &f, int nblocks, int style, int32_t *ws, int qlo,
                               int qhi)
{
    if (nblocks <= 0) return RG_SUCCESS;
    return launch_lists(s, FlatJobs{f}, nblocks, style, (int *)ws, qlo, qhi);
}

int rg_launch_chain_code_flat(hipStream_t s, const rg_t1_flat &f, int nblocks, const int32_t *ws)
{
    if (nblocks <= 0) return RG_SUCCESS;
    return launch_code(s, FlatJobs{f}, nblocks, (const int *)ws);
}

int rg_launch_chains_jobs(hipStream_t s, const void *jobs, int njobs, int style, int32_t *ws, int qlo, int qhi)
{
    if (njobs <= 0) return RG_SUCCESS;
    return launch_chains(s, ListJobs{(const T1Job *)jobs}, njobs, style, (int *)ws, qlo, qhi);
}
