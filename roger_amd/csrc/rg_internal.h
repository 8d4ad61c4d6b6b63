// Internal declarations shared by the HIP sources of libroger_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "../../include/roger_amd.h"

struct rg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    // side streams of the overlapped phased tier-1, shared by the context's
    // encoders (created on first use): streams per encoder would outnumber
    // the hardware queues and serialise the overlap
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    // host-to-device copies of the encoders' host sources (created on first
    // use): an upload there overlaps the kernels of the previous encode
    hipStream_t upload = nullptr;
};

// Thread-local last-error message; set by rg_fail().
int rg_fail(int code, const char *fmt, ...);

#define RG_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return rg_fail(RG_ERR_DEVICE, "%s failed: %s (%s:%d)", #call,            \
                           hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

#define RG_CHECK(cond, ...)                                                          \
    do {                                                                             \
        if (!(cond)) return rg_fail(RG_ERR_INVALID, __VA_ARGS__);                    \
    } while (0)

#define RG_LAUNCH_CHECK() RG_HIP(hipGetLastError())

// Whole-sample symmetric extension index for a signal of length n >= 1.
__host__ __device__ inline int rg_mirror(int i, int n)
{
    if (n == 1) return 0;
    int p = 2 * (n - 1);
    i %= p;
    if (i < 0) i += p;
    return i >= n ? p - i : i;
}

// Band geometry of the Mallat layout (origin 0): level l region sizes.
struct rg_level_dims {
    int w[33], h[33];
};
inline rg_level_dims rg_dims(int w, int h, int levels)
{
    rg_level_dims d;
    d.w[0] = w; d.h[0] = h;
    for (int l = 1; l <= levels && l < 33; ++l) {
        d.w[l] = (d.w[l - 1] + 1) / 2;
        d.h[l] = (d.h[l - 1] + 1) / 2;
    }
    return d;
}

// Inter-level LL buffers of the multi-level DWT (forward and inverse): the LL
// band after level l (l >= 1; the input of level l+1, ceil(w/2^l) x
// ceil(h/2^l) samples per plane, nplanes planes, 4 bytes a sample) has its own
// 256-byte-aligned region of the scratch buffer, so every level's LL stays
// readable after the transform (the reference's dwtIn pyramid,
// src/OCLMemoryManager.cpp:124-134).  About a third of the input's size.
inline size_t rg_ll_offset(int w, int h, int nplanes, int level)
{
    size_t off = 0;
    long lw = w, lh = h;
    for (int l = 1; l < level; ++l) {
        lw = (lw + 1) / 2;
        lh = (lh + 1) / 2;
        off += ((size_t)lw * lh * nplanes * 4 + 255) / 256 * 256;
    }
    return off;
}

// Kernel launchers (defined in the .hip files).
int rg_launch_colour_fwd(hipStream_t s, const uint8_t *src, int w, int h, int ncomp, int colour,
                         int out_float, void *planes);
int rg_launch_colour_inv(hipStream_t s, const void *planes, int w, int h, int ncomp, int colour,
                         int in_float, uint8_t *dst);
int rg_launch_planar_split_i16(hipStream_t s, const int16_t *src, int w, int h, int ncomp, int16_t *dst);

// One forward level: input region cw x ch (pitch in_pitch, plane stride in_ps),
// LL to ll_out (pitch ll_pitch, plane stride ll_ps), high bands into the
// Mallat plane q (pitch qp, plane stride q_ps) at offsets (wl, hl).
struct rg_fwd_level_args {
    int filter;             // RG_FILTER_53 / RG_FILTER_97 (read by the batched kernel)
    const void *in;
    long in_pitch, in_ps;
    int in_half;            // input LL stored as fp16 (9/7)
    void *ll_out;           // next-level LL (float / half / int32); nullptr: LL goes to q
    long ll_pitch, ll_ps;
    int ll_half;
    void *q;                // int32 indices (or float coefficients when q_float)
    long qp, q_ps;
    int q_float;            // write float coefficients (no quantisation)
    int cw, ch, nplanes;
    float inv_ll, inv_hl, inv_lh, inv_hh;  // quantisation (9/7 when !q_float)
};
int rg_launch_fwd53_level(hipStream_t s, const rg_fwd_level_args &a);
int rg_launch_fwd97_level(hipStream_t s, const rg_fwd_level_args &a);

// First forward level fused with the colour stage: the interleaved uint8
// image (w x h x ncomp, 4-byte aligned rows of w * ncomp bytes) is level
// shifted and colour transformed in registers and lifted at once; every
// component's level-1 outputs go out as with rg_fwd_level_args (LL to
// ll_out, or quantised with the bands into q when ll_out is null).
struct rg_fwd1c_args {
    const uint8_t *src;
    int w, h, ncomp, colour, filter;
    void *ll_out;           // float (9/7) / int32 (5/3)
    long ll_pitch, ll_ps;
    int32_t *q;
    long qp, q_ps;
    float inv_ll, inv_hl, inv_lh, inv_hh;  // 9/7 quantisation of level 1
};
// Shapes the fused kernel takes: 1 or 3 components, w * ncomp a multiple of
// 4, both sides longer than the filter's halo (4 for 9/7, 2 for 5/3).
bool rg_fwd1_colour_ok(int w, int h, int ncomp, int colour, int filter);
int rg_launch_fwd1_colour(hipStream_t s, const rg_fwd1c_args &a);

// One inverse level: reads the four bands of a cw x ch region (LL from ll_in,
// pitch ll_pitch; high bands from the Mallat plane q), writes the region to out.
struct rg_inv_level_args {
    int filter;
    const void *ll_in;
    long ll_pitch, ll_ps;
    int ll_is_q;            // LL read from q (coarsest level) with dequantisation
    const void *q;
    long qp, q_ps;
    int q_float;            // q holds float coefficients
    void *out;
    long out_pitch, out_ps;
    int cw, ch, nplanes;
    float s_ll, s_hl, s_lh, s_hh; // dequantisation steps (9/7 with int q)
};
int rg_launch_inv53_level(hipStream_t s, const rg_inv_level_args &a);
int rg_launch_inv97_level(hipStream_t s, const rg_inv_level_args &a);

// Last inverse level fused with the inverse colour stage: the level-1 bands
// of every component (LL from ll_in, or dequantised from q when ll_is_q) are
// synthesised in workgroup memory, colour transformed back and written as
// the interleaved uint8 image (w x h x ncomp, rows of w * ncomp bytes).
struct rg_inv1c_args {
    const void *ll_in;      // float (9/7) / int32 (5/3)
    long ll_pitch, ll_ps;
    int ll_is_q;
    const int32_t *q;
    long qp, q_ps;
    int w, h, ncomp, colour, filter;
    float s_ll, s_hl, s_lh, s_hh;   // 9/7 dequantisation steps of level 1
    uint8_t *dst;
};
// Shapes the fused kernel takes: as rg_fwd1_colour_ok.
bool rg_inv1_colour_ok(int w, int h, int ncomp, int colour, int filter);
int rg_launch_inv1_colour(hipStream_t s, const rg_inv1c_args &a);

// Tile shapes of the level kernels (the batched job tables use them too).
#define RG_DWT_FWD_TW 128
#define RG_DWT_FWD_TH 16
#define RG_DWT_INV_TW 64
#define RG_DWT_INV_TH 64
int rg_launch_fwd_level_batch(hipStream_t s, const rg_fwd_level_args *args, const int4 *jobs, int njobs);
int rg_launch_inv_level_batch(hipStream_t s, const rg_inv_level_args *args, const int4 *jobs, int njobs);

int rg_launch_t1_encode(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride,
                        const rg_cb_desc *desc, int ncb, uint8_t *out, int slot_bytes,
                        rg_cb_result *res);

// Batched tier-1: job records built on the host with rg_t1_job_fill.
int rg_t1_job_bytes();
void rg_t1_job_fill(void *dst, const int32_t *src, int pitch, int w, int h, int orient, uint8_t *out, int cap,
                    rg_cb_result *res, uint8_t *sym, int symcap, int4 *cm, int32_t *poff, uint8_t *seg,
                    int32_t *seglen, int style);
int rg_launch_t1_jobs(hipStream_t s, const void *jobs, int njobs);
// Split tier-1: context modelling writes per-block symbol streams (symcap
// bytes each), then one lane per block runs the MQ coder.
bool rg_t1_inline();
int rg_t1_symcap(int cbw, int cbh);
int rg_launch_t1_split(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc, int ncb,
                       uint8_t *syms, int symcap, int4 *cm, uint8_t *out, int slot_bytes, rg_cb_result *res);
int rg_launch_t1_jobs_split(hipStream_t s, const void *jobs, int njobs);
// Parallel code-block styles (RESET | TERMALL, optionally | BYPASS): context
// modelling, one MQ (or raw) lane per (block, pass), then concatenation of
// the pass segments into the slots.
// Per block: MAXP pass offsets / segment lengths, rg_t1_seg_bytes(symcap)
// bytes of segment arena.  chain_ws (rg_t1_chain_ws_bytes(blocks) bytes, or
// null for the pass-row grid): workspace of the sorted chain lists.
int rg_t1_maxp();
size_t rg_t1_seg_bytes(int symcap);
size_t rg_t1_chain_ws_bytes(int nblocks);
// Overlapped tier-1 (parallel styles, plain path): context modelling runs in
// phases that end before pass RG_T1_PHASE_SPLIT1 (three-phase form only) and
// before pass RG_T1_PHASE_SPLIT; the chains of each early phase are listed on
// s and coded on their own stream (se[k], with a chain workspace ws[k] and the
// events ev_in[k] / ev_out[k]) while s models the next passes.  cms: 64 x 16
// bytes per block of modelling state kept between the phases.
#define RG_T1_PHASE_SPLIT0 4   // four-phase form only
#define RG_T1_PHASE_SPLIT1 7
#define RG_T1_PHASE_SPLIT 16
void rg_t1_splits(int *s1, int *s2);
// number of modelling phases of a parallel code-block style: RG_T1_PHASES
// (1-4) if set, RG_T1_TWO_PHASE=0 one; else two with BYPASS and one without
int rg_t1_phases(int style);
constexpr int RG_T1_MAX_EARLY = 3;   // early phases (each with a side stream)
struct rg_t1_overlap {
    int nearly;              // early phases: 1 .. RG_T1_MAX_EARLY
    hipStream_t se[RG_T1_MAX_EARLY];
    hipEvent_t ev_in[RG_T1_MAX_EARLY], ev_out[RG_T1_MAX_EARLY];
    int32_t *ws[RG_T1_MAX_EARLY];
    void *cms;
};
// The encoder's flat per-block tier-1 arrays, for the chain kernels
// (t1_chain.hip); astride = rg_t1_seg_bytes(symcap).
struct rg_t1_flat {
    const uint8_t *syms;
    int symcap;
    const int4 *cm;
    const int32_t *poffs;
    uint8_t *arena;
    size_t astride;
    int32_t *seglens;
};
// Chain lists and coding of passes [qlo, qhi) of every block (t1_chain.hip).
int rg_launch_chains_flat(hipStream_t s, const rg_t1_flat &f, int nblocks, int style, int32_t *ws, int qlo, int qhi);
// The two halves of rg_launch_chains_flat: the sorted lists, then the coding.
int rg_launch_chain_lists_flat(hipStream_t s, const rg_t1_flat &f, int nblocks, int style, int32_t *ws, int qlo,
                               int qhi);
int rg_launch_chain_code_flat(hipStream_t s, const rg_t1_flat &f, int nblocks, const int32_t *ws);
int rg_launch_chains_jobs(hipStream_t s, const void *jobs, int njobs, int style, int32_t *ws, int qlo, int qhi);
// Standard-style lane MQ coders (t1_mq.hip).
int rg_launch_t1_mq(hipStream_t s, const uint8_t *syms, int symcap, const int4 *cm, int ncb, uint8_t *out,
                    int slot_bytes, rg_cb_result *res);
int rg_launch_t1_mq_jobs(hipStream_t s, const void *jobs, int njobs);
int rg_launch_t1_parallel(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc,
                          int ncb, uint8_t *syms, int symcap, int4 *cm, int32_t *poffs, uint8_t *arena,
                          int32_t *seglens, uint8_t *out, int slot_bytes, rg_cb_result *res, int style,
                          int32_t *chain_ws, const rg_t1_overlap *ov, uint64_t *sppm = nullptr);
// Rate control (t1_rate.hip): per-pass distortions from the coefficients and
// the context modeller's significance-pass masks (sppm, rg_rd_sppm_bytes),
// convex hulls, a slope threshold for `target` code-block bytes, and the
// results cut at it.  kept_out / err_out (optional) wait for the stream.
size_t rg_rd_ws_bytes(int ncb);
size_t rg_rd_sppm_bytes(int ncb);
int rg_launch_rate_control(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc,
                           int ncb, const int4 *cm, const uint64_t *sppm, const int32_t *seglens,
                           const double *weights, int nlayers, const long *targets, void *ws, int32_t *layer_np,
                           rg_cb_result *res, long *kept_out, int *err_out);
int rg_launch_t1_jobs_parallel(hipStream_t s, const void *jobs, int njobs, int style, int32_t *chain_ws);
// Persistent batched tier-1 (any style): nwaves workgroups pull job indices
// from *counter (reset by a kernel inside the launch sequence).
int rg_launch_t1_queue(hipStream_t s, const void *jobs, int njobs, int style, int *counter, int nwaves,
                       int32_t *chain_ws);

// Batched colour: per-image parameters; jobs {image, first chunk} cover 256
// chunks of 16 pixels each.
struct rg_colour_job_params {
    const uint8_t *src;
    long npix;
    int ncomp, colour, out_float;
    void *planes;
};
int rg_launch_colour_fwd_batch(hipStream_t s, const rg_colour_job_params *params, const int2 *jobs, int njobs);
