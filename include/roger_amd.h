/*
 * roger_amd — C ABI of the MI355X (gfx950) JPEG 2000 encode/decode hot path.
 *
 * Plain C: opaque handles, POD structs, plain pointers and sizes.  No HIP or
 * torch types appear here; the C++ drop-in classes in include/roger/ (named
 * after the reference's OCLEncoder, OCLDecoder, OCLDWTForward, OCLDWTRev,
 * OCLBPC, OCLRGBtoPlanar, OCLMemoryManager) forward to these entry points.
 *
 * Return convention: 0 (RG_SUCCESS) on success, a negative RG_ERR_* code on
 * failure, with a message available from rg_last_error().  This replaces the
 * reference's cl_int tDeviceRC / DeviceSuccess (src/ocl_platform.h:22-24).
 * All "d_" pointers are device (HBM) addresses; kernel launches are
 * asynchronous on the context's stream, rg_context_sync() waits.
 */
#ifndef ROGER_AMD_H
#define ROGER_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RG_SUCCESS 0
#define RG_ERR_INVALID (-1)     /* bad argument or unsupported shape */
#define RG_ERR_DEVICE (-2)      /* HIP runtime error */
#define RG_ERR_OOM (-3)         /* device or pinned allocation failed */
#define RG_ERR_OVERFLOW (-4)    /* a code-block did not fit its output slot */
#define RG_ERR_NOT_READY (-5)   /* result requested before the stage ran */

typedef struct rg_context rg_context;
typedef struct rg_encoder rg_encoder;
typedef struct rg_decoder rg_decoder;

/* Sample types for planar host components. */
#define RG_F32 0
#define RG_I16 1
#define RG_I32 2

/* Colour transforms. */
#define RG_COLOUR_NONE 0 /* level shift only, per channel */
#define RG_COLOUR_RCT 1  /* reversible (integer) component transform */
#define RG_COLOUR_ICT 2  /* irreversible (FP32) component transform */

/* Wavelet filters. */
#define RG_FILTER_53 0
#define RG_FILTER_97 1

/* One code-block: rectangle in the Mallat-packed coefficient plane of
 * component `comp`, plus its band orientation (0 LL, 1 HL, 2 LH, 3 HH). */
typedef struct {
    int32_t x0, y0, w, h;
    int32_t comp, orient, band, res;
} rg_cb_desc;

/* Tier-1 output of one code-block. */
typedef struct {
    int32_t len;     /* bytes of the MQ codeword (all coding passes) */
    int32_t npasses; /* 3*nbp-2, or 0 for an all-zero block */
    int32_t nbp;     /* magnitude bit-planes coded (msb index + 1) */
    int32_t flags;   /* bit 0: codeword larger than the slot (error) */
} rg_cb_result;

/* Encoder parameters (T.800 codec choices of the native path). */
typedef struct {
    int32_t levels;        /* DWT decomposition levels (0..15) */
    int32_t filter;        /* RG_FILTER_53 or RG_FILTER_97 */
    int32_t colour;        /* RG_COLOUR_* (RCT with 5/3, ICT with 9/7) */
    int32_t cbw, cbh;      /* code-block size: 32 or 64 (w*h <= 4096) */
    double delta;          /* 9/7 base quantisation step (e.g. 1/64) */
    int32_t ll_fp16;       /* 9/7: store the inter-level LL band as FP16 */
    int32_t bitdepth;      /* sample precision (8) */
    int32_t style;         /* code-block style: 0, RG_STYLE_PARALLEL or RG_STYLE_PARALLEL_BYPASS */
} rg_encode_params;

/* Code-block styles (T.800 Table A.19).  RG_STYLE_PARALLEL = RESET | TERMALL:
 * every coding pass is its own MQ codeword segment starting from the initial
 * contexts, so the passes of a block code in parallel (one GPU lane each);
 * the codestream signals the style in COD and each segment's length in the
 * packet headers.  About 1% larger codestreams than style 0.
 * RG_STYLE_PARALLEL_BYPASS adds BYPASS: the significance-propagation and
 * refinement passes from the fifth bit-plane on are sent as raw bits (T.800
 * D.6) instead of MQ-coded -- less coder work, codestreams within about 1% of
 * RG_STYLE_PARALLEL (smaller on most images, since reset contexts code
 * near-random low bit-planes poorly). */
#define RG_STYLE_BYPASS 0x01
#define RG_STYLE_RESET 0x02
#define RG_STYLE_TERMALL 0x04
#define RG_STYLE_PARALLEL 0x06
#define RG_STYLE_PARALLEL_BYPASS 0x07
#define RG_MAX_PASSES 96   /* stride of per-block pass arrays (>= 3*32-2) */

/* ---- errors ---- */
const char *rg_error_string(int code);
const char *rg_last_error(void);

/* ---- context: device + stream (replaces ocl_args_d_t, src/OCLUtil.h:31-39,
 *      and OCLDeviceManager::init/getInfo, src/OCLDeviceManager.h:27-39) ---- */
int rg_context_create(int device, rg_context **out);
int rg_context_destroy(rg_context *ctx);             /* after every encoder, decoder, batch and buffer made on it */
int rg_context_sync(rg_context *ctx);                 /* OCLEncodeDecode::finish, src/OCLEncodeDecode.cpp:33-36 */
int rg_context_device_name(rg_context *ctx, char *buf, int cap);
/* 1 when all work queued on the context's stream has finished, 0 if not. */
int rg_context_query(rg_context *ctx);

/* ---- device memory plumbing ---- */
int rg_malloc(rg_context *ctx, size_t bytes, void **d_ptr);
int rg_free(rg_context *ctx, void *d_ptr);
int rg_memcpy_h2d(rg_context *ctx, void *d_dst, const void *src, size_t bytes);   /* stream-ordered, waits */
/* Stream-ordered and asynchronous (the source must stay unchanged until the
 * stream reaches the copy; pinned memory makes it asynchronous to the host). */
int rg_memcpy_h2d_async(rg_context *ctx, void *d_dst, const void *src, size_t bytes);
int rg_memcpy_d2h(rg_context *ctx, void *dst, const void *d_src, size_t bytes);   /* stream-ordered, waits */
int rg_memset(rg_context *ctx, void *d_dst, int value, size_t bytes);
int rg_memcpy_d2d(rg_context *ctx, void *d_dst, const void *d_src, size_t bytes);  /* asynchronous */
/* Pinned (page-locked) host memory, for host sources whose upload should run
 * asynchronously at full PCIe rate (rg_encoder_encode_u8 with a host source). */
int rg_host_alloc(rg_context *ctx, size_t bytes, void **host_ptr);
int rg_host_free(rg_context *ctx, void *host_ptr);
/* Streaming device copy kernel (16-byte accesses): the measured HBM ceiling. */
int rg_stream_copy(rg_context *ctx, void *d_dst, const void *d_src, size_t bytes);

/* ---- timing on the context stream (HIP events) ---- */
int rg_timer_start(rg_context *ctx);
int rg_timer_stop(rg_context *ctx, float *ms);        /* waits for the stop event */

/* ---- stage kernels (device pointers, asynchronous) ---- */

/* Interleaved uint8 (ncomp channels, w*h pixels) -> level-shifted planes.
 * colour RG_COLOUR_NONE: each channel shifted by -128 into its plane;
 * RCT: int32 planes Y,U,V; ICT: float planes Y,Cb,Cr.  out_float selects
 * float (1) or int32 (0) planes for NONE.  Planes are contiguous w*h each.
 * Replaces the host level shift + interleave of src/OCLTest.cpp:81-90 and
 * src/OCLMemoryManager.cpp:41-54 (the reference has no RCT/ICT). */
int rg_colour_fwd(rg_context *ctx, const uint8_t *d_src, int w, int h, int ncomp,
                  int colour, int out_float, void *d_planes);
/* Planes -> interleaved uint8 with inverse transform, +128 and clamping. */
int rg_colour_inv(rg_context *ctx, const void *d_planes, int w, int h, int ncomp,
                  int colour, int in_float, uint8_t *d_dst);

/* Interleaved int16 coefficient image (ncomp channels) -> ncomp int16 planes.
 * Replaces OCLRGBtoPlanar::run (src/OCLRGBtoPlanar.cpp:36-80,
 * src/CL/oclplanar.cl:23-39). */
int rg_planar_split_i16(rg_context *ctx, const int16_t *d_src, int w, int h, int ncomp,
                        int16_t *d_planes);

/* Scratch bytes the multi-level DWT entry points need for a w x h plane
 * (nplanes planes transformed together). */
size_t rg_dwt_scratch_bytes(int w, int h, int nplanes, int filter);

/* Forward multi-level DWT of nplanes planes (each w*h, contiguous, plane
 * stride w*h) into Mallat-packed int32 coefficient planes d_q.
 *  5/3: d_in int32 samples, d_q receives the integer coefficients.
 *  9/7: d_in float samples, d_q receives quantisation indices
 *       sign*floor(|y| * inv_step[band]); inv_step has 3*levels+1 entries
 *       in band order (LL_N, then HL, LH, HH for level N..1).
 * Replaces OCLDWTForward::run (src/OCLDWTForward.cpp:124-137) with its
 * per-level kernels src/CL/ocldwt53.cl:167-292 and
 * src/CL/ocldwt97.cl:487-717. */
int rg_dwt_fwd(rg_context *ctx, int filter, const void *d_in, int w, int h, int nplanes,
               int levels, const float *inv_step, int ll_fp16, int32_t *d_q, void *d_scratch);
/* rg_dwt_fwd with no allocation and no workspace argument.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The inter-level LL scratch is a static device array of 96 MiB
 * that every GPU the library runs on holds until the process ends (enough for
 * 4096x4096 x 3 planes or 8192x8192 x 1 plane); rg_dwt_scratch_bytes larger
 * than that is refused with RG_ERR_INVALID before any work is queued.  All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller.  The
 * array keeps the last call's contents; nothing needs to start at zero. */
int rg_dwt_fwd_noalloc(rg_context *ctx, int filter, const void *d_in, int w, int h, int nplanes,
                       int levels, const float *inv_step, int ll_fp16, int32_t *d_q);
/* Colour stage and forward transform in one: interleaved uint8 image d_src
 * (w x h x ncomp, 4-byte aligned; device memory) -> level shift, colour
 * transform (RG_COLOUR_RCT with 5/3, RG_COLOUR_ICT with 9/7, or NONE) ->
 * levels (>= 1) of DWT (+ 9/7 quantisation as rg_dwt_fwd) into d_q.  Where the
 * shape allows (1 or 3 components, w * ncomp a multiple of 4, both sides
 * longer than the filter halo: 4 samples for 9/7, 2 for 5/3) the first level
 * reads the image itself and the colour planes never go to memory;
 * otherwise the planes pass through d_scratch.  d_scratch holds
 * rg_dwt_fwd_u8_scratch_bytes.  The native encoder's front end
 * (rg_encoder_encode_u8); it replaces the reference's host level shift and
 * upload (src/OCLTest.cpp:81-90, src/OCLMemoryManager.cpp:41-54,150-160)
 * followed by OCLDWTForward::run (src/OCLDWTForward.cpp:124-137). */
size_t rg_dwt_fwd_u8_scratch_bytes(int w, int h, int ncomp, int colour, int filter);
int rg_dwt_fwd_u8(rg_context *ctx, const uint8_t *d_src, int w, int h, int ncomp, int colour, int filter,
                  int levels, const float *inv_step, int32_t *d_q, void *d_scratch);
/* 9/7 only: forward DWT without quantisation, float Mallat output
 * (the reference's outputDwt path, src/CL/ocldwt97.cl:196-420). */
int rg_dwt97_fwd_float(rg_context *ctx, const float *d_in, int w, int h, int nplanes, int levels,
                       float *d_out, void *d_scratch);
/* Inverse multi-level DWT (the decoder side of OCLDWTRev, src/OCLDWTRev.cpp:39-59).
 *  5/3: d_q int32 coefficients -> int32 samples in d_out.
 *  9/7: d_q quantisation indices, dequantised as sign*(|q|+0.5)*step[band]
 *       (step == NULL: d_q is read as float coefficients), -> float samples. */
int rg_dwt_inv(rg_context *ctx, int filter, const void *d_q, int w, int h, int nplanes,
               int levels, const float *step, void *d_out, void *d_scratch);
/* Inverse transform and inverse colour stage in one: levels (>= 1) of 5/3
 * coefficients or 9/7 quantisation indices d_q (dequantised with step, as
 * rg_dwt_inv) -> inverse colour transform (RG_COLOUR_RCT with 5/3,
 * RG_COLOUR_ICT with 9/7, or NONE) -> interleaved uint8 image d_dst
 * (w x h x ncomp, 16-byte aligned; device memory).  Where the shape allows (as
 * rg_dwt_fwd_u8) the last level synthesises every component of a tile in
 * workgroup memory and writes the uint8 pixels itself, so the FP32 planes
 * never go to memory; otherwise they pass through d_scratch.  d_scratch holds
 * rg_dwt_inv_u8_scratch_bytes.  The native decoder's back end
 * (rg_decoder_decode_codeblocks); it replaces the reference's inverse DWT
 * (src/OCLDWTRev.cpp:39-59) and the host-side level shift back, which the
 * reference never built. */
size_t rg_dwt_inv_u8_scratch_bytes(int w, int h, int ncomp, int colour, int filter);
int rg_dwt_inv_u8(rg_context *ctx, const int32_t *d_q, int w, int h, int ncomp, int colour, int filter,
                  int levels, const float *step, uint8_t *d_dst, void *d_scratch);

/* Code-block list for a w x h image with `levels` levels and cbw x cbh
 * blocks, per component in resolution / band / raster order.  Returns the
 * count per component; fills at most cap descriptors (host memory) for
 * component 0 (callers replicate per component with comp set). */
int rg_codeblocks(int w, int h, int levels, int cbw, int cbh, rg_cb_desc *out, int cap);

/* EBCOT tier-1 encode of ncb code-blocks (descriptors in device memory)
 * over int32 coefficient planes (plane c at d_q + c*plane_stride, row
 * pitch `pitch`).  Each block's codeword goes to d_out + i*slot_bytes.
 * With d_out 16-byte aligned and slot_bytes a multiple of 16 the split
 * tier-1 (lane-parallel MQ coder) runs, otherwise the single-kernel one;
 * both give the same bytes.
 * Replaces OCLBPC::run (src/OCLBPC.cpp:38-51, src/CL/oclbpc.cl:279-746). */
int rg_t1_encode(rg_context *ctx, const int32_t *d_q, int pitch, size_t plane_stride,
                 const rg_cb_desc *d_desc, int ncb, uint8_t *d_out, int slot_bytes,
                 rg_cb_result *d_res);

/* ---- encoder: owns device buffers (OCLMemoryManager) and runs the plain
 *      per-image path (OCLEncoder, src/OCLEncoder.h:27-37) ---- */
int rg_encoder_create(rg_context *ctx, rg_encoder **out);
int rg_encoder_destroy(rg_encoder *enc);
/* Native path: interleaved uint8 image (device pointer if src_on_device,
 * else host) -> colour -> DWT (+quantisation) -> tier-1, asynchronous on the
 * context's stream.  A host source is uploaded on the context's upload
 * stream into one of two device input buffers used in turn, so the upload
 * of one image overlaps the kernels of the encode before it (double-buffered
 * ingest, the design of the reference's unbuilt
 * src/OCLDataTransferManager.cpp:18-30, replacing the blocking upload of
 * src/OCLMemoryManager.cpp:150-160); the host memory must stay unchanged
 * until the encode has completed (rg_context_sync / rg_context_query).
 * Pinned memory (rg_host_alloc) makes the copy asynchronous to the host. */
int rg_encoder_encode_u8(rg_encoder *enc, const uint8_t *src, int src_on_device, int w, int h,
                         int ncomp, const rg_encode_params *p);
/* Drop-in path of OCLEncoder<T>::run(components, w, h, levels, precision):
 * ncomp planar host components of sample type RG_F32/RG_I16/RG_I32, already
 * level shifted by the caller (src/OCLTest.cpp:87).  lossy selects 9/7
 * (+ quantisation unless output_dwt) else 5/3; code-blocks cbw x cbh. */
int rg_encoder_run_planar(rg_encoder *enc, const void *const *components, int ncomp, int sample_type,
                          int w, int h, int levels, int precision, int lossy, int output_dwt,
                          int cbw, int cbh, double delta);
/* The same path as separate stages, for the stage classes:
 *  load_planar  = OCLMemoryManager<T>::init + hostToDWTIn
 *                 (src/OCLMemoryManager.cpp:56-160): host components -> HBM;
 *  dwt          = OCLDWTForward<T>::run over all levels
 *                 (src/OCLDWTForward.cpp:124-137), quantised unless output_dwt;
 *  planar_i16   = OCLRGBtoPlanar<T>::run (src/OCLRGBtoPlanar.cpp:36-44): the
 *                 per-channel int16 planes (OCLMemoryManager::getDWTOutByChannel,
 *                 src/OCLMemoryManager.h:50-54), saturated from the int32 planes;
 *  bpc          = OCLBPC<T>::run(cbx, cby) (src/OCLBPC.cpp:38-51). */
int rg_encoder_load_planar(rg_encoder *enc, const void *const *components, int ncomp, int sample_type,
                           int w, int h, int precision, int lossy);
int rg_encoder_dwt(rg_encoder *enc, int levels, int output_dwt, double delta);
int rg_encoder_planar_i16(rg_encoder *enc);
int rg_encoder_channel_i16(rg_encoder *enc, int channel, int16_t **d_plane);
int rg_encoder_bpc(rg_encoder *enc, int cbw, int cbh);
/* Parallel-style tier-1 of this encoder in `phases` modelling phases: 1 runs
 * every kernel on the context's stream; 2-4 overlap the early passes' chain
 * coding with the later passes' modelling on 1-3 side streams of the context
 * (the default, 0, is three phases).  Callers that run several contexts at
 * once (a serving loop) use 1 to stay within the GPU's hardware queues. */
int rg_encoder_set_t1_phases(rg_encoder *enc, int phases);
/* Device pointers of the last run (valid until the next run / destroy). */
int rg_encoder_coefficients(rg_encoder *enc, int32_t **d_q, int *ncomp, int *w, int *h);
int rg_encoder_float_coefficients(rg_encoder *enc, float **d_coeff);
/* Input of DWT level `level` (OCLMemoryManager::getDwtIn,
 * src/OCLMemoryManager.h:55-59): level 0 is the transform's input planes
 * (float for 9/7, int32 for 5/3), level k in 1..levels-1 the LL band after k
 * levels (float, or FP16 when the encode stored inter-level LL as FP16;
 * int32 for 5/3), ncomp contiguous planes of *w x *h samples each.  Levels
 * beyond the transform return RG_ERR_INVALID and a null pointer. */
int rg_encoder_dwt_in(rg_encoder *enc, int level, void **d_ptr, int *w, int *h);
/* Number of code-blocks of the last run (all components). */
int rg_encoder_num_codeblocks(rg_encoder *enc);
/* Copies per-block results (ncb) and codewords to host; data gets the
 * concatenated codewords (cap bytes), offsets[i] their start.  Waits. */
int rg_encoder_get_codeblocks(rg_encoder *enc, rg_cb_desc *desc, rg_cb_result *res, long *offsets,
                              uint8_t *data, long cap, long *total);
/* TERMALL styles: the codeword segment length of every coding pass,
 * RG_MAX_PASSES entries per code-block (ncb * RG_MAX_PASSES, host).  Waits.
 * No counterpart in the reference (its bit-plane coder has no output). */
int rg_encoder_get_pass_lengths(rg_encoder *enc, int32_t *seglen);
/* Rate control (PCRD-opt).  No counterpart in the reference (its bit-plane
 * coder writes no bytes, src/CL/oclbpc.cl:279-746); T.800 leaves truncation
 * to the encoder.  enable_rate_control(enc, 1) before tier-1 runs makes the
 * context modeller record which samples each significance-propagation pass
 * made significant (512 bytes per bit-plane and code-block).  Then, for a
 * parallel style, rate_control cuts every code-block after the coding pass
 * that the slope threshold of the smallest total distortion within
 * `target_bytes` of code-block data selects (the codestream adds its
 * headers); weights: one positive factor per code-block scaling its
 * distortion (NULL: step^2 times the band's synthesis energy).  Results,
 * codewords and write_j2k reflect the cut; calling it again with another
 * target recuts from the full codewords.  *kept = code-block bytes kept.
 * Needs coefficient magnitudes below 2^24.  Waits. */
int rg_encoder_enable_rate_control(rg_encoder *enc, int on);
int rg_encoder_rate_control(rg_encoder *enc, long target_bytes, const double *weights, long *kept);
/* Quality layers: targets[0..nlayers) non-decreasing budgets of code-block
 * data; layer l holds every block's passes up to layer l's threshold beyond
 * those of layer l-1, and write_j2k writes nlayers layers (a decoder that
 * stops after layer l gets the cut at targets[l]).  *kept: the last layer's
 * bytes.  A new encode goes back to one layer. */
int rg_encoder_rate_control_layers(rg_encoder *enc, int nlayers, const long *targets, const double *weights,
                                   long *kept);
/* Tier-2: the JPEG 2000 codestream (.j2k: one tile, one layer, LRCP) of the
 * last encode.  out == NULL only reports the size.  The reference has no
 * tier-2 (it stops after its bit-plane coder, src/OCLEncoder.cpp:43-52). */
int rg_encoder_write_j2k(rg_encoder *enc, uint8_t *out, long cap, long *size);
/* mapDWTOut / unmapDWTOut (src/OCLEncodeDecode.cpp:42-48): host view of the
 * coefficients, interleaved by component — int16 saturated quantisation
 * indices / 5/3 coefficients, or float when output_dwt on the lossy path. */
int rg_encoder_map_dwt_out(rg_encoder *enc, void **host_ptr);
int rg_encoder_unmap_dwt_out(rg_encoder *enc, void *host_ptr);

/* ---- decoder: inverse DWT of planar coefficients (OCLDecoder,
 *      src/OCLDecoder.h:25-34) ---- */
int rg_decoder_create(rg_context *ctx, rg_decoder **out);
int rg_decoder_destroy(rg_decoder *dec);
/* components: ncomp Mallat-packed coefficient planes (host) of sample type
 * RG_F32 (lossy: 9/7 float coefficients) or RG_I16/RG_I32 (5/3). */
int rg_decoder_run_planar(rg_decoder *dec, const void *const *components, int ncomp, int sample_type,
                          int w, int h, int levels, int precision, int lossy);
/* Separate stages: load (OCLMemoryManager<T>::init) and the inverse DWT
 * (OCLDWTRev<T>::run, src/OCLDWTRev.cpp:39-59). */
int rg_decoder_load_planar(rg_decoder *dec, const void *const *components, int ncomp, int sample_type,
                           int w, int h, int precision, int lossy);
int rg_decoder_idwt(rg_decoder *dec, int levels);
/* Host view of the reconstructed samples, interleaved by component (float
 * for lossy, int16 for lossless), still level shifted. */
int rg_decoder_map_out(rg_decoder *dec, void **host_ptr);
int rg_decoder_unmap_out(rg_decoder *dec, void *host_ptr);

/* EBCOT tier-1 decode of ncb code-blocks (device memory): codeword i at
 * d_data + d_offsets[i] with results d_res[i] (len, npasses, nbp); signed
 * coefficients written into the Mallat planes like rg_t1_encode reads them.
 * No counterpart in the reference (its bit-plane coder has no output and
 * its decoder never decodes bits, src/OCLDecoder.cpp:35-38). */
int rg_t1_decode(rg_context *ctx, const uint8_t *d_data, const long *d_offsets, const rg_cb_result *d_res,
                 const rg_cb_desc *d_desc, int ncb, int32_t *d_q, int pitch, size_t plane_stride);
/* rg_t1_decode for a code-block style (0, RG_STYLE_PARALLEL or
 * RG_STYLE_PARALLEL_BYPASS); with TERMALL, d_seglen (device) holds each
 * block's pass segment lengths, RG_MAX_PASSES per block, as
 * rg_encoder_get_pass_lengths returns them.  No counterpart in the
 * reference (see rg_t1_decode). */
int rg_t1_decode_style(rg_context *ctx, const uint8_t *d_data, const long *d_offsets, const rg_cb_result *d_res,
                       const rg_cb_desc *d_desc, int ncb, int32_t *d_q, int pitch, size_t plane_stride, int style,
                       const int32_t *d_seglen);
/* Full native decode of one image from its code-blocks (host arrays, the
 * layout rg_encoder_get_codeblocks returns): tier-1 decode -> dequantisation
 * + inverse DWT -> inverse RCT/ICT -> interleaved uint8 on the device.
 * p->style selects the code-block style; with TERMALL, seglen gives each
 * block's pass segment lengths (RG_MAX_PASSES per block), else NULL. */
int rg_decoder_decode_codeblocks(rg_decoder *dec, const rg_cb_desc *desc, const rg_cb_result *res,
                                 const long *offsets, const uint8_t *data, long total, int ncb, int w, int h,
                                 int ncomp, const rg_encode_params *p, const int32_t *seglen);
/* Tier-2 reader (host only, no device work).  The reference has no tier-2
 * (src/OCLEncoder.cpp:43-52).  Reads a .j2k codestream: one tile in one
 * tile-part, LRCP, any number of quality layers, maximal precincts, unsigned
 * 8-bit samples without subsampling, 1-4 components, code-block styles 0,
 * RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS, optional SOP/EPH markers.
 * Call with only `info` to learn ncb and data_bytes, then again with desc /
 * res / offsets (ncb each), data (data_bytes) and, for TERMALL styles,
 * seglen (ncb * RG_MAX_PASSES); blocks come in the encoder's order and the
 * result feeds rg_decoder_decode_codeblocks. */
typedef struct {
    int32_t w, h, ncomp, levels, cbw, cbh;
    int32_t style;        /* code-block style byte of COD */
    int32_t filter;       /* RG_FILTER_53 / RG_FILTER_97 */
    int32_t colour;       /* RG_COLOUR_* from the COD component transform */
    int32_t nlayers, bitdepth, guard_bits;
    int32_t ncb;          /* code-blocks, all components */
    int64_t data_bytes;   /* codeword bytes, all blocks */
    float step[97];       /* quantisation step per band (3*levels+1; 1 for 5/3) */
} rg_j2k_info;
int rg_j2k_parse(const uint8_t *cs, long size, rg_j2k_info *info, rg_cb_desc *desc, rg_cb_result *res,
                 long *offsets, uint8_t *data, int32_t *seglen);
/* Whole decode of a codestream (host bytes) into the decoder's uint8 image:
 * rg_j2k_parse, then tier-1 decode, dequantisation with the codestream's
 * steps, inverse DWT and colour transform on the device. */
int rg_decoder_decode_j2k(rg_decoder *dec, const uint8_t *cs, long size, rg_j2k_info *info);
int rg_decoder_image_u8(rg_decoder *dec, uint8_t **d_img);
int rg_decoder_coefficients(rg_decoder *dec, int32_t **d_q);

/* ---- batched encode of many images (no counterpart in the reference, which
 *      encodes one image per OCLEncoder::run call) ---- */
typedef struct rg_batch rg_batch;
typedef struct {
    const uint8_t *d_src;  /* interleaved uint8 image in device memory, 16-byte aligned */
    int32_t w, h, ncomp;   /* ncomp 1 or 3 */
    int32_t filter, colour, levels;  /* levels 1..15 */
    double delta;          /* 9/7 base quantisation step */
} rg_batch_image;
int rg_batch_create(rg_context *ctx, rg_batch **out);
int rg_batch_destroy(rg_batch *b);
/* Lays out all buffers and job tables (tier-1 jobs sorted largest block
 * first); with use_graph the pipeline (colour, every DWT level, tier-1) is
 * captured once into a HIP graph.  style: code-block style of every image
 * (0, RG_STYLE_PARALLEL or RG_STYLE_PARALLEL_BYPASS).  waves_per_cu > 0: tier-1 context modelling runs
 * as a persistent grid of that many wavefronts per CU pulling code-blocks
 * from a device work queue; 0: one workgroup per code-block. */
int rg_batch_plan(rg_batch *b, const rg_batch_image *imgs, int n, int cbw, int cbh, int use_graph, int style,
                  int waves_per_cu);
/* Runs (replays) the plan asynchronously on the context stream. */
int rg_batch_run(rg_batch *b);
/* Code-blocks of one image (image = -1: the whole batch). */
int rg_batch_num_codeblocks(rg_batch *b, int image);
int rg_batch_coefficients(rg_batch *b, int image, int32_t **d_q);
int rg_batch_get_codeblocks(rg_batch *b, int image, rg_cb_desc *desc, rg_cb_result *res, long *offsets,
                            uint8_t *data, long cap, long *total);
/* TERMALL style: segment lengths of image `image`'s blocks (RG_MAX_PASSES
 * each, host).  Waits. */
int rg_batch_get_pass_lengths(rg_batch *b, int image, int32_t *seglen);
int rg_batch_write_j2k(rg_batch *b, int image, uint8_t *out, long cap, long *size);

/* ---- synthetic test images generated on the device (the seeded recipe of
 *      roger_amd/synth.py: params are 16 doubles per channel = 4 x (cos theta,
 *      sin theta, period, phase); noise_base one counter base per channel;
 *      walk_x[w] / walk_y[h] the luminance walks or both NULL) ---- */
int rg_synth_u8(rg_context *ctx, uint8_t *d_dst, int w, int h, int ncomp, const double *params,
                const uint64_t *noise_base, const double *walk_x, const double *walk_y);

#ifdef __cplusplus
}
#endif
#endif
