/*
 * CPU oracle for the JPEG 2000 encode/decode hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a clean-room restatement of ITU-T T.800 (JPEG 2000 part 1) for the
 * stages on the path the reference (knopkem/roger) implements as OpenCL
 * kernels: level shift and RCT/ICT (Annex G), the 5/3 and 9/7 lifting DWT
 * (Annex F), scalar quantisation (Annex E), EBCOT tier-1 coding with the MQ
 * coder (Annexes D and C), plus a minimal tier-2 codestream writer (Annex B)
 * used only so that an independent decoder (OpenJPEG, through Pillow) can
 * check the output.
 *
 * Only the test-suite, the smoke check and the benchmark's CPU-baseline leg
 * may load this library; the product path never calls it.
 *
 * Pinning (what each function is checked against, in tests/):
 *   colour    : closed forms (T.800 G.2/G.3), RCT exact inverse over all 2^24
 *               RGB triples, ICT against an FP64 matrix product.
 *   dwt53     : brute-force 1-D/2-D restatement for all sizes 1..33, exact
 *               inverse, and lossless decode by OpenJPEG of the codestream.
 *   dwt97     : FP64 lifting == FP64 convolution with the T.800 Table F.4
 *               derived 9/7 filter taps, perfect reconstruction, decode by
 *               OpenJPEG (PSNR agreement).
 *   quant     : closed form sign*floor(|y|/delta); round trip via OpenJPEG.
 *   t1 (EBCOT): own decoder inverts own encoder, and OpenJPEG decodes the
 *               codestreams bit-exactly (lossless) / within PSNR (lossy).
 *   t2 writer : OpenJPEG decodes it.
 */
#ifndef J2K_ORACLE_H
#define J2K_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- colour (T.800 Annex G) ---------------- */
/* Interleaved uint8 (ncomp channels) -> level-shifted planes. */
void orc_shift_i32(const uint8_t *src, int npix, int ncomp, int c, int32_t *dst);
void orc_shift_f32(const uint8_t *src, int npix, int ncomp, int c, float *dst);
void orc_rct_fwd(const uint8_t *rgb, int npix, int32_t *y, int32_t *u, int32_t *v);
void orc_rct_inv(const int32_t *y, const int32_t *u, const int32_t *v, int npix, uint8_t *rgb);
void orc_ict_fwd_f32(const uint8_t *rgb, int npix, float *y, float *cb, float *cr);
void orc_ict_inv_f32(const float *y, const float *cb, const float *cr, int npix, uint8_t *rgb);
void orc_ict_fwd_f64(const uint8_t *rgb, int npix, double *y, double *cb, double *cr);
void orc_ict_inv_f64(const double *y, const double *cb, const double *cr, int npix, double *rgb_f);
/* Planar inverse level shift with clamping to 0..255 into interleaved output. */
void orc_unshift_i32(const int32_t *src, int npix, int ncomp, int c, uint8_t *dst);
void orc_unshift_f32(const float *src, int npix, int ncomp, int c, uint8_t *dst);

/* ---------------- DWT (T.800 Annex F), in place, Mallat layout ---------------- */
/* Plane is w x h with row stride `stride` (elements).  One level transforms the
 * top-left ceil(w/2^l) x ceil(h/2^l) region: columns (VER_SD) then rows (HOR_SD). */
void orc_dwt53_fwd(int32_t *p, int w, int h, int stride, int levels);
void orc_dwt53_inv(int32_t *p, int w, int h, int stride, int levels);
/* ll_fp16 != 0: the LL band handed to the next level is rounded to FP16. */
void orc_dwt97_fwd_f32(float *p, int w, int h, int stride, int levels, int ll_fp16);
void orc_dwt97_inv_f32(float *p, int w, int h, int stride, int levels);
void orc_dwt97_fwd_f64(double *p, int w, int h, int stride, int levels);
void orc_dwt97_inv_f64(double *p, int w, int h, int stride, int levels);
/* 1-D single-level transforms (interleaved in, deinterleaved low|high out). */
void orc_dwt53_fwd_1d(const int32_t *x, int n, int32_t *out);
void orc_dwt97_fwd_1d_f64(const double *x, int n, double *out);

/* ---------------- subband geometry ---------------- */
typedef struct {
    int x0, y0, w, h; /* rectangle in the Mallat plane */
    int orient;       /* 0 LL, 1 HL, 2 LH, 3 HH */
    int level;        /* decomposition level (LL: = levels) */
    int res;          /* resolution index (0 = LL) */
    int band;         /* band index: 0 = LL_N, then HL/LH/HH for l = N..1 */
} orc_band;
/* Fills bands[0 .. 3*levels] in band-index order; returns the count. */
int orc_bands(int w, int h, int levels, orc_band *bands);

/* ---------------- quantisation (T.800 Annex E) ---------------- */
/* Step size for a band from its exponent/mantissa and nominal range R_b. */
double orc_step_from_expmant(int Rb, int eps, int mu);
/* Picks (eps, mu) so that the signalled step is the closest to `delta`. */
void orc_expmant_for_step(int Rb, double delta, int *eps, int *mu);
/* q = sign(y) * floor(|y| * inv_step[band]) over the Mallat plane. */
void orc_quantize_f32(const float *p, int w, int h, int stride, int levels,
                      const float *inv_step, int32_t *q);
/* y = sign(q) * (|q| + 0.5) * step[band] for q != 0, 0 otherwise. */
void orc_dequantize_f32(const int32_t *q, int w, int h, int stride, int levels,
                        const float *step, float *y);

/* ---------------- EBCOT tier-1 (Annexes C, D) ---------------- */
typedef struct {
    int32_t len;     /* bytes of the terminated MQ codeword (all passes) */
    int32_t npasses; /* coding passes: 3*nbp-2, 0 if the block is all zero */
    int32_t nbp;     /* magnitude bit-planes actually coded (msb+1) */
    int32_t flags;   /* bit0: output did not fit the slot */
} orc_cb_result;

/* Code-block style bits (T.800 Table A.19) the oracle implements. */
#define ORC_STYLE_BYPASS  0x01  /* raw SPP/MRP passes from the fifth bit-plane (with TERMALL) */
#define ORC_STYLE_RESET   0x02  /* contexts reset after every coding pass */
#define ORC_STYLE_TERMALL 0x04  /* codeword terminated after every coding pass */
#define ORC_MAX_PASSES    96    /* >= 3*32-2: stride of per-pass length arrays */

/* Encodes one code-block of w x h (<= 64 x 64, w*h <= 4096) signed
 * coefficients (q, row stride `stride`) of orientation `orient`.  Writes at
 * most `cap` bytes to out. */
void orc_t1_encode(const int32_t *q, int w, int h, int stride, int orient,
                   uint8_t *out, int cap, orc_cb_result *res);
/* Same with a code-block style; with TERMALL, passlen[i] receives the length
 * of pass i's codeword segment (the codeword is their concatenation). */
void orc_t1_encode_style(const int32_t *q, int w, int h, int stride, int orient, int style,
                         uint8_t *out, int cap, orc_cb_result *res, int32_t *passlen);
/* orc_t1_encode_style that also writes, per coding pass, the distortion
 * reduction of the pass (dist[i]; units of orc_t1_dsig / orc_t1_dref). */
void orc_t1_encode_rd(const int32_t *q, int w, int h, int stride, int orient, int style,
                      uint8_t *out, int cap, orc_cb_result *res, int32_t *passlen, int64_t *dist);
/* Squared-error reductions, in doubled index units (integers), of a magnitude
 * m becoming significant at plane p / of refining its bit p; the decoder
 * reconstructs the known bits plus half the remaining interval. */
int64_t orc_t1_dsig(uint32_t m, int p);
int64_t orc_t1_dref(uint32_t m, int p);
/* Rate control (PCRD-opt): the upper convex hull of one block's (rate,
 * weighted distortion reduction) truncation points; hn/hr/hs get the pass
 * count, cumulative bytes and slope of each hull point (npasses + 1 entries
 * at most).  Returns the number of hull points. */
int orc_rd_hull(int npasses, const int32_t *passlen, const int64_t *dist, double w, int32_t *hn, int64_t *hr,
                double *hs);
/* Truncation of nblocks code-blocks (per-pass arrays with row stride `stride`)
 * at one slope threshold: the smallest whose total rate is <= target.
 * ntrunc / rtrunc: passes and bytes kept per block.  Returns the total. */
long orc_rate_control(int nblocks, const int32_t *npasses, const int32_t *passlen, const int64_t *dist, int stride,
                      const double *weight, long target, int32_t *ntrunc, int32_t *rtrunc);
/* Decodes all `npasses` passes of a code-block back into q (row stride). */
int orc_t1_decode(const uint8_t *data, int len, int w, int h, int orient,
                  int nbp, int npasses, int32_t *q, int stride);
/* Same with a code-block style (passlen: segment lengths when TERMALL). */
int orc_t1_decode_style(const uint8_t *data, int len, int w, int h, int orient, int nbp, int npasses,
                        int style, const int32_t *passlen, int32_t *q, int stride);
/* ZC context lookup used by both coders (exported for table tests). */
int orc_zc_context(int orient, int h, int v, int d);
/* Sign context (9..13) and xor bit for H, V contributions in -1..1. */
int orc_sc_context(int H, int V, int *xorbit);

/* Code-block enumeration for one component plane: resolution 0..N, bands in
 * resolution order (LL | HL, LH, HH), raster order inside each band. */
typedef struct {
    int x0, y0, w, h; /* rectangle in the Mallat plane */
    int orient, band, res;
    int cbx, cby;     /* code-block grid position inside its band */
} orc_cb_desc;
int orc_codeblocks(int w, int h, int levels, int cbw, int cbh, orc_cb_desc *out, int cap);

/* ---------------- tier-2 (Annex B), single tile / layer / precinct ---------------- */
typedef struct {
    int w, h, ncomp, levels;
    int cbw_log2, cbh_log2;
    int reversible;  /* 1: 5/3 + RCT, 0: 9/7 + ICT */
    int mct;         /* component transform flag (ncomp == 3) */
    int guard_bits;
    int bitdepth;
    int eps[97], mu[97]; /* per band */
    int cblk_style;      /* code-block style byte of COD (ORC_STYLE_*) */
} orc_j2k_params;
/* Nominal dynamic range R_b of a band (bitdepth + log2 gain). */
int orc_band_range(int bitdepth, int orient);
/* Number of magnitude bit-planes M_b = G + eps_b - 1. */
int orc_band_mb(const orc_j2k_params *p, int band);
/* cbres/cbdata are per component in orc_codeblocks order; cbdata holds each
 * block's bytes at cboff[i]; with TERMALL, passlen holds each block's segment
 * lengths (ORC_MAX_PASSES per block).  Returns the codestream size or -1. */
long orc_t2_write(const orc_j2k_params *p, const orc_cb_result *cbres,
                  const long *cboff, const uint8_t *cbdata, const int32_t *passlen,
                  int ncb_per_comp, uint8_t *out, long cap);
long orc_t2_write_layers(const orc_j2k_params *p, const orc_cb_result *cbres,
                         const long *cboff, const uint8_t *cbdata, const int32_t *passlen,
                         int ncb_per_comp, int nlayers, const int32_t *layer_np, uint8_t *out, long cap);

/* ---------------- synthetic inputs ---------------- */
uint64_t orc_splitmix64(uint64_t *state);

/* ---------------- whole pipelines (CPU baseline) ---------------- */
typedef struct {
    int levels, cbw, cbh, reversible, ncomp;
    double delta;  /* 9/7 base quantisation step */
    int nthreads;  /* OpenMP threads (<= 0: library default) */
    int style;     /* code-block style (ORC_STYLE_*) */
} orc_encode_cfg;
/* Encodes an interleaved uint8 image; writes per-code-block results and
 * bytes (slot_bytes each, component-major).  Returns #code-blocks or <0. */
int orc_encode_image(const uint8_t *img, int w, int h, const orc_encode_cfg *cfg,
                     orc_cb_result *res, uint8_t *slots, int slot_bytes, int ncb_cap,
                     int32_t *qplanes_out, int32_t *passlen_out);

#ifdef __cplusplus
}
#endif
#endif
