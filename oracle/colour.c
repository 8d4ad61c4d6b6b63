/*
 * Oracle: DC level shift and component transforms (T.800 Annex G).
 * TEST INFRASTRUCTURE ONLY (see j2k_oracle.h).
 *
 * The reference has no colour transform at all: its harness subtracts 128 on
 * the host (src/OCLTest.cpp:87) and src/CL/oclplanar.cl:23-39 only splits an
 * RGBA coefficient image into planes.  The forms below follow T.800 G.1.2
 * (level shift), G.2 (RCT) and G.3 (ICT).
 *
 * FP32 op order (fmaf chains, left to right) is the contract shared with the
 * HIP colour kernel so that both produce identical bits.
 */
#include "j2k_oracle.h"
#include <math.h>

void orc_shift_i32(const uint8_t *src, int npix, int ncomp, int c, int32_t *dst)
{
    for (int i = 0; i < npix; ++i) dst[i] = (int32_t)src[(size_t)i * ncomp + c] - 128;
}

void orc_shift_f32(const uint8_t *src, int npix, int ncomp, int c, float *dst)
{
    for (int i = 0; i < npix; ++i) dst[i] = (float)((int)src[(size_t)i * ncomp + c] - 128);
}

/* G.2: Y = floor((R + 2G + B) / 4), U = B - G, V = R - G (after level shift). */
void orc_rct_fwd(const uint8_t *rgb, int npix, int32_t *y, int32_t *u, int32_t *v)
{
    for (int i = 0; i < npix; ++i) {
        int r = (int)rgb[3 * (size_t)i] - 128, g = (int)rgb[3 * (size_t)i + 1] - 128,
            b = (int)rgb[3 * (size_t)i + 2] - 128;
        y[i] = (r + 2 * g + b) >> 2;
        u[i] = b - g;
        v[i] = r - g;
    }
}

static inline uint8_t clamp_u8(int x) { return (uint8_t)(x < 0 ? 0 : (x > 255 ? 255 : x)); }

void orc_rct_inv(const int32_t *y, const int32_t *u, const int32_t *v, int npix, uint8_t *rgb)
{
    for (int i = 0; i < npix; ++i) {
        int g = y[i] - ((u[i] + v[i]) >> 2);
        int r = v[i] + g, b = u[i] + g;
        rgb[3 * (size_t)i] = clamp_u8(r + 128);
        rgb[3 * (size_t)i + 1] = clamp_u8(g + 128);
        rgb[3 * (size_t)i + 2] = clamp_u8(b + 128);
    }
}

/* G.3 coefficients. */
#define ICT_YR 0.299
#define ICT_YG 0.587
#define ICT_YB 0.114
#define ICT_CBR (-0.16875)
#define ICT_CBG (-0.33126)
#define ICT_CBB 0.5
#define ICT_CRR 0.5
#define ICT_CRG (-0.41869)
#define ICT_CRB (-0.08131)
#define ICT_RCR 1.402
#define ICT_GCB (-0.34413)
#define ICT_GCR (-0.71414)
#define ICT_BCB 1.772

void orc_ict_fwd_f32(const uint8_t *rgb, int npix, float *y, float *cb, float *cr)
{
    for (int i = 0; i < npix; ++i) {
        float r = (float)((int)rgb[3 * (size_t)i] - 128), g = (float)((int)rgb[3 * (size_t)i + 1] - 128),
              b = (float)((int)rgb[3 * (size_t)i + 2] - 128);
        y[i] = fmaf((float)ICT_YB, b, fmaf((float)ICT_YG, g, (float)ICT_YR * r));
        cb[i] = fmaf((float)ICT_CBB, b, fmaf((float)ICT_CBG, g, (float)ICT_CBR * r));
        cr[i] = fmaf((float)ICT_CRB, b, fmaf((float)ICT_CRG, g, (float)ICT_CRR * r));
    }
}

/* Round half away from zero, then clamp: the decoder-side sample output. */
static inline uint8_t out_u8f(float x)
{
    float s = x + 128.0f;
    int v = (int)floorf(s + 0.5f);
    return clamp_u8(v);
}

void orc_ict_inv_f32(const float *y, const float *cb, const float *cr, int npix, uint8_t *rgb)
{
    for (int i = 0; i < npix; ++i) {
        float r = fmaf((float)ICT_RCR, cr[i], y[i]);
        float g = fmaf((float)ICT_GCR, cr[i], fmaf((float)ICT_GCB, cb[i], y[i]));
        float b = fmaf((float)ICT_BCB, cb[i], y[i]);
        rgb[3 * (size_t)i] = out_u8f(r);
        rgb[3 * (size_t)i + 1] = out_u8f(g);
        rgb[3 * (size_t)i + 2] = out_u8f(b);
    }
}

void orc_ict_fwd_f64(const uint8_t *rgb, int npix, double *y, double *cb, double *cr)
{
    for (int i = 0; i < npix; ++i) {
        double r = (int)rgb[3 * (size_t)i] - 128, g = (int)rgb[3 * (size_t)i + 1] - 128,
               b = (int)rgb[3 * (size_t)i + 2] - 128;
        y[i] = ICT_YR * r + ICT_YG * g + ICT_YB * b;
        cb[i] = ICT_CBR * r + ICT_CBG * g + ICT_CBB * b;
        cr[i] = ICT_CRR * r + ICT_CRG * g + ICT_CRB * b;
    }
}

void orc_ict_inv_f64(const double *y, const double *cb, const double *cr, int npix, double *rgb_f)
{
    for (int i = 0; i < npix; ++i) {
        rgb_f[3 * (size_t)i] = y[i] + ICT_RCR * cr[i] + 128.0;
        rgb_f[3 * (size_t)i + 1] = y[i] + ICT_GCB * cb[i] + ICT_GCR * cr[i] + 128.0;
        rgb_f[3 * (size_t)i + 2] = y[i] + ICT_BCB * cb[i] + 128.0;
    }
}

void orc_unshift_i32(const int32_t *src, int npix, int ncomp, int c, uint8_t *dst)
{
    for (int i = 0; i < npix; ++i) dst[(size_t)i * ncomp + c] = clamp_u8(src[i] + 128);
}

void orc_unshift_f32(const float *src, int npix, int ncomp, int c, uint8_t *dst)
{
    for (int i = 0; i < npix; ++i) dst[(size_t)i * ncomp + c] = out_u8f(src[i]);
}
