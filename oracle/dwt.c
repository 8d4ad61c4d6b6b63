/*
 * Oracle: reversible 5/3 and irreversible 9/7 lifting DWT (T.800 Annex F).
 * TEST INFRASTRUCTURE ONLY (see j2k_oracle.h).
 *
 * Reference counterparts: src/CL/ocldwt53.cl:201-278 (5/3 lifting, predict
 * -(a+b)>>1, update +(a+b+2)>>2), src/CL/ocldwt97.cl:69-78 (9/7 constants),
 * src/OCLDWTForward.cpp:124-137 (level recursion on ceil(w/2) x ceil(h/2)),
 * src/CL/ocldwt53.cl:110-157,171-175 (Mallat-packed output).  The reference
 * kernels run rows before columns and apply the 9/7 scaling twice; this
 * restatement follows T.800: VER_SD (columns) then HOR_SD (rows) per level
 * (F.4.2), whole-sample symmetric extension (F.4.7 / 1D_EXTR), a length-1
 * signal passes through unchanged (F.4.6), low band scaled by 1/K and high
 * band by K.
 *
 * FP32 contract shared with the HIP kernels: every lifting step is
 * x[i] = fmaf(c, x[i-1] + x[i+1], x[i]); scaling multiplies by the float
 * constants below; inverse steps use fmaf(-c, ...).
 */
#include "j2k_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A97 (-1.586134342059924)
#define B97 (-0.052980118572961)
#define G97 (0.882911075530934)
#define D97 (0.443506852043971)
#define K97 (1.230174104914001)

static const float A97F = (float)A97, B97F = (float)B97, G97F = (float)G97, D97F = (float)D97;
static const float K97F = (float)K97;
static const float IK97F = (float)(1.0 / K97);

/* Whole-sample symmetric periodic extension index for n >= 2. */
static inline int mirror(int i, int n)
{
    int p = 2 * (n - 1);
    i %= p;
    if (i < 0) i += p;
    return i >= n ? p - i : i;
}

#define EXT 4

/* ---- 5/3 ---- */
static void fwd53_1d(const int32_t *x, int sx, int n, int32_t *lo, int32_t *hi, int s_out, int32_t *buf)
{
    int32_t *e = buf + EXT;
    if (n == 1) { lo[0] = x[0]; return; }
    for (int i = -2; i < n + 2; ++i) e[i] = x[(size_t)mirror(i, n) * sx];
    for (int i = -1; i <= n; ++i)
        if (i & 1) e[i] -= (e[i - 1] + e[i + 1]) >> 1;
    for (int i = 0; i < n; i += 2) e[i] += (e[i - 1] + e[i + 1] + 2) >> 2;
    int nl = (n + 1) / 2;
    for (int k = 0; k < nl; ++k) lo[(size_t)k * s_out] = e[2 * k];
    for (int k = 0; 2 * k + 1 < n; ++k) hi[(size_t)k * s_out] = e[2 * k + 1];
}

static void inv53_1d(const int32_t *lo, const int32_t *hi, int s_in, int n, int32_t *x, int sx, int32_t *buf)
{
    int32_t *e = buf + EXT;
    if (n == 1) { x[0] = lo[0]; return; }
    int32_t *t = buf + EXT + n + EXT + 8; /* interleaved copy */
    for (int i = 0; i < n; ++i) t[i] = (i & 1) ? hi[(size_t)(i >> 1) * s_in] : lo[(size_t)(i >> 1) * s_in];
    for (int i = -2; i < n + 2; ++i) e[i] = t[mirror(i, n)];
    for (int i = -1; i <= n; ++i)
        if (!(i & 1)) e[i] -= (e[i - 1] + e[i + 1] + 2) >> 2;
    for (int i = 1; i < n; i += 2) e[i] += (e[i - 1] + e[i + 1]) >> 1;
    for (int i = 0; i < n; ++i) x[(size_t)i * sx] = e[i];
}

void orc_dwt53_fwd_1d(const int32_t *x, int n, int32_t *out)
{
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (3 * n + 64));
    fwd53_1d(x, 1, n, out, out + (n + 1) / 2, 1, buf);
    free(buf);
}

/* One 2-D level on the top-left cw x ch region: columns then rows. */
void orc_dwt53_fwd(int32_t *p, int w, int h, int stride, int levels)
{
    int maxn = (w > h ? w : h);
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (3 * maxn + 64));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (maxn + 8));
    int cw = w, ch = h;
    for (int l = 0; l < levels; ++l) {
        int hl = (ch + 1) / 2, wl = (cw + 1) / 2;
        for (int x = 0; x < cw; ++x) {
            for (int y = 0; y < ch; ++y) tmp[y] = p[(size_t)y * stride + x];
            fwd53_1d(tmp, 1, ch, p + x, p + (size_t)hl * stride + x, stride, buf);
        }
        for (int y = 0; y < ch; ++y) {
            int32_t *row = p + (size_t)y * stride;
            memcpy(tmp, row, sizeof(int32_t) * cw);
            fwd53_1d(tmp, 1, cw, row, row + wl, 1, buf);
        }
        cw = wl; ch = hl;
    }
    free(buf); free(tmp);
}

void orc_dwt53_inv(int32_t *p, int w, int h, int stride, int levels)
{
    int maxn = (w > h ? w : h);
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (3 * maxn + 64));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (maxn + 8));
    for (int l = levels; l >= 1; --l) {
        /* region size at level l-1 */
        int cw = w, ch = h;
        for (int k = 0; k < l - 1; ++k) { cw = (cw + 1) / 2; ch = (ch + 1) / 2; }
        int hl = (ch + 1) / 2, wl = (cw + 1) / 2;
        for (int y = 0; y < ch; ++y) {
            int32_t *row = p + (size_t)y * stride;
            memcpy(tmp, row, sizeof(int32_t) * cw);
            inv53_1d(tmp, tmp + wl, 1, cw, row, 1, buf);
        }
        for (int x = 0; x < cw; ++x) {
            for (int y = 0; y < ch; ++y) tmp[y] = p[(size_t)y * stride + x];
            inv53_1d(tmp, tmp + hl, 1, ch, p + x, stride, buf);
        }
    }
    free(buf); free(tmp);
}

/* ---- 9/7 FP32 ---- */
static void fwd97_1d_f32(const float *x, int n, float *lo, float *hi, int s_out, float *buf)
{
    float *e = buf + EXT;
    if (n == 1) { lo[0] = x[0]; return; }
    for (int i = -4; i < n + 4; ++i) e[i] = x[mirror(i, n)];
    for (int i = -3; i <= n + 2; ++i) if (i & 1) e[i] = fmaf(A97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = -2; i <= n + 1; ++i) if (!(i & 1)) e[i] = fmaf(B97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = -1; i <= n; ++i) if (i & 1) e[i] = fmaf(G97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = 0; i < n; i += 2) e[i] = fmaf(D97F, e[i - 1] + e[i + 1], e[i]);
    int nl = (n + 1) / 2;
    for (int k = 0; k < nl; ++k) lo[(size_t)k * s_out] = e[2 * k] * IK97F;
    for (int k = 0; 2 * k + 1 < n; ++k) hi[(size_t)k * s_out] = e[2 * k + 1] * K97F;
}

static void inv97_1d_f32(const float *lo, const float *hi, int n, float *x, int sx, float *buf)
{
    float *e = buf + EXT;
    if (n == 1) { x[0] = lo[0]; return; }
    float *t = buf + EXT + n + EXT + 8;
    for (int i = 0; i < n; ++i) t[i] = (i & 1) ? hi[i >> 1] * IK97F : lo[i >> 1] * K97F;
    for (int i = -4; i < n + 4; ++i) e[i] = t[mirror(i, n)];
    for (int i = -3; i <= n + 2; ++i) if (!(i & 1)) e[i] = fmaf(-D97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = -2; i <= n + 1; ++i) if (i & 1) e[i] = fmaf(-G97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = -1; i <= n; ++i) if (!(i & 1)) e[i] = fmaf(-B97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = 1; i < n; i += 2) e[i] = fmaf(-A97F, e[i - 1] + e[i + 1], e[i]);
    for (int i = 0; i < n; ++i) x[(size_t)i * sx] = e[i];
}

/* Round an FP32 value to the nearest FP16 value (ties to even) and back. */
float orc_round_f16(float f);
float orc_round_f16(float f)
{
    union { float f; uint32_t u; } v = { f };
    uint32_t sign = v.u & 0x80000000u;
    uint32_t a = v.u & 0x7fffffffu;
    if (a >= 0x7f800000u) return f;           /* inf / nan */
    if (a >= 0x477ff000u) {                   /* >= 65520: overflows to inf */
        v.u = sign | 0x7f800000u; return v.f;
    }
    if (a < 0x38800000u) {                    /* below 2^-14: FP16 subnormal grid 2^-24 */
        float m = fabsf(f) * 16777216.0f;     /* exact scaling by 2^24 */
        float r = nearbyintf(m);              /* ties to even (default rounding mode) */
        float out = r / 16777216.0f;
        return sign ? -out : out;
    }
    /* normal: keep 10 mantissa bits, round to nearest even */
    uint32_t lsb = (a >> 13) & 1u;
    uint32_t rounded = (a + 0xfffu + lsb) & ~0x1fffu;
    v.u = sign | rounded;
    return v.f;
}

void orc_dwt97_fwd_f32(float *p, int w, int h, int stride, int levels, int ll_fp16)
{
    int maxn = (w > h ? w : h);
    float *buf = (float *)malloc(sizeof(float) * (3 * maxn + 64));
    float *tmp = (float *)malloc(sizeof(float) * (maxn + 8));
    float *tmp2 = (float *)malloc(sizeof(float) * (maxn + 8));
    int cw = w, ch = h;
    for (int l = 0; l < levels; ++l) {
        int hl = (ch + 1) / 2, wl = (cw + 1) / 2;
        for (int x = 0; x < cw; ++x) {
            for (int y = 0; y < ch; ++y) tmp[y] = p[(size_t)y * stride + x];
            fwd97_1d_f32(tmp, ch, tmp2, tmp2 + hl, 1, buf);
            for (int y = 0; y < ch; ++y) p[(size_t)y * stride + x] = tmp2[y];
        }
        for (int y = 0; y < ch; ++y) {
            float *row = p + (size_t)y * stride;
            memcpy(tmp, row, sizeof(float) * cw);
            fwd97_1d_f32(tmp, cw, row, row + wl, 1, buf);
        }
        if (ll_fp16 && l + 1 < levels)
            for (int y = 0; y < hl; ++y)
                for (int x = 0; x < wl; ++x) p[(size_t)y * stride + x] = orc_round_f16(p[(size_t)y * stride + x]);
        cw = wl; ch = hl;
    }
    free(buf); free(tmp); free(tmp2);
}

void orc_dwt97_inv_f32(float *p, int w, int h, int stride, int levels)
{
    int maxn = (w > h ? w : h);
    float *buf = (float *)malloc(sizeof(float) * (3 * maxn + 64));
    float *tmp = (float *)malloc(sizeof(float) * (maxn + 8));
    float *tmp2 = (float *)malloc(sizeof(float) * (maxn + 8));
    for (int l = levels; l >= 1; --l) {
        int cw = w, ch = h;
        for (int k = 0; k < l - 1; ++k) { cw = (cw + 1) / 2; ch = (ch + 1) / 2; }
        int hl = (ch + 1) / 2, wl = (cw + 1) / 2;
        for (int y = 0; y < ch; ++y) {
            float *row = p + (size_t)y * stride;
            memcpy(tmp, row, sizeof(float) * cw);
            inv97_1d_f32(tmp, tmp + wl, cw, row, 1, buf);
        }
        for (int x = 0; x < cw; ++x) {
            for (int y = 0; y < ch; ++y) tmp[y] = p[(size_t)y * stride + x];
            inv97_1d_f32(tmp, tmp + hl, ch, tmp2, 1, buf);
            for (int y = 0; y < ch; ++y) p[(size_t)y * stride + x] = tmp2[y];
        }
    }
    free(buf); free(tmp); free(tmp2);
}

/* ---- 9/7 FP64 (accuracy reference) ---- */
static void fwd97_1d_f64(const double *x, int n, double *lo, double *hi, double *buf)
{
    double *e = buf + EXT;
    if (n == 1) { lo[0] = x[0]; return; }
    for (int i = -4; i < n + 4; ++i) e[i] = x[mirror(i, n)];
    for (int i = -3; i <= n + 2; ++i) if (i & 1) e[i] += A97 * (e[i - 1] + e[i + 1]);
    for (int i = -2; i <= n + 1; ++i) if (!(i & 1)) e[i] += B97 * (e[i - 1] + e[i + 1]);
    for (int i = -1; i <= n; ++i) if (i & 1) e[i] += G97 * (e[i - 1] + e[i + 1]);
    for (int i = 0; i < n; i += 2) e[i] += D97 * (e[i - 1] + e[i + 1]);
    int nl = (n + 1) / 2;
    for (int k = 0; k < nl; ++k) lo[k] = e[2 * k] / K97;
    for (int k = 0; 2 * k + 1 < n; ++k) hi[k] = e[2 * k + 1] * K97;
}

static void inv97_1d_f64(const double *lo, const double *hi, int n, double *x, double *buf)
{
    double *e = buf + EXT;
    if (n == 1) { x[0] = lo[0]; return; }
    double *t = buf + EXT + n + EXT + 8;
    for (int i = 0; i < n; ++i) t[i] = (i & 1) ? hi[i >> 1] / K97 : lo[i >> 1] * K97;
    for (int i = -4; i < n + 4; ++i) e[i] = t[mirror(i, n)];
    for (int i = -3; i <= n + 2; ++i) if (!(i & 1)) e[i] -= D97 * (e[i - 1] + e[i + 1]);
    for (int i = -2; i <= n + 1; ++i) if (i & 1) e[i] -= G97 * (e[i - 1] + e[i + 1]);
    for (int i = -1; i <= n; ++i) if (!(i & 1)) e[i] -= B97 * (e[i - 1] + e[i + 1]);
    for (int i = 1; i < n; i += 2) e[i] -= A97 * (e[i - 1] + e[i + 1]);
    for (int i = 0; i < n; ++i) x[i] = e[i];
}

void orc_dwt97_fwd_1d_f64(const double *x, int n, double *out)
{
    double *buf = (double *)malloc(sizeof(double) * (3 * n + 64));
    fwd97_1d_f64(x, n, out, out + (n + 1) / 2, buf);
    free(buf);
}

void orc_dwt97_fwd_f64(double *p, int w, int h, int stride, int levels)
{
    int maxn = (w > h ? w : h);
    double *buf = (double *)malloc(sizeof(double) * (3 * maxn + 64));
    double *tmp = (double *)malloc(sizeof(double) * (maxn + 8));
    double *tmp2 = (double *)malloc(sizeof(double) * (maxn + 8));
    int cw = w, ch = h;
    for (int l = 0; l < levels; ++l) {
        int hl = (ch + 1) / 2, wl = (cw + 1) / 2;
        for (int x = 0; x < cw; ++x) {
            for (int y = 0; y < ch; ++y) tmp[y] = p[(size_t)y * stride + x];
            fwd97_1d_f64(tmp, ch, tmp2, tmp2 + hl, buf);
            for (int y = 0; y < ch; ++y) p[(size_t)y * stride + x] = tmp2[y];
        }
        for (int y = 0; y < ch; ++y) {
            double *row = p + (size_t)y * stride;
            memcpy(tmp, row, sizeof(double) * cw);
            fwd97_1d_f64(tmp, cw, row, row + wl, buf);
        }
        cw = wl; ch = hl;
    }
    free(buf); free(tmp); free(tmp2);
}

void orc_dwt97_inv_f64(double *p, int w, int h, int stride, int levels)
{
    int maxn = (w > h ? w : h);
    double *buf = (double *)malloc(sizeof(double) * (3 * maxn + 64));
    double *tmp = (double *)malloc(sizeof(double) * (maxn + 8));
    double *tmp2 = (double *)malloc(sizeof(double) * (maxn + 8));
    for (int l = levels; l >= 1; --l) {
        int cw = w, ch = h;
        for (int k = 0; k < l - 1; ++k) { cw = (cw + 1) / 2; ch = (ch + 1) / 2; }
        int hl = (ch + 1) / 2, wl = (cw + 1) / 2;
        for (int y = 0; y < ch; ++y) {
            double *row = p + (size_t)y * stride;
            memcpy(tmp, row, sizeof(double) * cw);
            inv97_1d_f64(tmp, tmp + wl, cw, row, buf);
        }
        for (int x = 0; x < cw; ++x) {
            for (int y = 0; y < ch; ++y) tmp[y] = p[(size_t)y * stride + x];
            inv97_1d_f64(tmp, tmp + hl, ch, tmp2, buf);
            for (int y = 0; y < ch; ++y) p[(size_t)y * stride + x] = tmp2[y];
        }
    }
    free(buf); free(tmp); free(tmp2);
}

/* ---- band geometry (Mallat layout, image origin 0) ---- */
int orc_bands(int w, int h, int levels, orc_band *b)
{
    int ws[40], hs[40];
    ws[0] = w; hs[0] = h;
    for (int l = 1; l <= levels; ++l) { ws[l] = (ws[l - 1] + 1) / 2; hs[l] = (hs[l - 1] + 1) / 2; }
    int n = 0;
    b[n].x0 = 0; b[n].y0 = 0; b[n].w = ws[levels]; b[n].h = hs[levels];
    b[n].orient = 0; b[n].level = levels; b[n].res = 0; b[n].band = n; ++n;
    for (int l = levels; l >= 1; --l) {
        int wl = ws[l], hl = hs[l], wp = ws[l - 1], hp = hs[l - 1];
        int res = levels - l + 1;
        /* HL: right of LL, LH: below LL, HH: diagonal */
        b[n] = (orc_band){ wl, 0, wp - wl, hl, 1, l, res, n }; ++n;
        b[n] = (orc_band){ 0, hl, wl, hp - hl, 2, l, res, n }; ++n;
        b[n] = (orc_band){ wl, hl, wp - wl, hp - hl, 3, l, res, n }; ++n;
    }
    return n;
}
