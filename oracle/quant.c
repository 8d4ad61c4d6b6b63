/*
 * Oracle: scalar dead-zone quantisation (T.800 Annex E).
 * TEST INFRASTRUCTURE ONLY (see j2k_oracle.h).
 *
 * Reference counterpart: src/OCLDWTForward.cpp:99-122 (getStep) and the
 * quantised writes of src/CL/ocldwt97.cl:429-481 (convert_int_rte(ceil(q*v))
 * with a hard-coded exponent/mantissa).  This restatement follows T.800
 * E.1: Delta_b = 2^(R_b - eps_b) * (1 + mu_b / 2^11), q = sign(y) *
 * floor(|y| / Delta_b), and the decoder-side reconstruction
 * y = sign(q) * (|q| + r) * Delta_b with r = 1/2 (E.1.1.2).
 *
 * FP32 contract shared with the HIP kernels: q = (int)(fabsf(y) * inv_step)
 * with inv_step = (float)(1.0 / Delta_b) computed in double on the host.
 */
#include "j2k_oracle.h"
#include <math.h>

int orc_band_range(int bitdepth, int orient)
{
    int gain = orient == 0 ? 0 : (orient == 3 ? 2 : 1);
    return bitdepth + gain;
}

double orc_step_from_expmant(int Rb, int eps, int mu)
{
    return ldexp(1.0 + mu / 2048.0, Rb - eps);
}

void orc_expmant_for_step(int Rb, double delta, int *eps, int *mu)
{
    int e;
    double m = frexp(delta, &e); /* delta = m * 2^e, m in [0.5, 1) */
    m *= 2.0; e -= 1;            /* m in [1, 2) */
    int u = (int)floor((m - 1.0) * 2048.0 + 0.5);
    if (u >= 2048) { u = 0; e += 1; }
    int ep = Rb - e;
    if (ep < 0) ep = 0;
    if (ep > 31) ep = 31;
    *eps = ep; *mu = u;
}

void orc_quantize_f32(const float *p, int w, int h, int stride, int levels,
                      const float *inv_step, int32_t *q)
{
    orc_band b[3 * 32 + 1];
    int nb = orc_bands(w, h, levels, b);
    for (int k = 0; k < nb; ++k) {
        float inv = inv_step[k];
        for (int y = b[k].y0; y < b[k].y0 + b[k].h; ++y)
            for (int x = b[k].x0; x < b[k].x0 + b[k].w; ++x) {
                float v = p[(size_t)y * stride + x];
                int32_t m = (int32_t)(fabsf(v) * inv);
                q[(size_t)y * stride + x] = v < 0.0f ? -m : m;
            }
    }
}

void orc_dequantize_f32(const int32_t *q, int w, int h, int stride, int levels,
                        const float *step, float *y)
{
    orc_band b[3 * 32 + 1];
    int nb = orc_bands(w, h, levels, b);
    for (int k = 0; k < nb; ++k) {
        float s = step[k];
        for (int yy = b[k].y0; yy < b[k].y0 + b[k].h; ++yy)
            for (int x = b[k].x0; x < b[k].x0 + b[k].w; ++x) {
                int32_t v = q[(size_t)yy * stride + x];
                float m = ((float)(v < 0 ? -v : v) + 0.5f) * s;
                y[(size_t)yy * stride + x] = v == 0 ? 0.0f : (v < 0 ? -m : m);
            }
    }
}
