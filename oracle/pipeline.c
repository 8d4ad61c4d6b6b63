/*
 * Oracle: whole-image encode (colour -> DWT -> quantisation -> tier-1), the
 * CPU restatement timed as the CPU baseline.  TEST INFRASTRUCTURE ONLY.
 *
 * Stage order follows src/OCLEncoder.cpp:43-52 (upload -> DWT over all
 * levels -> planar split -> bit-plane coder); the colour transform the
 * reference lacks is added in front (T.800 G.1-G.3), and code-blocks are
 * tier-1 coded in parallel with OpenMP.
 */
#include "j2k_oracle.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int orc_encode_image(const uint8_t *img, int w, int h, const orc_encode_cfg *cfg,
                     orc_cb_result *res, uint8_t *slots, int slot_bytes, int ncb_cap,
                     int32_t *qplanes_out, int32_t *passlen_out)
{
    int nc = cfg->ncomp, L = cfg->levels;
    size_t npix = (size_t)w * h;
    if (nc != 1 && nc != 3) return -1;
#ifdef _OPENMP
    if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
    int32_t *q = qplanes_out ? qplanes_out : (int32_t *)malloc(sizeof(int32_t) * npix * nc);
    if (cfg->reversible) {
        if (nc == 3) orc_rct_fwd(img, (int)npix, q, q + npix, q + 2 * npix);
        else orc_shift_i32(img, (int)npix, 1, 0, q);
#pragma omp parallel for
        for (int c = 0; c < nc; ++c) orc_dwt53_fwd(q + c * npix, w, h, w, L);
    } else {
        float *f = (float *)malloc(sizeof(float) * npix * nc);
        if (nc == 3) orc_ict_fwd_f32(img, (int)npix, f, f + npix, f + 2 * npix);
        else orc_shift_f32(img, (int)npix, 1, 0, f);
        float inv[3 * 32 + 1];
        orc_band b[3 * 32 + 1];
        int nb = orc_bands(w, h, L, b);
        for (int k = 0; k < nb; ++k) {
            int Rb = orc_band_range(8, b[k].orient), eps, mu;
            orc_expmant_for_step(Rb, cfg->delta, &eps, &mu);
            inv[k] = (float)(1.0 / orc_step_from_expmant(Rb, eps, mu));
        }
#pragma omp parallel for
        for (int c = 0; c < nc; ++c) {
            orc_dwt97_fwd_f32(f + c * npix, w, h, w, L, 0);
            orc_quantize_f32(f + c * npix, w, h, w, L, inv, q + c * npix);
        }
        free(f);
    }
    int ncb = orc_codeblocks(w, h, L, cfg->cbw, cfg->cbh, NULL, 0);
    if (ncb * nc > ncb_cap) { if (!qplanes_out) free(q); return -2; }
    orc_cb_desc *d = (orc_cb_desc *)malloc(sizeof(orc_cb_desc) * ncb);
    orc_codeblocks(w, h, L, cfg->cbw, cfg->cbh, d, ncb);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < ncb * nc; ++i) {
        int c = i / ncb, k = i % ncb;
        const int32_t *src = q + c * npix + (size_t)d[k].y0 * w + d[k].x0;
        orc_t1_encode_style(src, d[k].w, d[k].h, w, d[k].orient, cfg->style,
                            slots + (size_t)i * slot_bytes, slot_bytes, &res[i],
                            passlen_out ? passlen_out + (size_t)i * ORC_MAX_PASSES : NULL);
    }
    free(d);
    if (!qplanes_out) free(q);
    return ncb * nc;
}
