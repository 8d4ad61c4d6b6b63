/*
 * Oracle: rate control by post-compression rate-distortion optimisation
 * (EBCOT's PCRD-opt).  TEST INFRASTRUCTURE ONLY.
 *
 * The reference has no rate control (its bit-plane coder writes no bytes,
 * src/CL/oclbpc.cl:279-746); this is the published algorithm (Taubman,
 * "High performance scalable image compression with EBCOT", IEEE TIP 2000,
 * section IV; T.800 leaves the choice of truncation points to the encoder):
 *   * every coding pass end of a code-block is a candidate truncation point
 *     n with cumulative rate R_n (bytes of the first n pass segments, exact
 *     with TERMALL) and distortion reduction D_n = w_b * sum of the passes'
 *     reductions (orc_t1_dsig / orc_t1_dref units, w_b the block's weight);
 *   * only points on the upper convex hull of (R, D) are kept, with the
 *     slope to the previous hull point, decreasing along the hull;
 *   * a slope threshold lambda keeps, in every block, the last hull point
 *     whose slope is >= lambda; the threshold is the smallest one whose
 *     total rate is <= the target.  Slopes are compared through the bit
 *     pattern of the (non-negative) double, which orders like the value.
 */
#include "j2k_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint64_t slope_key(double s)
{
    uint64_t k;
    memcpy(&k, &s, sizeof k);
    return k;
}

int orc_rd_hull(int npasses, const int32_t *passlen, const int64_t *dist, double w, int32_t *hn, int64_t *hr,
                double *hs)
{
    int m = 0;              /* hull points so far (the origin is implicit) */
    int64_t hd[ORC_MAX_PASSES + 1];
    long r = 0;
    int64_t d = 0;
    for (int n = 1; n <= npasses; ++n) {
        r += passlen[n - 1];
        d += dist[n - 1];
        for (;;) {
            const long r0 = m ? (long)hr[m - 1] : 0;
            const int64_t d0 = m ? hd[m - 1] : 0;
            if (d <= d0) goto skip;                      /* no gain over the hull end */
            const double s = r > r0 ? (w * (double)(d - d0)) / (double)(r - r0) : HUGE_VAL;
            if (m && s >= hs[m - 1]) {                   /* the hull end lies under the chord: drop it */
                --m;
                continue;
            }
            hn[m] = n; hr[m] = r; hd[m] = d; hs[m] = s;
            ++m;
            break;
        }
    skip:;
    }
    return m;
}

long orc_rate_control(int nblocks, const int32_t *npasses, const int32_t *passlen, const int64_t *dist, int stride,
                      const double *weight, long target, int32_t *ntrunc, int32_t *rtrunc)
{
    int32_t *cnt = (int32_t *)calloc((size_t)nblocks + 1, sizeof(int32_t));
    int32_t *hn = (int32_t *)malloc(sizeof(int32_t) * (size_t)nblocks * (ORC_MAX_PASSES + 1));
    int64_t *hr = (int64_t *)malloc(sizeof(int64_t) * (size_t)nblocks * (ORC_MAX_PASSES + 1));
    double *hs = (double *)malloc(sizeof(double) * (size_t)nblocks * (ORC_MAX_PASSES + 1));
    long npts = 0;
    for (int b = 0; b < nblocks; ++b) {
        const size_t o = (size_t)b * (ORC_MAX_PASSES + 1);
        cnt[b] = orc_rd_hull(npasses[b], passlen + (size_t)b * stride, dist + (size_t)b * stride, weight[b], hn + o,
                             hr + o, hs + o);
        npts += cnt[b];
    }
    /* every hull point as (key, rate increment over its predecessor), highest key first */
    typedef struct { uint64_t key; long dr; } seg;
    seg *sg = (seg *)malloc(sizeof(seg) * (size_t)(npts ? npts : 1));
    long k = 0;
    for (int b = 0; b < nblocks; ++b) {
        const size_t o = (size_t)b * (ORC_MAX_PASSES + 1);
        for (int i = 0; i < cnt[b]; ++i) {
            sg[k].key = slope_key(hs[o + i]);
            sg[k].dr = (long)hr[o + i] - (i ? (long)hr[o + i - 1] : 0);
            ++k;
        }
    }
    /* insertion into key order (descending); sizes here are small */
    for (long i = 1; i < npts; ++i) {
        seg x = sg[i];
        long j = i - 1;
        while (j >= 0 && sg[j].key < x.key) { sg[j + 1] = sg[j]; --j; }
        sg[j + 1] = x;
    }
    /* smallest threshold key K with total rate <= target: walk the distinct
     * keys downwards; the rate at threshold K counts every point with key >= K */
    uint64_t K = UINT64_MAX;      /* nothing kept */
    long total = 0;
    if (target >= 0) {
        long i = 0;
        K = npts ? sg[0].key + 1 : 0;
        while (i < npts) {
            long j = i, add = 0;
            while (j < npts && sg[j].key == sg[i].key) add += sg[j++].dr;
            if (total + add > target) break;
            total += add;
            K = sg[i].key;
            i = j;
        }
    }
    for (int b = 0; b < nblocks; ++b) {
        const size_t o = (size_t)b * (ORC_MAX_PASSES + 1);
        int keep = 0;
        while (keep < cnt[b] && slope_key(hs[o + keep]) >= K) ++keep;
        ntrunc[b] = keep ? hn[o + keep - 1] : 0;
        rtrunc[b] = keep ? (int32_t)hr[o + keep - 1] : 0;
    }
    free(sg); free(hs); free(hr); free(hn); free(cnt);
    return total;
}
