/*
 * Oracle: EBCOT tier-1 (T.800 Annex D) with the MQ arithmetic coder (Annex C).
 * TEST INFRASTRUCTURE ONLY (see j2k_oracle.h).
 *
 * Reference counterpart: src/CL/oclbpc.cl:279-746 (intended tier-1 coder:
 * per-code-block MSB at :287-344, bit-plane loop at :454-746, context tables
 * only as comments at :204-275, no MQ coder and no output) driven by
 * src/OCLBPC.cpp:38-51.  This restatement is the sequential scan of T.800
 * D.3-D.5: 4-row stripes, column-major inside a stripe; significance
 * propagation, magnitude refinement and cleanup passes (first bit-plane:
 * cleanup only); ZC/SC/MR/RL/UNIFORM contexts 0..18 (Tables D.1-D.4, D.7);
 * MQ coder INITENC/ENCODE/RENORME/BYTEOUT/FLUSH and the decoder (C.2, C.3)
 * with the 47-state probability table C.2.
 */
#include "j2k_oracle.h"
#include <stdlib.h>
#include <string.h>

/* ---------------- MQ probability estimation table (T.800 Table C.2) ---------------- */
static const uint16_t QE[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801,
    0x3001, 0x2401, 0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401,
    0x3001, 0x2801, 0x2401, 0x2201, 0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101,
    0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1, 0x0221, 0x0141, 0x0111, 0x0085,
    0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601 };
static const uint8_t NMPS[47] = {
    1, 2, 3, 4, 5, 38, 7, 8, 9, 10, 11, 12, 13, 29, 15, 16, 17, 18, 19, 20,
    21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40,
    41, 42, 43, 44, 45, 45, 46 };
static const uint8_t NLPS[47] = {
    1, 6, 9, 12, 29, 33, 6, 14, 14, 14, 17, 18, 20, 21, 14, 14, 15, 16, 17, 18,
    19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 37,
    38, 39, 40, 41, 42, 43, 46 };
static const uint8_t SWITCH_[47] = {
    1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0 };

#define NCTX 19
#define CTX_RL 17
#define CTX_UNI 18

/* ---------------- MQ encoder ---------------- */
typedef struct {
    uint32_t A, C;
    int CT;
    uint8_t *buf; /* buf[0] is the byte before the codeword (fake, = 0) */
    long bp, cap;
    int overflow;
    uint8_t I[NCTX], MPS[NCTX];
} mq_enc;

static void mq_reset_ctx(uint8_t *I, uint8_t *M)
{
    memset(I, 0, NCTX);
    memset(M, 0, NCTX);
    I[0] = 4;        /* all-zero-neighbourhood ZC context */
    I[CTX_RL] = 3;
    I[CTX_UNI] = 46;
}

static void mqe_init(mq_enc *e, uint8_t *buf, long cap)
{
    e->A = 0x8000; e->C = 0; e->CT = 12;
    e->buf = buf; e->bp = 0; e->cap = cap; e->overflow = 0;
    buf[0] = 0;
    mq_reset_ctx(e->I, e->MPS);
}

/* Re-initialises the coder registers for a new codeword segment (after a
 * termination) and keeps the context states (C.2.8 INITENC without the
 * context reset; the byte before the segment is never 0xFF, see mqe_flush). */
static void mqe_restart(mq_enc *e, uint8_t *buf, long cap)
{
    e->A = 0x8000; e->C = 0; e->CT = 12;
    e->buf = buf; e->bp = 0; e->cap = cap; e->overflow = 0;
    buf[0] = 0;
}

static inline void mqe_put(mq_enc *e, uint32_t v)
{
    if (e->bp + 1 >= e->cap) { e->overflow = 1; return; }
    e->buf[++e->bp] = (uint8_t)v;
}

static void mqe_byteout(mq_enc *e)
{
    if (e->buf[e->bp] == 0xFF) {
        mqe_put(e, e->C >> 20); e->C &= 0xFFFFF; e->CT = 7;
    } else if (e->C < 0x8000000) {
        mqe_put(e, e->C >> 19); e->C &= 0x7FFFF; e->CT = 8;
    } else {
        e->buf[e->bp]++;
        if (e->buf[e->bp] == 0xFF) {
            e->C &= 0x7FFFFFF;
            mqe_put(e, e->C >> 20); e->C &= 0xFFFFF; e->CT = 7;
        } else {
            mqe_put(e, e->C >> 19); e->C &= 0x7FFFF; e->CT = 8;
        }
    }
}

static inline void mqe_renorm(mq_enc *e)
{
    do {
        e->A <<= 1; e->C <<= 1; e->CT--;
        if (e->CT == 0) mqe_byteout(e);
    } while (!(e->A & 0x8000));
}

static void mqe_encode(mq_enc *e, int cx, int d)
{
    int i = e->I[cx];
    uint32_t qe = QE[i];
    e->A -= qe;
    if (d == e->MPS[cx]) {
        if (!(e->A & 0x8000)) {
            if (e->A < qe) e->A = qe; else e->C += qe;
            e->I[cx] = NMPS[i];
            mqe_renorm(e);
        } else {
            e->C += qe;
        }
    } else {
        if (e->A < qe) e->C += qe; else e->A = qe;
        if (SWITCH_[i]) e->MPS[cx] ^= 1;
        e->I[cx] = NLPS[i];
        mqe_renorm(e);
    }
}

/* Returns the codeword length in bytes. */
static long mqe_flush(mq_enc *e)
{
    uint32_t tmp = e->C + e->A;
    e->C |= 0xFFFF;
    if (e->C >= tmp) e->C -= 0x8000;
    e->C <<= e->CT; mqe_byteout(e);
    e->C <<= e->CT; mqe_byteout(e);
    if (e->buf[e->bp] != 0xFF) e->bp++;
    return e->bp - 1;
}

/* ---------------- MQ decoder ---------------- */
typedef struct {
    uint32_t A, C;
    int CT;
    const uint8_t *d;
    long len, bp;
    uint8_t I[NCTX], MPS[NCTX];
} mq_dec;

static inline uint32_t mqd_byte(const mq_dec *m, long i) { return i < m->len ? m->d[i] : 0xFF; }

static void mqd_bytein(mq_dec *m)
{
    uint32_t b = mqd_byte(m, m->bp), b1 = mqd_byte(m, m->bp + 1);
    if (b == 0xFF) {
        if (b1 > 0x8F) { m->C += 0xFF00; m->CT = 8; }
        else { m->bp++; m->C += b1 << 9; m->CT = 7; }
    } else {
        m->bp++; m->C += b1 << 8; m->CT = 8;
    }
}

/* INITDEC (C.3.5) on a codeword segment; contexts are left untouched. */
static void mqd_start(mq_dec *m, const uint8_t *d, long len)
{
    m->d = d; m->len = len; m->bp = 0;
    m->C = mqd_byte(m, 0) << 16;
    mqd_bytein(m);
    m->C <<= 7; m->CT -= 7; m->A = 0x8000;
}

static void mqd_init(mq_dec *m, const uint8_t *d, long len)
{
    mqd_start(m, d, len);
    mq_reset_ctx(m->I, m->MPS);
}

static inline void mqd_renorm(mq_dec *m)
{
    do {
        if (m->CT == 0) mqd_bytein(m);
        m->A <<= 1; m->C <<= 1; m->CT--;
    } while (!(m->A & 0x8000));
}

static int mqd_decode(mq_dec *m, int cx)
{
    int i = m->I[cx];
    uint32_t qe = QE[i];
    int d;
    m->A -= qe;
    if ((m->C >> 16) < qe) {
        if (m->A < qe) { m->A = qe; d = m->MPS[cx]; m->I[cx] = NMPS[i]; }
        else { m->A = qe; d = 1 - m->MPS[cx]; if (SWITCH_[i]) m->MPS[cx] ^= 1; m->I[cx] = NLPS[i]; }
        mqd_renorm(m);
    } else {
        m->C -= qe << 16;
        if (!(m->A & 0x8000)) {
            if (m->A < qe) { d = 1 - m->MPS[cx]; if (SWITCH_[i]) m->MPS[cx] ^= 1; m->I[cx] = NLPS[i]; }
            else { d = m->MPS[cx]; m->I[cx] = NMPS[i]; }
            mqd_renorm(m);
        } else {
            d = m->MPS[cx];
        }
    }
    return d;
}

/* ---------------- raw (bypass) coding (T.800 D.6) ----------------
 * With the BYPASS code-block style the significance-propagation and
 * magnitude-refinement passes from the fifth bit-plane on (pass index >= 10,
 * counting the first cleanup pass as 0) carry their decisions as plain bits:
 * MSB first, and after an 0xFF byte the next byte holds only 7 (its MSB is a
 * stuffed 0).  A sign is sent as the sign itself (1 = negative), with no
 * context and no XOR prediction.  Termination pads an unfinished byte -- also
 * the 7-bit byte owed after a final 0xFF -- with the alternating pattern
 * 0,1,0,... so a segment never ends in 0xFF. */
typedef struct {
    uint8_t *buf;
    long bp, cap;     /* next byte index */
    uint32_t c;       /* bits of the byte being filled */
    int ct;           /* free bits left in it (8, or 7 after 0xFF) */
    int overflow;
} raw_enc;

static void rawe_init(raw_enc *r, uint8_t *buf, long cap)
{
    r->buf = buf; r->bp = 0; r->cap = cap; r->c = 0; r->ct = 8; r->overflow = 0;
}

static void rawe_emit(raw_enc *r)
{
    if (r->bp < r->cap) r->buf[r->bp] = (uint8_t)r->c;
    else r->overflow = 1;
    r->bp++;
    r->ct = r->c == 0xFF ? 7 : 8;
    r->c = 0;
}

static inline void rawe_put(raw_enc *r, int bit)
{
    r->c = (r->c << 1) | (uint32_t)(bit & 1);
    if (--r->ct == 0) rawe_emit(r);
}

/* Returns the segment length in bytes. */
static long rawe_flush(raw_enc *r)
{
    if (r->ct < 8) {
        const int n = r->ct;   /* the last put of the pattern completes the byte */
        for (int k = 0; k < n; ++k) rawe_put(r, k & 1);
    }
    return r->bp;
}

typedef struct {
    const uint8_t *d;
    long len, bp;
    uint32_t c;       /* current byte */
    int ct;           /* bits of it not yet read */
} raw_dec;

static void rawd_start(raw_dec *r, const uint8_t *d, long len)
{
    r->d = d; r->len = len; r->bp = 0; r->c = 0; r->ct = 0;
}

static inline int rawd_get(raw_dec *r)
{
    if (r->ct == 0) {
        const int after_ff = r->c == 0xFF;
        r->c = r->bp < r->len ? r->d[r->bp] : 0xFF;
        r->bp++;
        r->ct = after_ff ? 7 : 8;
    }
    r->ct--;
    return (int)((r->c >> r->ct) & 1u);
}

/* Pass index k (0 = the first cleanup pass) is raw-coded under BYPASS.  Only
 * BYPASS together with TERMALL is implemented (every pass its own segment);
 * BYPASS alone groups passes into segments differently and is not offered. */
static inline int pass_is_raw(int style, int k)
{
    return (style & ORC_STYLE_BYPASS) && (style & ORC_STYLE_TERMALL) && k >= 10 && k % 3 != 0;
}

/* ---------------- contexts (Tables D.1, D.3, D.4) ---------------- */
int orc_zc_context(int orient, int h, int v, int d)
{
    if (orient == 1) { int t = h; h = v; v = t; } /* HL: roles of h and v swap */
    if (orient == 3) {                            /* HH */
        int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv >= 1 ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv == 1 ? 4 : 3);
        return hv >= 2 ? 2 : (hv == 1 ? 1 : 0);
    }
    /* LL and LH */
    if (h == 2) return 8;
    if (h == 1) return v >= 1 ? 7 : (d >= 1 ? 6 : 5);
    if (v == 2) return 4;
    if (v == 1) return 3;
    return d >= 2 ? 2 : (d == 1 ? 1 : 0);
}

int orc_sc_context(int H, int V, int *xorbit)
{
    /* Table D.3; negative H (or H == 0 with V < 0) flips the sign prediction */
    int x = 0, ctx;
    if (H < 0 || (H == 0 && V < 0)) { x = 1; H = -H; V = -V; }
    if (H == 1) ctx = V == 1 ? 13 : (V == 0 ? 12 : 11);
    else ctx = V == 1 ? 10 : 9;
    *xorbit = x;
    return ctx;
}

/* ---------------- coefficient state ---------------- */
enum { F_SIG = 1, F_VIS = 2, F_REF = 4, F_NEG = 8 };

typedef struct {
    uint8_t *f;   /* (w+2) x (h+2) with a zero border */
    uint32_t *m;  /* magnitudes, w x h */
    int w, h, fs, orient;
    int64_t dacc; /* distortion reduction of the current pass (orc_t1_dsig / orc_t1_dref units) */
} cbstate;

#define FL(s, x, y) ((s)->f[((y) + 1) * (s)->fs + (x) + 1])

static inline int zc_of(const cbstate *s, int x, int y)
{
    const uint8_t *c = &FL(s, x, y);
    int fs = s->fs;
    int h = (c[-1] & F_SIG) + (c[1] & F_SIG);
    int v = (c[-fs] & F_SIG) + (c[fs] & F_SIG);
    int d = (c[-fs - 1] & F_SIG) + (c[-fs + 1] & F_SIG) + (c[fs - 1] & F_SIG) + (c[fs + 1] & F_SIG);
    return orc_zc_context(s->orient, h, v, d);
}

static inline int contrib(uint8_t f) { return (f & F_SIG) ? ((f & F_NEG) ? -1 : 1) : 0; }

static inline int sc_of(const cbstate *s, int x, int y, int *xr)
{
    const uint8_t *c = &FL(s, x, y);
    int fs = s->fs;
    int H = contrib(c[-1]) + contrib(c[1]);
    int V = contrib(c[-fs]) + contrib(c[fs]);
    H = H > 1 ? 1 : (H < -1 ? -1 : H);
    V = V > 1 ? 1 : (V < -1 ? -1 : V);
    return orc_sc_context(H, V, xr);
}

static inline int any_sig_nbr(const cbstate *s, int x, int y)
{
    const uint8_t *c = &FL(s, x, y);
    int fs = s->fs;
    return ((c[-1] | c[1] | c[-fs] | c[fs] | c[-fs - 1] | c[-fs + 1] | c[fs - 1] | c[fs + 1]) & F_SIG) != 0;
}

/* ---------------- distortion (rate control) ----------------
 * Squared error of a magnitude m against the decoder's reconstruction, in
 * index units doubled (so every value is an integer): with the bits of m at
 * and above plane p known, the reconstruction is their value plus half the
 * remaining interval, 2^(p-1) (the exact m at p = 0), and an insignificant
 * sample reconstructs as 0. */
int64_t orc_t1_dsig(uint32_t m, int p)
{
    /* m becomes significant at plane p (m >> p == 1): error 2m -> 2m - 3*2^p */
    const int64_t e0 = 2 * (int64_t)m;
    const int64_t e1 = p > 0 ? e0 - (3ll << p) : 0;
    return e0 * e0 - e1 * e1;
}

int64_t orc_t1_dref(uint32_t m, int p)
{
    /* bit p of a sample significant above p is refined */
    const int64_t m2 = 2 * (int64_t)m;
    const int64_t eb = m2 - ((((int64_t)(m >> (p + 1))) << (p + 2)) + (2ll << p));
    const int64_t ea = p > 0 ? m2 - ((((int64_t)(m >> p)) << (p + 1)) + (1ll << p)) : 0;
    return eb * eb - ea * ea;
}

/* ---------------- encoder passes ---------------- */
/* raw != NULL: the pass is raw-coded (bypass), contexts are not used */
static void enc_spp(cbstate *s, mq_enc *e, raw_enc *raw, int p)
{
    for (int y0 = 0; y0 < s->h; y0 += 4)
        for (int x = 0; x < s->w; ++x)
            for (int y = y0; y < y0 + 4 && y < s->h; ++y) {
                uint8_t *f = &FL(s, x, y);
                if (*f & F_SIG) continue;
                int cx = zc_of(s, x, y);
                if (cx == 0) continue;
                int bit = (s->m[y * s->w + x] >> p) & 1;
                if (raw) rawe_put(raw, bit);
                else mqe_encode(e, cx, bit);
                *f |= F_VIS;
                if (bit) {
                    const int neg = (*f & F_NEG) ? 1 : 0;
                    s->dacc += orc_t1_dsig(s->m[y * s->w + x], p);
                    if (raw) {
                        rawe_put(raw, neg);
                    } else {
                        int xr, sc = sc_of(s, x, y, &xr);
                        mqe_encode(e, sc, neg ^ xr);
                    }
                    *f |= F_SIG;
                }
            }
}

static void enc_mrp(cbstate *s, mq_enc *e, raw_enc *raw, int p)
{
    for (int y0 = 0; y0 < s->h; y0 += 4)
        for (int x = 0; x < s->w; ++x)
            for (int y = y0; y < y0 + 4 && y < s->h; ++y) {
                uint8_t *f = &FL(s, x, y);
                if ((*f & (F_SIG | F_VIS)) != F_SIG) continue;
                const int bit = (s->m[y * s->w + x] >> p) & 1;
                s->dacc += orc_t1_dref(s->m[y * s->w + x], p);
                if (raw) {
                    rawe_put(raw, bit);
                } else {
                    int cx = (*f & F_REF) ? 16 : (any_sig_nbr(s, x, y) ? 15 : 14);
                    mqe_encode(e, cx, bit);
                }
                *f |= F_REF;
            }
}

static void enc_cup(cbstate *s, mq_enc *e, int p)
{
    for (int y0 = 0; y0 < s->h; y0 += 4)
        for (int x = 0; x < s->w; ++x) {
            int y = y0, yend = y0 + 4 < s->h ? y0 + 4 : s->h;
            if (y0 + 4 <= s->h) {
                int rl = 1;
                for (int k = 0; k < 4 && rl; ++k)
                    if ((FL(s, x, y0 + k) & (F_SIG | F_VIS)) || zc_of(s, x, y0 + k) != 0) rl = 0;
                if (rl) {
                    int k = 0;
                    while (k < 4 && !((s->m[(y0 + k) * s->w + x] >> p) & 1)) ++k;
                    if (k == 4) { mqe_encode(e, CTX_RL, 0); continue; }
                    mqe_encode(e, CTX_RL, 1);
                    mqe_encode(e, CTX_UNI, k >> 1);
                    mqe_encode(e, CTX_UNI, k & 1);
                    uint8_t *f = &FL(s, x, y0 + k);
                    int xr, sc = sc_of(s, x, y0 + k, &xr);
                    mqe_encode(e, sc, ((*f & F_NEG) ? 1 : 0) ^ xr);
                    *f |= F_SIG;
                    s->dacc += orc_t1_dsig(s->m[(y0 + k) * s->w + x], p);
                    y = y0 + k + 1;
                }
            }
            for (; y < yend; ++y) {
                uint8_t *f = &FL(s, x, y);
                if (*f & (F_SIG | F_VIS)) continue;
                int bit = (s->m[y * s->w + x] >> p) & 1;
                mqe_encode(e, zc_of(s, x, y), bit);
                if (bit) {
                    int xr, sc = sc_of(s, x, y, &xr);
                    mqe_encode(e, sc, ((*f & F_NEG) ? 1 : 0) ^ xr);
                    *f |= F_SIG;
                    s->dacc += orc_t1_dsig(s->m[y * s->w + x], p);
                }
            }
            for (int yy = y0; yy < yend; ++yy) FL(s, x, yy) &= (uint8_t)~F_VIS;
        }
}

/* Code-block styles (Table A.19): RESET resets every context to its initial
 * state after each coding pass; TERMALL terminates the codeword (C.2.9) after
 * each pass and restarts the coder (C.2.8) for the next one, so every pass is
 * its own codeword segment.  Together they make the passes of a block
 * independent MQ chains.  BYPASS (with TERMALL) sends the significance and
 * refinement passes from the fifth bit-plane on as raw bits (D.6). */
void orc_t1_encode_style(const int32_t *q, int w, int h, int stride, int orient, int style,
                         uint8_t *out, int cap, orc_cb_result *res, int32_t *passlen)
{
    orc_t1_encode_rd(q, w, h, stride, orient, style, out, cap, res, passlen, NULL);
}

void orc_t1_encode_rd(const int32_t *q, int w, int h, int stride, int orient, int style,
                      uint8_t *out, int cap, orc_cb_result *res, int32_t *passlen, int64_t *dist)
{
    cbstate s;
    s.w = w; s.h = h; s.fs = w + 2; s.orient = orient; s.dacc = 0;
    s.f = (uint8_t *)calloc((size_t)(w + 2) * (h + 2), 1);
    s.m = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)w * h + 4);
    uint32_t maxm = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int32_t v = q[(size_t)y * stride + x];
            uint32_t m = v < 0 ? (uint32_t)(-(int64_t)v) : (uint32_t)v;
            s.m[y * w + x] = m;
            if (v < 0) FL(&s, x, y) |= F_NEG;
            if (m > maxm) maxm = m;
        }
    int nbp = 0;
    while (nbp < 32 && (maxm >> nbp)) ++nbp;
    res->flags = 0;
    res->nbp = nbp;
    if (nbp == 0) {
        res->len = 0; res->npasses = 0;
        free(s.f); free(s.m);
        return;
    }
    long icap = (long)w * h * (nbp + 2) + 1024;
    uint8_t *buf = (uint8_t *)malloc((size_t)icap);
    uint8_t *code = (uint8_t *)malloc((size_t)icap);   /* concatenated segments */
    long total = 0;
    int overflow = 0, pass = 0;
    mq_enc e;
    mqe_init(&e, buf, icap);
    for (int p = nbp - 1; p >= 0; --p)
        for (int t = (p == nbp - 1) ? 2 : 0; t < 3; ++t) {
            const int last = (p == 0 && t == 2);
            if (pass_is_raw(style, pass)) {
                /* a raw pass is its own segment (BYPASS is used with TERMALL) */
                raw_enc r;
                rawe_init(&r, code + total, icap - total);
                if (t == 0) enc_spp(&s, &e, &r, p);
                else enc_mrp(&s, &e, &r, p);
                const long len = rawe_flush(&r);
                overflow |= r.overflow;
                total += len;
                if (passlen) passlen[pass] = (int32_t)len;
            } else {
                if (t == 0) enc_spp(&s, &e, NULL, p);
                else if (t == 1) enc_mrp(&s, &e, NULL, p);
                else enc_cup(&s, &e, p);
                if ((style & ORC_STYLE_TERMALL) || last) {
                    long len = mqe_flush(&e);
                    overflow |= e.overflow;
                    memcpy(code + total, buf + 1, (size_t)len);
                    total += len;
                    if (passlen) passlen[pass] = (int32_t)len;
                    if (!last) mqe_restart(&e, buf, icap);
                }
            }
            if (style & ORC_STYLE_RESET) mq_reset_ctx(e.I, e.MPS);
            if (dist) dist[pass] = s.dacc;
            s.dacc = 0;
            ++pass;
        }
    res->npasses = 3 * nbp - 2;
    res->len = (int32_t)total;
    if (overflow || total > cap) res->flags |= 1;
    else memcpy(out, code, (size_t)total);
    free(code); free(buf); free(s.f); free(s.m);
}

void orc_t1_encode(const int32_t *q, int w, int h, int stride, int orient,
                   uint8_t *out, int cap, orc_cb_result *res)
{
    orc_t1_encode_style(q, w, h, stride, orient, 0, out, cap, res, NULL);
}

/* ---------------- decoder passes ---------------- */
/* raw != NULL: the pass is raw-coded (bypass) */
static void dec_spp(cbstate *s, mq_dec *m, raw_dec *raw, int p)
{
    for (int y0 = 0; y0 < s->h; y0 += 4)
        for (int x = 0; x < s->w; ++x)
            for (int y = y0; y < y0 + 4 && y < s->h; ++y) {
                uint8_t *f = &FL(s, x, y);
                if (*f & F_SIG) continue;
                int cx = zc_of(s, x, y);
                if (cx == 0) continue;
                int bit = raw ? rawd_get(raw) : mqd_decode(m, cx);
                *f |= F_VIS;
                if (bit) {
                    int neg;
                    if (raw) {
                        neg = rawd_get(raw);
                    } else {
                        int xr, sc = sc_of(s, x, y, &xr);
                        neg = mqd_decode(m, sc) ^ xr;
                    }
                    *f |= F_SIG | (neg ? F_NEG : 0);
                    s->m[y * s->w + x] |= 1u << p;
                }
            }
}

static void dec_mrp(cbstate *s, mq_dec *m, raw_dec *raw, int p)
{
    for (int y0 = 0; y0 < s->h; y0 += 4)
        for (int x = 0; x < s->w; ++x)
            for (int y = y0; y < y0 + 4 && y < s->h; ++y) {
                uint8_t *f = &FL(s, x, y);
                if ((*f & (F_SIG | F_VIS)) != F_SIG) continue;
                int cx = (*f & F_REF) ? 16 : (any_sig_nbr(s, x, y) ? 15 : 14);
                if (raw ? rawd_get(raw) : mqd_decode(m, cx)) s->m[y * s->w + x] |= 1u << p;
                *f |= F_REF;
            }
}

static void dec_cup(cbstate *s, mq_dec *m, int p)
{
    for (int y0 = 0; y0 < s->h; y0 += 4)
        for (int x = 0; x < s->w; ++x) {
            int y = y0, yend = y0 + 4 < s->h ? y0 + 4 : s->h;
            if (y0 + 4 <= s->h) {
                int rl = 1;
                for (int k = 0; k < 4 && rl; ++k)
                    if ((FL(s, x, y0 + k) & (F_SIG | F_VIS)) || zc_of(s, x, y0 + k) != 0) rl = 0;
                if (rl) {
                    if (!mqd_decode(m, CTX_RL)) continue;
                    int k = mqd_decode(m, CTX_UNI) << 1;
                    k |= mqd_decode(m, CTX_UNI);
                    uint8_t *f = &FL(s, x, y0 + k);
                    int xr, sc = sc_of(s, x, y0 + k, &xr);
                    int neg = mqd_decode(m, sc) ^ xr;
                    *f |= F_SIG | (neg ? F_NEG : 0);
                    s->m[(y0 + k) * s->w + x] |= 1u << p;
                    y = y0 + k + 1;
                }
            }
            for (; y < yend; ++y) {
                uint8_t *f = &FL(s, x, y);
                if (*f & (F_SIG | F_VIS)) continue;
                if (mqd_decode(m, zc_of(s, x, y))) {
                    int xr, sc = sc_of(s, x, y, &xr);
                    int neg = mqd_decode(m, sc) ^ xr;
                    *f |= F_SIG | (neg ? F_NEG : 0);
                    s->m[y * s->w + x] |= 1u << p;
                }
            }
            for (int yy = y0; yy < yend; ++yy) FL(s, x, yy) &= (uint8_t)~F_VIS;
        }
}

int orc_t1_decode_style(const uint8_t *data, int len, int w, int h, int orient, int nbp, int npasses,
                        int style, const int32_t *passlen, int32_t *q, int stride)
{
    cbstate s;
    s.w = w; s.h = h; s.fs = w + 2; s.orient = orient;
    s.f = (uint8_t *)calloc((size_t)(w + 2) * (h + 2), 1);
    s.m = (uint32_t *)calloc((size_t)w * h + 4, sizeof(uint32_t));
    if (nbp > 0 && npasses > 0) {
        const int termall = (style & ORC_STYLE_TERMALL) && passlen;
        mq_dec m;
        raw_dec r;
        long off = 0;
        mqd_init(&m, data, termall ? passlen[0] : len);
        int pass = 0;
        for (int p = nbp - 1; p >= 0 && pass < npasses; --p)
            for (int t = (p == nbp - 1) ? 2 : 0; t < 3 && pass < npasses; ++t) {
                const int raw = termall && pass_is_raw(style, pass);
                if (pass > 0) {
                    if (style & ORC_STYLE_RESET) mq_reset_ctx(m.I, m.MPS);
                    if (termall) {
                        off += passlen[pass - 1];
                        if (raw) rawd_start(&r, data + off, passlen[pass]);
                        else mqd_start(&m, data + off, passlen[pass]);
                    }
                }
                if (t == 0) dec_spp(&s, &m, raw ? &r : NULL, p);
                else if (t == 1) dec_mrp(&s, &m, raw ? &r : NULL, p);
                else dec_cup(&s, &m, p);
                ++pass;
            }
    }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int32_t v = (int32_t)s.m[y * w + x];
            q[(size_t)y * stride + x] = (FL(&s, x, y) & F_NEG) ? -v : v;
        }
    free(s.f); free(s.m);
    return 0;
}

int orc_t1_decode(const uint8_t *data, int len, int w, int h, int orient,
                  int nbp, int npasses, int32_t *q, int stride)
{
    return orc_t1_decode_style(data, len, w, h, orient, nbp, npasses, 0, NULL, q, stride);
}

/* ---------------- code-block enumeration ---------------- */
int orc_codeblocks(int w, int h, int levels, int cbw, int cbh, orc_cb_desc *out, int cap)
{
    orc_band b[3 * 32 + 1];
    int nb = orc_bands(w, h, levels, b), n = 0;
    for (int k = 0; k < nb; ++k) {
        if (b[k].w <= 0 || b[k].h <= 0) continue;
        int nx = (b[k].w + cbw - 1) / cbw, ny = (b[k].h + cbh - 1) / cbh;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (n < cap) {
                    orc_cb_desc *d = &out[n];
                    d->x0 = b[k].x0 + i * cbw; d->y0 = b[k].y0 + j * cbh;
                    d->w = (b[k].w - i * cbw) < cbw ? (b[k].w - i * cbw) : cbw;
                    d->h = (b[k].h - j * cbh) < cbh ? (b[k].h - j * cbh) : cbh;
                    d->orient = b[k].orient; d->band = b[k].band; d->res = b[k].res;
                    d->cbx = i; d->cby = j;
                }
                ++n;
            }
    }
    return n;
}
