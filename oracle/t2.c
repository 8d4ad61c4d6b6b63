/*
 * Oracle: minimal tier-2 codestream writer (T.800 Annex A markers, Annex B
 * packets): one tile, one quality layer, LRCP order, default (maximal)
 * precincts, no SOP/EPH, default code-block style.
 * TEST INFRASTRUCTURE ONLY (see j2k_oracle.h).  The reference stops before
 * tier-2 (src/OCLEncoder.cpp:43-52 ends at the bit-plane coder); this writer
 * exists so that an independent JPEG 2000 decoder can check the tier-1
 * output, the quantiser signalling and the DWT conventions end to end.
 *
 * Packet headers follow B.10: tag-tree coding of inclusion and of the
 * number of missing MSBs (B.10.2), the coding-pass codewords of Table B.4,
 * Lblock length signalling (B.10.7.1) and bit stuffing after 0xFF (B.10.1).
 */
#include "j2k_oracle.h"
#include <stdlib.h>
#include <string.h>

/* ---------------- byte / bit output ---------------- */
typedef struct {
    uint8_t *b;
    long pos, cap;
    int err;
} wbuf;

static void put8(wbuf *w, int v) { if (w->pos < w->cap) w->b[w->pos] = (uint8_t)v; else w->err = 1; w->pos++; }
static void put16(wbuf *w, int v) { put8(w, v >> 8); put8(w, v); }
static void put32(wbuf *w, uint32_t v) { put16(w, (int)(v >> 16)); put16(w, (int)(v & 0xffff)); }

typedef struct {
    wbuf *w;
    int cur, nbits, room; /* room: bits in the current byte (8, or 7 after 0xFF) */
    int last;             /* last emitted byte */
} bitw;

static void bw_init(bitw *b, wbuf *w) { b->w = w; b->cur = 0; b->nbits = 0; b->room = 8; b->last = 0; }

static void bw_emit(bitw *b)
{
    put8(b->w, b->cur);
    b->last = b->cur;
    b->room = (b->cur == 0xFF) ? 7 : 8;
    b->cur = 0; b->nbits = 0;
}

static void bw_bit(bitw *b, int v)
{
    b->cur = (b->cur << 1) | (v & 1);
    if (++b->nbits == b->room) bw_emit(b);
}

static void bw_bits(bitw *b, uint32_t v, int n)
{
    for (int i = n - 1; i >= 0; --i) bw_bit(b, (v >> i) & 1);
}

static void bw_flush(bitw *b)
{
    if (b->nbits > 0) {
        while (b->nbits < b->room) { b->cur <<= 1; b->nbits++; }
        bw_emit(b);
    }
    if (b->last == 0xFF) { b->cur = 0; put8(b->w, 0); b->last = 0; }
}

/* ---------------- tag tree (B.10.2) ---------------- */
typedef struct {
    int nlev, w[20], h[20], off[20], total;
    int *val, *low, *known;
} tagtree;

static void tt_init(tagtree *t, int w, int h)
{
    int n = 0, total = 0;
    do {
        t->w[n] = w; t->h[n] = h; t->off[n] = total;
        total += w * h; ++n;
        if (w == 1 && h == 1) break;
        w = (w + 1) / 2; h = (h + 1) / 2;
    } while (1);
    t->nlev = n; t->total = total;
    t->val = (int *)malloc(sizeof(int) * total);
    t->low = (int *)calloc(total, sizeof(int));
    t->known = (int *)calloc(total, sizeof(int));
    for (int i = 0; i < total; ++i) t->val[i] = 0x7fffffff;
}

static void tt_free(tagtree *t) { free(t->val); free(t->low); free(t->known); }

static void tt_set(tagtree *t, int x, int y, int v)
{
    for (int l = 0; l < t->nlev; ++l) {
        int *node = &t->val[t->off[l] + y * t->w[l] + x];
        if (v < *node) *node = v;
        x >>= 1; y >>= 1;
    }
}

/* Codes whether leaf (x,y) has value < threshold, from the root down. */
static void tt_encode(tagtree *t, bitw *b, int x, int y, int threshold)
{
    int idx[20];
    for (int l = 0; l < t->nlev; ++l) { idx[l] = t->off[l] + (y >> l) * t->w[l] + (x >> l); }
    int low = 0;
    for (int l = t->nlev - 1; l >= 0; --l) {
        int n = idx[l];
        if (low > t->low[n]) t->low[n] = low; else low = t->low[n];
        while (low < threshold) {
            if (low >= t->val[n]) {
                if (!t->known[n]) { bw_bit(b, 1); t->known[n] = 1; }
                break;
            }
            bw_bit(b, 0);
            ++low;
        }
        t->low[n] = low;
    }
}

/* ---------------- packet helpers ---------------- */
static void put_npasses(bitw *b, int n)
{
    if (n == 1) bw_bit(b, 0);
    else if (n == 2) bw_bits(b, 2, 2);
    else if (n <= 5) { bw_bits(b, 3, 2); bw_bits(b, (uint32_t)(n - 3), 2); }
    else if (n <= 36) { bw_bits(b, 15, 4); bw_bits(b, (uint32_t)(n - 6), 5); }
    else { bw_bits(b, 0x1ff, 9); bw_bits(b, (uint32_t)(n - 37), 7); }
}

static int floorlog2(int v) { int l = -1; while (v) { v >>= 1; ++l; } return l; }

int orc_band_mb(const orc_j2k_params *p, int band) { return p->guard_bits + p->eps[band] - 1; }

long orc_t2_write(const orc_j2k_params *p, const orc_cb_result *cbres,
                  const long *cboff, const uint8_t *cbdata, const int32_t *passlen,
                  int ncb_per_comp, uint8_t *out, long cap)
{
    return orc_t2_write_layers(p, cbres, cboff, cbdata, passlen, ncb_per_comp, 1, NULL, out, cap);
}

/* Quality layers (B.9-B.10, LRCP): layer_np[blk * nlayers + l] is the number
 * of coding passes block blk has after layer l (non-decreasing in l; NULL:
 * one layer with every block's npasses).  A packet of layer l carries each
 * block's passes beyond the previous layer's.  Inclusion: the tag tree holds
 * the first layer with a pass (nlayers if none) and is coded against l + 1
 * until the block is included, then one bit per layer says whether the
 * block contributes; the zero-bit-plane tag tree is coded at first
 * inclusion; Lblock (3 at first inclusion) persists across layers.  Several
 * layers need the TERMALL style (segment lengths at every pass end). */
long orc_t2_write_layers(const orc_j2k_params *p, const orc_cb_result *cbres,
                         const long *cboff, const uint8_t *cbdata, const int32_t *passlen,
                         int ncb_per_comp, int nlayers, const int32_t *layer_np, uint8_t *out, long cap)
{
    if (nlayers < 1 || (nlayers > 1 && (!layer_np || !passlen || !(p->cblk_style & ORC_STYLE_TERMALL)))) return -1;
    wbuf w = { out, 0, cap, 0 };
    int nbands = 3 * p->levels + 1;
    /* SOC, SIZ */
    put16(&w, 0xFF4F);
    put16(&w, 0xFF51); put16(&w, 38 + 3 * p->ncomp); put16(&w, 0);
    put32(&w, (uint32_t)p->w); put32(&w, (uint32_t)p->h); put32(&w, 0); put32(&w, 0);
    put32(&w, (uint32_t)p->w); put32(&w, (uint32_t)p->h); put32(&w, 0); put32(&w, 0);
    put16(&w, p->ncomp);
    for (int c = 0; c < p->ncomp; ++c) { put8(&w, p->bitdepth - 1); put8(&w, 1); put8(&w, 1); }
    /* COD */
    put16(&w, 0xFF52); put16(&w, 12); put8(&w, 0);
    put8(&w, 0); put16(&w, nlayers); put8(&w, p->mct);
    put8(&w, p->levels); put8(&w, p->cbw_log2 - 2); put8(&w, p->cbh_log2 - 2); put8(&w, p->cblk_style);
    put8(&w, p->reversible ? 1 : 0);
    /* QCD */
    put16(&w, 0xFF5C);
    if (p->reversible) {
        put16(&w, 3 + nbands); put8(&w, (p->guard_bits << 5) | 0);
        for (int k = 0; k < nbands; ++k) put8(&w, p->eps[k] << 3);
    } else {
        put16(&w, 3 + 2 * nbands); put8(&w, (p->guard_bits << 5) | 2);
        for (int k = 0; k < nbands; ++k) put16(&w, (p->eps[k] << 11) | p->mu[k]);
    }
    /* SOT (Psot patched below), SOD */
    long sot = w.pos;
    put16(&w, 0xFF90); put16(&w, 10); put16(&w, 0); put32(&w, 0); put8(&w, 0); put8(&w, 1);
    put16(&w, 0xFF93);

    orc_band bands[3 * 32 + 1];
    orc_bands(p->w, p->h, p->levels, bands);
    /* first code-block index of each band inside a component's list */
    int bstart[3 * 32 + 2], bnx[3 * 32 + 1], bny[3 * 32 + 1];
    int cbw = 1 << p->cbw_log2, cbh = 1 << p->cbh_log2, acc = 0;
    for (int k = 0; k < nbands; ++k) {
        bstart[k] = acc;
        if (bands[k].w <= 0 || bands[k].h <= 0) { bnx[k] = bny[k] = 0; continue; }
        bnx[k] = (bands[k].w + cbw - 1) / cbw; bny[k] = (bands[k].h + cbh - 1) / cbh;
        acc += bnx[k] * bny[k];
    }
    if (acc != ncb_per_comp) return -1;

    const long ncb_all = (long)ncb_per_comp * p->ncomp;
    /* passes of block i after layer l */
#define NP(i, l) (layer_np ? layer_np[(long)(i) * nlayers + (l)] : cbres[i].npasses)
    /* per (component, band) tag trees, per block: passes sent, Lblock */
    tagtree *incl = (tagtree *)calloc((size_t)p->ncomp * nbands, sizeof(tagtree));
    tagtree *imsb = (tagtree *)calloc((size_t)p->ncomp * nbands, sizeof(tagtree));
    int *sent = (int *)calloc((size_t)ncb_all, sizeof(int));
    int *lblk = (int *)malloc(sizeof(int) * (size_t)ncb_all);
    for (long i = 0; i < ncb_all; ++i) lblk[i] = 3;
    for (int c = 0; c < p->ncomp; ++c)
        for (int k = 0; k < nbands; ++k) {
            if (bnx[k] == 0) continue;
            tagtree *ti = &incl[c * nbands + k], *tm = &imsb[c * nbands + k];
            tt_init(ti, bnx[k], bny[k]);
            tt_init(tm, bnx[k], bny[k]);
            const int Mb = orc_band_mb(p, k);
            for (int j = 0; j < bny[k]; ++j)
                for (int i = 0; i < bnx[k]; ++i) {
                    const long blk = (long)c * ncb_per_comp + bstart[k] + j * bnx[k] + i;
                    int first = nlayers;
                    for (int l = nlayers - 1; l >= 0; --l) if (NP(blk, l) > 0) first = l;
                    tt_set(ti, i, j, first);
                    tt_set(tm, i, j, Mb - cbres[blk].nbp);
                }
        }
    for (int l = 0; l < nlayers; ++l)
        for (int r = 0; r <= p->levels; ++r) {
            int b0 = r == 0 ? 0 : 1 + 3 * (r - 1), b1 = r == 0 ? 1 : b0 + 3;
            for (int c = 0; c < p->ncomp; ++c) {
                const long cb0 = (long)c * ncb_per_comp;
                int any = 0;
                for (int k = b0; k < b1; ++k)
                    for (int i = 0; i < bnx[k] * bny[k]; ++i)
                        if (NP(cb0 + bstart[k] + i, l) > sent[cb0 + bstart[k] + i]) any = 1;
                bitw b;
                bw_init(&b, &w);
                if (!any) { bw_bit(&b, 0); bw_flush(&b); continue; }
                bw_bit(&b, 1);
                for (int k = b0; k < b1; ++k) {
                    if (bnx[k] == 0) continue;
                    for (int j = 0; j < bny[k]; ++j)
                        for (int i = 0; i < bnx[k]; ++i) {
                            const long blk = cb0 + bstart[k] + j * bnx[k] + i;
                            const orc_cb_result *q = &cbres[blk];
                            const int prev = sent[blk], now = NP(blk, l), nnew = now - prev;
                            if (prev == 0) {
                                tt_encode(&incl[c * nbands + k], &b, i, j, l + 1);
                                if (nnew <= 0) continue;
                                tt_encode(&imsb[c * nbands + k], &b, i, j, 999);
                            } else {
                                bw_bit(&b, nnew > 0);
                                if (nnew <= 0) continue;
                            }
                            put_npasses(&b, nnew);
                            /* B.10.7: one Lblock increment, then the length of each
                             * codeword segment (every new pass with TERMALL, else one
                             * segment of npasses passes) in Lblock + log2(passes) bits */
                            const int32_t *segs = passlen ? passlen + blk * ORC_MAX_PASSES : NULL;
                            const int termall = (p->cblk_style & ORC_STYLE_TERMALL) && segs;
                            const int nseg = termall ? nnew : 1;
                            const int seg_passes_log2 = termall ? 0 : floorlog2(nnew);
                            int inc = 0;
                            for (int sgi = 0; sgi < nseg; ++sgi) {
                                int len = termall ? segs[prev + sgi] : q->len;
                                int nb = floorlog2(len) + 1, have = lblk[blk] + seg_passes_log2;
                                if (nb - have > inc) inc = nb - have;
                            }
                            for (int t = 0; t < inc; ++t) bw_bit(&b, 1);
                            bw_bit(&b, 0);
                            lblk[blk] += inc;
                            for (int sgi = 0; sgi < nseg; ++sgi)
                                bw_bits(&b, (uint32_t)(termall ? segs[prev + sgi] : q->len),
                                        lblk[blk] + seg_passes_log2);
                        }
                }
                bw_flush(&b);
                /* body: the new passes' bytes, block order */
                for (int k = b0; k < b1; ++k)
                    for (int i = 0; i < bnx[k] * bny[k]; ++i) {
                        const long blk = cb0 + bstart[k] + i;
                        const int prev = sent[blk], now = NP(blk, l);
                        if (now <= prev) continue;
                        long from = 0, to = cbres[blk].len;
                        if (passlen && (p->cblk_style & ORC_STYLE_TERMALL)) {
                            const int32_t *segs = passlen + blk * ORC_MAX_PASSES;
                            to = 0;
                            for (int t = 0; t < now; ++t) { if (t == prev) from = to; to += segs[t]; }
                        }
                        const uint8_t *src = cbdata + cboff[blk];
                        for (long t = from; t < to; ++t) put8(&w, src[t]);
                        sent[blk] = now;
                    }
            }
        }
#undef NP
    for (int c = 0; c < p->ncomp; ++c)
        for (int k = 0; k < nbands; ++k)
            if (bnx[k]) { tt_free(&incl[c * nbands + k]); tt_free(&imsb[c * nbands + k]); }
    free(incl); free(imsb); free(sent); free(lblk);
    long psot = w.pos - sot;
    if (sot + 10 <= cap) {
        out[sot + 6] = (uint8_t)(psot >> 24); out[sot + 7] = (uint8_t)(psot >> 16);
        out[sot + 8] = (uint8_t)(psot >> 8); out[sot + 9] = (uint8_t)psot;
    }
    put16(&w, 0xFFD9);
    return w.err ? -1 : w.pos;
}
