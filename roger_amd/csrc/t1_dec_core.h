// EBCOT tier-1 decoder core (T.800 Annex D passes, MQ decoder C.3, raw
// segments D.6), written against the wavefront primitives of wave_hip.h so
// the kernel in t1_decode.hip and the host lane emulator build the same code.
//
// The reference has no decoder at all (its OCLDecoder only runs the DWT
// stage, src/OCLDecoder.cpp:35-38, and even that is never called).
//
// Design: one wavefront per code-block, lane j owns column j with its state
// as 64-bit row masks (significance, sign, visited, refined, bit-plane) and
// its 64 magnitudes in registers, folded in MSB first one bit-plane at a
// time.  Decoding is sequential in scan order, so the MQ / raw decoders and
// the per-sample decisions of the significance-propagation and cleanup
// passes run as wave-uniform scalar code; the lanes do everything that does
// not depend on a decision of the same pass:
//   * per stripe, a ballot gives the columns that hold a candidate (SPP: an
//     insignificant sample with a significant neighbour; cleanup: an
//     insignificant sample not visited in this bit-plane), and the scalar
//     walk visits only those; a sample that becomes significant in the SPP
//     adds the next column (whose candidates it can create);
//   * a visited column's 6-row windows come from its lane and its neighbours'
//     (readlane), the left one carried over when it was the column just
//     decoded; ZC and SC contexts are looked up in 256-entry tables of 4-bit
//     entries kept in a VGPR (readlane + shift);
//   * MQ refinement passes form their contexts in parallel (no intra-pass
//     dependency) and only decode in order; raw refinement passes (BYPASS)
//     are fully parallel: the segment is de-stuffed into a bit array in
//     workgroup memory and every lane reads its bits at its prefix-sum offset.
// Codeword bytes are read through the scalar cache; the 19 context states and
// the 47-entry probability table sit in VGPR lanes.
#pragma once
#include "t1_core.h"

namespace rg_t1 {

// Workgroup memory of the decoder: the context tables (LUT_BYTES) and the
// de-stuffed raw refinement bits (DEC_RAW_WORDS words: a refinement pass codes
// at most 4096 bits, read from at most 4096/7 + 2 bytes).
constexpr int DEC_RAW_WORDS = 160;

struct DecJob {
    const uint8_t *data;    // codeword (wave-uniform address)
    int len, npasses, nbp, w, h, orient, pitch;
    int32_t *dst;
    int style;              // code-block style bits (BYPASS / RESET / TERMALL)
    const int32_t *seglen;  // TERMALL: codeword segment length of each pass
};

struct MQDec {
    // MQ: C.3's C register in bits 63..32 (the compared half in 63..48) with
    // the codeword bits already read in below it, zero below bit E.  Raw: the
    // current byte.
    uint64_t C;
    uint32_t A;
    int E;            // MQ: lowest bit holding codeword data
    int CT;           // raw: unread bits of the current byte
    int bp, len;      // next byte to read; segment length
    int ff, marker;   // MQ: the last byte read was 0xFF / a marker was reached
    const uint8_t *buf;
};

// Bytes past the segment end read as 0xFF (C.3.4).
RG_DEV uint32_t mqd_byte(const MQDec &m, int i) { return i < m.len ? wv_sload_u8(m.buf + i) : 0xFFu; }

// Reads codeword bytes in below bit E while E > 32, so a renormalisation's
// (at most 15) shifts always find bits, each byte where BYTEIN (C.3.4) puts
// it: 8 bits below the previous byte, or 7 after an 0xFF (its first bit a
// carry into the 0xFF).  A byte after an 0xFF is added only once the 0xFF's
// last bit has passed bit 48, which is when BYTEIN adds it, since its carry
// can reach the compared bits; other bytes land on zero bits, so adding them
// early changes nothing that DECODE compares.  After a marker (0xFF, then a
// byte above 0x8F) the reader supplies 0xFF bytes for ever.
RG_DEV void mqd_fill(MQDec &m)
{
    while (m.E > 32) {
        if (m.marker) {
            m.C += 0xFFull << (m.E - 8);
            m.E -= 8;
            continue;
        }
        if (m.ff) {
            if (m.E <= 48) break;
            const uint32_t b1 = mqd_byte(m, m.bp);
            if (b1 > 0x8F) {
                m.marker = 1;
                continue;
            }
            m.bp++;
            m.C += (uint64_t)b1 << (m.E - 7);
            m.E -= 7;
            m.ff = 0;
            continue;
        }
        const uint32_t b = mqd_byte(m, m.bp);
        m.bp++;
        m.C += (uint64_t)b << (m.E - 8);
        m.E -= 8;
        m.ff = b == 0xFFu;
    }
}

// INITDEC (C.3.5) on the segment [p, p + len): the first byte, BYTEIN and
// the shift by 7.
RG_DEV void mqd_start(MQDec &m, const uint8_t *p, int len)
{
    m.buf = p; m.len = len;
    const uint32_t b0 = mqd_byte(m, 0);
    m.bp = 1; m.ff = b0 == 0xFFu; m.marker = 0;
    m.C = (uint64_t)b0 << 55;
    m.E = 55;
    mqd_fill(m);
    m.A = 0x8000;
}

// DECODE (C.3.2) of one decision in context cx.  Lane c of ctxv holds
// context c's state as its probability-table entry (Qe | NMPS << 16 | NLPS
// << 22 | SWITCH << 28) with the MPS in bit 31, so a decision reads its state
// with one readlane; lane k of mqtv holds table entry k, read only when the
// state changes.  The exchanges are written as selects: the decision is the
// LPS when the lower sub-interval is kept with A - Qe >= Qe or the upper one
// with A - Qe < Qe.  RENORMD (C.3.3) shifts A and C in one go.
RG_DEV int mqd_decode(MQDec &m, int &ctxv, int mqtv, int cx)
{
    const uint32_t e = (uint32_t)wv_readlane(ctxv, cx);
    const uint32_t mps = e >> 31;
    const uint32_t qe = e & 0xFFFFu;
    const uint32_t a = m.A - qe;
    const uint32_t chigh = (uint32_t)(m.C >> 48);
    if (chigh >= qe) {
        m.C -= (uint64_t)qe << 48;
        if (a & 0x8000u) {
            m.A = a;
            return (int)mps;
        }
    }
    const uint32_t lower = chigh < qe ? 1u : 0u;
    const uint32_t lps = lower ^ (a < qe ? 1u : 0u);
    m.A = lower ? qe : a;
    const uint32_t idx = (e >> (lps ? 22 : 16)) & 63u;
    const uint32_t nm = mps ^ (lps & (e >> 28) & 1u);
    ctxv = wv_writelane(ctxv, cx, (int)((uint32_t)wv_readlane(mqtv, (int)idx) | (nm << 31)));
    const int n = wv_clz(m.A) - 16;
    m.A <<= n;
    m.C <<= n;
    m.E += n;
    if (m.E > 32) mqd_fill(m);
    return (int)(mps ^ lps);
}

// Raw segment reader (D.6): bits MSB first, 7 in the byte after an 0xFF,
// 0xFF bytes past the segment end.  Reuses MQDec: C holds the current byte,
// CT its unread bits.
RG_DEV void raw_start(MQDec &m, const uint8_t *p, int len)
{
    m.buf = p; m.len = len; m.bp = 0; m.C = 0; m.CT = 0; m.A = 0;
}

RG_DEV int raw_bit(MQDec &m)
{
    if (m.CT == 0) {
        const bool after_ff = m.C == 0xFFu;
        m.C = mqd_byte(m, m.bp);
        m.bp++;
        m.CT = after_ff ? 7 : 8;
    }
    m.CT--;
    return (int)((m.C >> m.CT) & 1u);
}

// The decoder's context tables, 4 bits an entry, 8 entries per lane: entry
// (lane l, nibble n) of a table held in one VGPR.  The fields are placed so
// that the lane and the nibble come out of the windows with few scalar
// operations (windows: bit 0 = the row above the stripe; row i):
//   ZC: lane = right column rows i-1..i+1 (bits 0-2), own rows i-1 (bit 3)
//       and i+1 (bit 5); nibble = left column rows i-1..i+1; entry = context;
//   SC: lane = own row i-1 significant / negative (bits 0 / 1), row i+1
//       (bits 2 / 3), left row i significant (bit 4); nibble = left row i
//       negative, right row i significant, right row i negative; entry =
//       (context - 9) << 1 | XOR bit.
RG_DEV uint32_t lut_nib(int v, uint32_t l, uint32_t n)
{
    return ((uint32_t)wv_readlane(v, (int)l) >> (4 * n)) & 0xFu;
}

// This lane's words of the two tables, from t1_lut_fill's byte tables.
RG_DEV void dec_lut_pack(const uint8_t *lut, int &zcn, int &scn)
{
    const uint32_t l = (uint32_t)wv_lane();
    uint32_t z = 0, c = 0;
#pragma unroll
    for (uint32_t n = 0; n < 8; ++n) {
        const uint32_t zx = n | ((l & 7u) << 3) | (((l >> 3) & 1u) << 6) | (((l >> 5) & 1u) << 7);
        z |= (uint32_t)lut[zx] << (4 * n);
        const uint32_t sx = ((l >> 4) & 1u) | ((n & 1u) << 1) | (((n >> 1) & 1u) << 2) | (((n >> 2) & 1u) << 3) |
                            ((l & 1u) << 4) | (((l >> 1) & 1u) << 5) | (((l >> 2) & 1u) << 6) | (((l >> 3) & 1u) << 7);
        const uint32_t sc = lut[256 + sx];
        c |= (((((sc >> 1) - 9u) & 7u) << 1) | (sc & 1u)) << (4 * n);
    }
    zcn = (int)z;
    scn = (int)c;
}

// ZC context of window row i of a column.
RG_DEV int dec_zc(int zcn, int i, uint32_t Lw, uint32_t Cw, uint32_t Rw)
{
    return (int)lut_nib(zcn, ((Rw >> (i - 1)) & 7u) | (((Cw >> (i - 1)) & 5u) << 3), (Lw >> (i - 1)) & 7u);
}

// Sign of window row i: MQ-decoded with its SC context and XOR bit.
RG_DEV int dec_sign(MQDec &m, int &ctxv, int mqtv, int scn, int i, uint32_t Lw, uint32_t Lneg, uint32_t Cw,
                    uint32_t Cneg, uint32_t Rw, uint32_t Rneg)
{
    const uint32_t l = ((Cw >> (i - 1)) & 5u) | (((Cneg >> (i - 1)) & 5u) << 1) | (b1(Lw, i) << 4);
    const uint32_t n = b1(Lneg, i) | (b1(Rw, i) << 1) | (b1(Rneg, i) << 2);
    const uint32_t sc = lut_nib(scn, l, n);
    return mqd_decode(m, ctxv, mqtv, 9 + (int)(sc >> 1)) ^ (int)(sc & 1u);
}

// Whole raw refinement pass: segment [p, p + n); every significant, not yet
// visited sample of the block takes the next bit in scan order.
RG_DEV void dec_raw_mrp(const uint8_t *p, int n, uint32_t *rawbuf, bool col_ok, uint64_t rowmask, int nstripes,
                        uint64_t S, uint64_t V, uint64_t &R, uint64_t &B)
{
    const int lane = wv_lane();
    const uint64_t Mall = col_ok ? (S & ~V & rowmask) : 0;
    const int need = wv_readlane(wv_incl_scan(__builtin_popcountll(Mall)), 63);
    if (need == 0) return;
    // de-stuff enough bytes for `need` bits (each gives 7 or 8) into rawbuf,
    // MSB first: bit q at word q >> 5, bit 31 - (q & 31)
    const int nbytes = (need + 6) / 7 + 1;
    const int nwords = (nbytes * 8 + 31) / 32 + 1;
    for (int i = lane; i < nwords; i += 64) rawbuf[i] = 0;
    wv_sync();
    int carry = 0;
    uint32_t prevlast = 0;   // the byte before the first: not 0xFF
    for (int b0 = 0; b0 < nbytes; b0 += 64) {
        const int b = b0 + lane;
        const uint32_t byte = b < n ? (uint32_t)p[b] : 0xFFu;
        uint32_t prev = wv_from_left(byte);
        prev = lane == 0 ? prevlast : prev;
        const int nb = b < nbytes ? (prev == 0xFFu ? 7 : 8) : 0;
        const uint32_t val = prev == 0xFFu ? (byte & 0x7Fu) : byte;
        const int incl = wv_incl_scan(nb);
        if (nb) {
            const int pos = carry + incl - nb;
            const uint32_t al = val << (32 - nb);
            const int wi = pos >> 5, sh = pos & 31;
            wv_lds_or(&rawbuf[wi], al >> sh);
            if (sh + nb > 32) wv_lds_or(&rawbuf[wi + 1], al << (32 - sh));
        }
        carry += wv_readlane(incl, 63);
        prevlast = (uint32_t)wv_readlane((int)byte, 63);
    }
    wv_sync();
    int base = 0;
    for (int s = 0; s < nstripes; ++s) {
        const int r0 = 4 * s;
        const uint32_t M4 = (uint32_t)(Mall >> r0) & 0xFu;
        const int mc = __builtin_popcount(M4);
        const int incl = wv_incl_scan(mc);
        // branch-free: off <= need, so both words lie inside the de-stuffed range
        const int off = base + incl - mc;
        const int wi = off >> 5, sh = off & 31;
        uint32_t x = (rawbuf[wi] << sh) | (sh ? rawbuf[wi + 1] >> (32 - sh) : 0u);
        uint32_t bits = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t mk = (M4 >> k) & 1u;
            bits |= ((x >> 31) & mk) << k;
            x <<= mk;
        }
        B |= (uint64_t)bits << r0;
        base += wv_readlane(incl, 63);
    }
    R |= Mall;
    wv_sync();
}

// Decodes one code-block into job.dst.  lut: LUT_BYTES, rawbuf: DEC_RAW_WORDS
// words of workgroup memory.
RG_DEV void t1_dec_block(const DecJob &job, uint8_t *lut, uint32_t *rawbuf)
{
    const int lane = wv_lane();
    const int w = job.w, h = job.h, nbp = job.nbp;
    t1_lut_fill(lut, job.orient);
    int zcv, scv;
    dec_lut_pack(lut, zcv, scv);

    uint32_t mag[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) mag[r] = 0;
    uint64_t S = 0, NEG = 0, V = 0, R = 0;
    const bool col_ok = lane < w;
    const uint64_t rowmask = h >= 64 ? ~0ull : ((1ull << h) - 1);
    const int nstripes = (h + 3) >> 2;
    // initial states (Table D.7) as table entries, MPS 0
    const int ctx0 = (int)MQT[lane == 0 ? 4 : lane == CTX_RL ? 3 : lane == CTX_UNI ? 46 : 0];
    int ctxv = ctx0;
    const int mqtv = lane < 47 ? (int)MQT[lane] : 0;
    const bool termall = (job.style & STYLE_TERMALL) && job.seglen;
    const int len = job.len;

    MQDec m;
    int seg_off = 0;
    mqd_start(m, job.data, termall ? (job.seglen[0] < len ? job.seglen[0] : len) : len);

    int passes = 0;
    for (int p = nbp - 1; p >= 0; --p) {
        uint64_t B = 0;   // bit p of this lane's column
        for (int pass = (p == nbp - 1) ? 2 : 0; pass < 3 && passes < job.npasses; ++pass, ++passes) {
            const bool raw = termall && pass < 2 && pass_is_raw(job.style, passes);
            int so = 0, sn = 0;   // raw segment
            if (passes > 0) {
                if (job.style & STYLE_RESET) ctxv = ctx0;
                if (termall) {   // next codeword segment
                    seg_off += job.seglen[passes - 1];
                    const int rest = len - seg_off;
                    const int sl = job.seglen[passes];
                    so = seg_off < len ? seg_off : len;
                    sn = rest <= 0 ? 0 : (sl < rest ? sl : rest);
                    if (raw) raw_start(m, job.data + so, sn);
                    else mqd_start(m, job.data + so, sn);
                }
            }
            if (raw && pass == 1) {
                dec_raw_mrp(job.data + so, sn, rawbuf, col_ok, rowmask, nstripes, S, V, R, B);
                continue;
            }
            for (int s = 0; s < nstripes; ++s) {
                const int r0 = 4 * s;
                const uint32_t vrow = (uint32_t)(rowmask >> r0) & 0xFu;
                const uint32_t swv = win6(S, r0), negv = win6(NEG, r0);
                const uint32_t lw = wv_from_left(swv), rw = wv_from_right(swv);
                if (pass == 1) {
                    // ---- MQ magnitude refinement: contexts in parallel, bits in order ----
                    const uint32_t vw = (uint32_t)(V >> r0) & 0xFu, rf = (uint32_t)(R >> r0) & 0xFu;
                    const uint32_t M = col_ok ? ((swv >> 1) & ~vw & vrow) : 0u;
                    uint32_t cxs = 0;   // 4 contexts x 5 bits
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int i = k + 1;
                        const uint32_t nb = b1(swv, i - 1) | b1(swv, i + 1) | ((lw >> (i - 1)) & 7u) |
                                            ((rw >> (i - 1)) & 7u);
                        const uint32_t cx = b1(rf, k) ? 16u : (nb ? 15u : 14u);
                        cxs |= cx << (5 * k);
                    }
                    uint64_t mask = wv_ballot(M != 0);
                    while (mask) {
                        const int j = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const uint32_t Mj = (uint32_t)wv_readlane((int)M, j);
                        const uint32_t cj = (uint32_t)wv_readlane((int)cxs, j);
                        uint32_t bits = 0;
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (b1(Mj, k) && mqd_decode(m, ctxv, mqtv, (int)((cj >> (5 * k)) & 31u)))
                                bits |= 1u << k;
                        if (lane == j) B |= (uint64_t)bits << r0;
                    }
                    R |= (uint64_t)M << r0;
                    continue;
                }
                // ---- significance propagation (pass 0) / cleanup (pass 2): columns with candidates ----
                const uint32_t vwv = (uint32_t)(V >> r0) & 0xFu;
                uint32_t cand;
                if (pass == 0) {
                    const uint32_t nbr = ((swv | (swv >> 2)) | (lw | (lw >> 1) | (lw >> 2)) | (rw | (rw >> 1) | (rw >> 2)));
                    cand = ~(swv >> 1) & nbr & vrow;
                } else {
                    cand = ~(swv >> 1) & ~vwv & vrow;
                }
                cand = col_ok ? cand & 0xFu : 0u;
                uint64_t cols = wv_ballot(cand != 0);
                const bool full = r0 + 4 <= h;
                // Cleanup: columns in run-length mode (D.3.4: four insignificant,
                // unvisited rows and no significant sample in the 3 x 6
                // neighbourhood) as of the stripe's start.  During the walk only
                // the left neighbour of a column can change, and only when that
                // column has just become significant, so a column of this mask
                // is still in run-length mode unless column j - 1 was the last
                // one to gain a significant sample (sigj), and no other column
                // enters the mode.
                const uint64_t rlm =
                    pass == 2 && full ? wv_ballot(cand == 0xFu && ((swv | lw | rw) & 0x3Fu) == 0u) : 0ull;
                int lastj = -2, sigj = -2;
                uint32_t Lw = 0, Lneg = 0;
                while (cols) {
                    const int j = __builtin_ctzll(cols);
                    cols &= cols - 1;
                    int rl = 0;   // run-length decision 1: a significant row follows
                    if (((rlm >> j) & 1u) && sigj != j - 1) {
                        // the run-length decision, before any neighbourhood read:
                        // a 0 leaves the column as it is (no window of it changes)
                        rl = mqd_decode(m, ctxv, mqtv, CTX_RL);
                        if (!rl) {
                            lastj = -2;
                            continue;
                        }
                    }
                    uint32_t Cw = (uint32_t)wv_readlane((int)swv, j);
                    uint32_t Cneg = (uint32_t)wv_readlane((int)negv, j);
                    if (j != lastj + 1) {   // else: carried from the column just decoded
                        Lw = j > 0 ? (uint32_t)wv_readlane((int)swv, j - 1) : 0u;
                        Lneg = j > 0 ? (uint32_t)wv_readlane((int)negv, j - 1) : 0u;
                    }
                    uint32_t Rw = 0, Rneg = 0;
                    if (j + 1 < w) {
                        Rw = (uint32_t)wv_readlane((int)swv, j + 1);
                        Rneg = (uint32_t)wv_readlane((int)negv, j + 1);
                    }
                    uint32_t vis = 0, nsig = 0;
                    if (pass == 0) {
                        // window rows 1-4 in the pass: valid, insignificant, with a
                        // significant neighbour (all four at once); a row that
                        // becomes significant adds the row below it
                        const uint32_t LR = Lw | Rw, vw = vrow << 1;
                        uint32_t candm = (LR | (LR >> 1) | (LR << 1) | (Cw >> 1) | (Cw << 1)) & ~Cw & vw;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const int i = k + 1;
                            if (!b1(candm, i)) continue;
                            vis |= 1u << k;
                            const int bit = raw ? raw_bit(m) : mqd_decode(m, ctxv, mqtv, dec_zc(zcv, i, Lw, Cw, Rw));
                            if (bit) {
                                const int neg = raw ? raw_bit(m)
                                                    : dec_sign(m, ctxv, mqtv, scv, i, Lw, Lneg, Cw, Cneg, Rw, Rneg);
                                Cw |= 1u << i;
                                Cneg |= (uint32_t)neg << i;
                                nsig |= 1u << k;
                                candm |= (2u << i) & ~Cw & vw;
                            }
                        }
                        // a new significant sample can make the next column a candidate
                        if (nsig && j + 1 < w) cols |= 1ull << (j + 1);
                    } else {
                        const uint32_t cj = (uint32_t)wv_readlane((int)cand, j);
                        int k0 = 0;
                        if (rl) {
                            // run-length mode, decision 1: two UNIFORM decisions give
                            // the first significant row, then its sign
                            int k = mqd_decode(m, ctxv, mqtv, CTX_UNI) << 1;
                            k |= mqd_decode(m, ctxv, mqtv, CTX_UNI);
                            const int neg = dec_sign(m, ctxv, mqtv, scv, k + 1, Lw, Lneg, Cw, Cneg, Rw, Rneg);
                            Cw |= 1u << (k + 1);
                            Cneg |= (uint32_t)neg << (k + 1);
                            nsig |= 1u << k;
                            k0 = k + 1;
                        }
                        for (int k = k0; k < 4; ++k) {
                            if (!b1(cj, k)) continue;
                            const int i = k + 1;
                            if (mqd_decode(m, ctxv, mqtv, dec_zc(zcv, i, Lw, Cw, Rw))) {
                                const int neg = dec_sign(m, ctxv, mqtv, scv, i, Lw, Lneg, Cw, Cneg, Rw, Rneg);
                                Cw |= 1u << i;
                                Cneg |= (uint32_t)neg << i;
                                nsig |= 1u << k;
                            }
                        }
                    }
                    if (nsig) sigj = j;
                    if (vis | nsig) {
                        if (lane == j) {
                            V |= (uint64_t)vis << r0;
                            S |= (uint64_t)nsig << r0;
                            NEG |= (uint64_t)((Cneg >> 1) & nsig) << r0;
                            B |= (uint64_t)nsig << r0;
                        }
                    }
                    lastj = j;
                    Lw = Cw;
                    Lneg = Cneg;
                }
            }
            if (pass == 2) V = 0;
        }
        // fold bit-plane p into the magnitudes (MSB first)
        const uint32_t lo = (uint32_t)B, hi = (uint32_t)(B >> 32);
#pragma unroll
        for (int r = 0; r < 32; ++r) mag[r] = (mag[r] << 1) | ((lo >> r) & 1u);
#pragma unroll
        for (int r = 0; r < 32; ++r) mag[r + 32] = (mag[r + 32] << 1) | ((hi >> r) & 1u);
    }
    if (col_ok) {
#pragma unroll
        for (int r = 0; r < 64; ++r) {
            if (r < h) {
                const int v = (int)mag[r];
                job.dst[(size_t)r * job.pitch + lane] = ((NEG >> r) & 1) ? -v : v;
            }
        }
    }
}

}  // namespace rg_t1
