// EBCOT tier-1 decoder, one code-block per LANE (T.800 Annex D passes, MQ
// decoder C.3, raw segments D.6).
//
// The wavefront decoder (t1_dec_core.h) runs a block's decisions as
// wave-uniform scalar code: every workgroup of a CU shares one scalar unit,
// which bounds it (six scalar instructions per vector one).  Here each lane
// decodes a block of its own, entirely in vector registers: the MQ / raw
// decoders, the context formation and the scan are ordinary per-lane code,
// so the blocks of a wavefront advance together on the vector unit and the
// four SIMDs of a CU each run their own wavefront.  The block's state lives
// in workgroup memory (the Mem policy): per column, 64-bit row masks of
// significance (S), sign (N), visited-in-this-bit-plane (V), refined-once (R)
// and the current bit-plane's bits (B); per block, the 19 context states.
// Each finished bit-plane's bits go to `planes` (64 columns per plane) and
// the signs to `negs`; a lane-per-column epilogue folds them into the
// coefficients.  The same source builds for the host (plain C++ policies)
// so the decoder is checked against the oracle on the CPU.
//
// Scan, contexts and segment rules follow t1_dec_core.h exactly: stripes of
// four rows, columns left to right, rows top to bottom; the left column's
// windows are those just decoded, the right column's those before this
// pass; significance propagation, refinement and cleanup (with run-length
// mode) per bit-plane; RESET / TERMALL / BYPASS segments.
#pragma once
#include "t1_core.h"

namespace rg_t1 {

// Entry x of the block's context tables (the layout t1_lut_fill builds):
// x < 256 the ZC context of a neighbourhood pattern, 256 + x the SC context
// << 1 | XOR bit of a sign pattern.
RG_DEV uint32_t lut_entry(int orient, int x)
{
    if (x < 256) {
        const int h = ((x >> 1) & 1) + ((x >> 4) & 1);
        const int v = ((x >> 6) & 1) + ((x >> 7) & 1);
        const int d = (x & 1) + ((x >> 2) & 1) + ((x >> 3) & 1) + ((x >> 5) & 1);
        return (uint32_t)zc_ctx(orient, h, v, d);
    }
    x -= 256;
    int c[4];
    for (int k = 0; k < 4; ++k) c[k] = ((x >> (2 * k)) & 1) ? (((x >> (2 * k + 1)) & 1) ? -1 : 1) : 0;
    int H = c[0] + c[1], V = c[2] + c[3];
    H = H > 1 ? 1 : (H < -1 ? -1 : H);
    V = V > 1 ? 1 : (V < -1 ? -1 : V);
    return (uint32_t)sc_ctx(H, V);
}

struct LaneJob {
    int len, npasses, nbp, w, h, orient;
    int style;              // code-block style bits (BYPASS / RESET / TERMALL)
    const int32_t *seglen;  // TERMALL: codeword segment length of each pass
    uint64_t *planes;       // out: bits of plane p, column c at planes[p * 64 + c] (p < nbp)
    uint64_t *negs;         // out: sign masks of the 64 columns
};

// MQ / raw decoder state of one lane; bytes come from Rd::get(absolute
// offset into the block's codeword), past the segment end they read 0xFF.
struct LaneMQ {
    uint32_t A, C;
    int CT, bp, len, so;
};

template <class Rd>
RG_DEV uint32_t lmq_byte(Rd &rd, const LaneMQ &m, int i)
{
    return i < m.len ? rd.get(m.so + i) : 0xFFu;
}

template <class Rd>
RG_DEV void lmq_bytein(Rd &rd, LaneMQ &m)
{
    const uint32_t b = lmq_byte(rd, m, m.bp), b1n = lmq_byte(rd, m, m.bp + 1);
    if (b == 0xFF) {
        if (b1n > 0x8F) { m.C += 0xFF00; m.CT = 8; }
        else { m.bp++; m.C += b1n << 9; m.CT = 7; }
    } else {
        m.bp++; m.C += b1n << 8; m.CT = 8;
    }
}

template <class Rd>
RG_DEV void lmq_start(Rd &rd, LaneMQ &m, int so, int len)
{
    m.so = so; m.len = len; m.bp = 0;
    m.C = lmq_byte(rd, m, 0) << 16;
    lmq_bytein(rd, m);
    m.C <<= 7; m.CT -= 7; m.A = 0x8000;
}

// DECODE (C.3.2) in context cx; the context's state (index << 1 | MPS) is
// mem.ctx(cx), table entries mem.mqt(k) (Qe | NMPS << 16 | NLPS << 22 |
// SWITCH << 28).  RENORMD shifts in one go up to the next BYTEIN.
template <class Mem, class Rd>
RG_DEV int lmq_decode(Mem &mem, Rd &rd, LaneMQ &m, int cx)
{
    const int st = mem.ctx(cx);
    const int mps = st & 1;
    const uint32_t t = mem.mqt(st >> 1);
    const uint32_t qe = t & 0xFFFFu;
    int d, ns;
    m.A -= qe;
    if ((m.C >> 16) < qe) {
        const bool cond = m.A < qe;   // LPS exchange: the MPS if A < Qe
        m.A = qe;
        d = cond ? mps : 1 - mps;
        ns = cond ? (int)(((t >> 16) & 63u) << 1) | mps : (int)(((t >> 22) & 63u) << 1) | (mps ^ (int)(t >> 28));
    } else {
        m.C -= qe << 16;
        if (m.A & 0x8000u) return mps;
        const bool cond = m.A < qe;   // MPS exchange: the LPS if A < Qe
        d = cond ? 1 - mps : mps;
        ns = cond ? (int)(((t >> 22) & 63u) << 1) | (mps ^ (int)(t >> 28)) : (int)(((t >> 16) & 63u) << 1) | mps;
    }
    mem.set_ctx(cx, ns);
    int n = wv_clz(m.A) - 16;
    m.A <<= n;
    while (n > 0) {
        if (m.CT == 0) lmq_bytein(rd, m);
        const int s = n < m.CT ? n : m.CT;
        m.C <<= s; m.CT -= s; n -= s;
    }
    return d;
}

// Raw segment (D.6): bits MSB first, 7 in the byte after an 0xFF.  C holds
// the current byte, CT its unread bits.
template <class Rd>
RG_DEV void lraw_start(LaneMQ &m, int so, int len)
{
    m.so = so; m.len = len; m.bp = 0; m.C = 0; m.CT = 0; m.A = 0;
}

template <class Rd>
RG_DEV int lraw_bit(Rd &rd, LaneMQ &m)
{
    if (m.CT == 0) {
        const bool after_ff = m.C == 0xFFu;
        m.C = lmq_byte(rd, m, m.bp);
        m.bp++;
        m.CT = after_ff ? 7 : 8;
    }
    m.CT--;
    return (int)((m.C >> m.CT) & 1u);
}

template <class Mem>
RG_DEV void lctx_reset(Mem &mem)
{
    for (int k = 0; k < 19; ++k) mem.set_ctx(k, 0);
    mem.set_ctx(0, 4 << 1);
    mem.set_ctx(CTX_RL, 3 << 1);
    mem.set_ctx(CTX_UNI, 46 << 1);
}

// Decodes one block.  mem: the lane's state (S, N, V, R, B column masks,
// contexts) and the shared tables (lut(x): ZC / SC context tables of the
// block's orientation, t1_lut_fill layout; mqt(k)); rd: the block's codeword.
template <class Mem, class Rd>
RG_DEV void t1_dec_lane(const LaneJob &job, Mem &mem, Rd &rd)
{
    const int w = job.w, h = job.h, nbp = job.nbp, len = job.len;
    const uint64_t rowmask = h >= 64 ? ~0ull : ((1ull << h) - 1);
    const int nstripes = (h + 3) >> 2;
    for (int c = 0; c < 64; ++c) {
        mem.S(c) = 0; mem.N(c) = 0; mem.V(c) = 0; mem.R(c) = 0;
    }
    lctx_reset(mem);
    const bool termall = (job.style & STYLE_TERMALL) && job.seglen;
    LaneMQ m;
    int seg_off = 0;
    if (nbp > 0 && job.npasses > 0) lmq_start(rd, m, 0, termall ? (job.seglen[0] < len ? job.seglen[0] : len) : len);

    int passes = 0;
    for (int p = nbp - 1; p >= 0; --p) {
        for (int c = 0; c < w; ++c) mem.B(c) = 0;
        for (int pass = (p == nbp - 1) ? 2 : 0; pass < 3 && passes < job.npasses; ++pass, ++passes) {
            const bool raw = termall && pass < 2 && pass_is_raw(job.style, passes);
            if (passes > 0) {
                if (job.style & STYLE_RESET) lctx_reset(mem);
                if (termall) {   // next codeword segment
                    seg_off += job.seglen[passes - 1];
                    const int rest = len - seg_off;
                    const int sl = job.seglen[passes];
                    const int so = seg_off < len ? seg_off : len;
                    const int sn = rest <= 0 ? 0 : (sl < rest ? sl : rest);
                    if (raw) lraw_start<Rd>(m, so, sn);
                    else lmq_start(rd, m, so, sn);
                }
            }
            for (int s = 0; s < nstripes; ++s) {
                const int r0 = 4 * s;
                const uint32_t vrow = (uint32_t)(rowmask >> r0) & 0xFu;
                uint32_t Lw = 0, Lneg = 0;   // the column just decoded (updated windows)
                uint32_t Cw = win6(mem.S(0), r0), Cneg = win6(mem.N(0), r0);
                for (int c = 0; c < w; ++c) {
                    uint32_t Rw = 0, Rneg = 0;
                    if (c + 1 < w) {
                        Rw = win6(mem.S(c + 1), r0);
                        Rneg = win6(mem.N(c + 1), r0);
                    }
                    if (pass == 1) {
                        // ---- magnitude refinement ----
                        const uint32_t vw = (uint32_t)(mem.V(c) >> r0) & 0xFu;
                        const uint32_t M = (Cw >> 1) & ~vw & vrow;
                        if (M) {
                            const uint32_t rf = (uint32_t)(mem.R(c) >> r0) & 0xFu;
                            uint32_t bits = 0;
                            for (int k = 0; k < 4; ++k) {
                                if (!b1(M, k)) continue;
                                int bit;
                                if (raw) {
                                    bit = lraw_bit(rd, m);
                                } else {
                                    const int i = k + 1;
                                    const uint32_t nb = b1(Cw, i - 1) | b1(Cw, i + 1) | ((Lw >> (i - 1)) & 7u) |
                                                        ((Rw >> (i - 1)) & 7u);
                                    bit = lmq_decode(mem, rd, m, b1(rf, k) ? 16 : (nb ? 15 : 14));
                                }
                                bits |= (uint32_t)bit << k;
                            }
                            mem.B(c) |= (uint64_t)bits << r0;
                            mem.R(c) |= (uint64_t)M << r0;
                        }
                    } else {
                        // ---- significance propagation (0) / cleanup (2) ----
                        uint32_t vis = 0, nsig = 0;
                        if (pass == 0) {
                            for (int k = 0; k < 4; ++k) {
                                const int i = k + 1;
                                if (!b1(vrow, k) || b1(Cw, i)) continue;
                                const uint32_t nb = b1(Cw, i - 1) | b1(Cw, i + 1) | ((Lw >> (i - 1)) & 7u) |
                                                    ((Rw >> (i - 1)) & 7u);
                                if (!nb) continue;
                                vis |= 1u << k;
                                const int bit = raw ? lraw_bit(rd, m)
                                                    : lmq_decode(mem, rd, m,
                                                                 mem.lut(((Lw >> (i - 1)) & 7u) |
                                                                         (((Rw >> (i - 1)) & 7u) << 3) |
                                                                         (b1(Cw, i - 1) << 6) | (b1(Cw, i + 1) << 7)));
                                if (bit) {
                                    int neg;
                                    if (raw) {
                                        neg = lraw_bit(rd, m);
                                    } else {
                                        const uint32_t x = b1(Lw, i) | (b1(Lneg, i) << 1) | (b1(Rw, i) << 2) |
                                                           (b1(Rneg, i) << 3) | (b1(Cw, i - 1) << 4) |
                                                           (b1(Cneg, i - 1) << 5) | (b1(Cw, i + 1) << 6) |
                                                           (b1(Cneg, i + 1) << 7);
                                        const uint32_t sc = mem.lut(256 + x);
                                        neg = lmq_decode(mem, rd, m, (int)(sc >> 1)) ^ (int)(sc & 1u);
                                    }
                                    Cw |= 1u << i;
                                    Cneg |= (uint32_t)neg << i;
                                    nsig |= 1u << k;
                                }
                            }
                        } else {
                            const uint32_t vwv = (uint32_t)(mem.V(c) >> r0) & 0xFu;
                            const uint32_t cj = ~(Cw >> 1) & ~vwv & vrow & 0xFu;
                            if (cj) {
                                int k0 = 0;
                                bool done = false;
                                if (r0 + 4 <= h && cj == 0xFu && ((Lw | Rw) & 0x3Fu) == 0 && (Cw & 0x21u) == 0) {
                                    // run-length mode (D.3.4)
                                    if (!lmq_decode(mem, rd, m, CTX_RL)) {
                                        done = true;
                                    } else {
                                        int k = lmq_decode(mem, rd, m, CTX_UNI) << 1;
                                        k |= lmq_decode(mem, rd, m, CTX_UNI);
                                        const int i = k + 1;
                                        const uint32_t x = b1(Lw, i) | (b1(Lneg, i) << 1) | (b1(Rw, i) << 2) |
                                                           (b1(Rneg, i) << 3) | (b1(Cw, i - 1) << 4) |
                                                           (b1(Cneg, i - 1) << 5) | (b1(Cw, i + 1) << 6) |
                                                           (b1(Cneg, i + 1) << 7);
                                        const uint32_t sc = mem.lut(256 + x);
                                        const int neg = lmq_decode(mem, rd, m, (int)(sc >> 1)) ^ (int)(sc & 1u);
                                        Cw |= 1u << i;
                                        Cneg |= (uint32_t)neg << i;
                                        nsig |= 1u << k;
                                        k0 = k + 1;
                                    }
                                }
                                if (!done) {
                                    for (int k = k0; k < 4; ++k) {
                                        if (!b1(cj, k)) continue;
                                        const int i = k + 1;
                                        const int zc = mem.lut(((Lw >> (i - 1)) & 7u) | (((Rw >> (i - 1)) & 7u) << 3) |
                                                               (b1(Cw, i - 1) << 6) | (b1(Cw, i + 1) << 7));
                                        if (lmq_decode(mem, rd, m, zc)) {
                                            const uint32_t x = b1(Lw, i) | (b1(Lneg, i) << 1) | (b1(Rw, i) << 2) |
                                                               (b1(Rneg, i) << 3) | (b1(Cw, i - 1) << 4) |
                                                               (b1(Cneg, i - 1) << 5) | (b1(Cw, i + 1) << 6) |
                                                               (b1(Cneg, i + 1) << 7);
                                            const uint32_t sc = mem.lut(256 + x);
                                            const int neg = lmq_decode(mem, rd, m, (int)(sc >> 1)) ^ (int)(sc & 1u);
                                            Cw |= 1u << i;
                                            Cneg |= (uint32_t)neg << i;
                                            nsig |= 1u << k;
                                        }
                                    }
                                }
                            }
                        }
                        if (vis | nsig) {
                            mem.V(c) |= (uint64_t)vis << r0;
                            mem.S(c) |= (uint64_t)nsig << r0;
                            mem.N(c) |= (uint64_t)((Cneg >> 1) & nsig) << r0;
                            mem.B(c) |= (uint64_t)nsig << r0;
                        }
                    }
                    Lw = Cw;
                    Lneg = Cneg;
                    Cw = Rw;
                    Cneg = Rneg;
                }
            }
            if (pass == 2)
                for (int c = 0; c < w; ++c) mem.V(c) = 0;
        }
        for (int c = 0; c < 64; ++c) job.planes[p * 64 + c] = c < w ? mem.B(c) : 0;
    }
    for (int c = 0; c < 64; ++c) job.negs[c] = c < w ? mem.N(c) : 0;
}

}  // namespace rg_t1
