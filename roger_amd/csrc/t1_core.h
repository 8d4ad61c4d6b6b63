// EBCOT tier-1 encoder cores (T.800 Annex D) with the MQ coder (Annex C).
//
// Replaces the reference's bit-plane coder skeleton src/CL/oclbpc.cl:279-746
// (launched per channel by src/OCLBPC.cpp:38-51), which stops after the
// per-block MSB and has no context formation, MQ coder or output.
//
// Written against the wavefront primitives of wave_hip.h (RG_DEV, wv_*), so
// the kernels in t1_encode.hip and the host lane emulator build the same code.
//
// Design (gfx950, wave64): one 64-lane wavefront per code-block (<= 64x64).
// Lane j owns column j; every per-sample state (significance, visited,
// refined, sign, current magnitude bit-plane) is a 64-bit row mask in
// registers, so the 8-neighbourhood of a 4-row stripe is a 6-bit window of
// the lane's own masks plus the windows of lanes j-1 / j+1 (one cross-lane
// shuffle each).  Per stripe and pass:
//   * significance propagation: a sample joins the pass if any neighbour is
//     significant at its turn in scan order.  The only cross-lane dependency
//     is "left column became significant earlier in this stripe"; it is
//     resolved by a wave-wide fix-point (recompute with the left lane's
//     newly-significant set until no lane changes, checked with wv_any), which
//     converges in (longest left-to-right chain + 1) rounds, usually 1-2.
//   * magnitude refinement: no intra-pass dependency.
//   * cleanup: newly significant samples are known up front (bit == 1), so
//     contexts and the run-length test come straight from the windows.
// Each lane keeps its column's (context, bit) symbols for the stripe in two
// 64-bit registers.  Two ways to MQ-code them:
//   * inline: the coder walks the lanes in scan order (ballot mask, readlane)
//     as wave-uniform scalar code whose context states and probability table
//     sit in VGPR lanes;
//   * split: the stripe's symbols are appended to a per-block stream, and a
//     second kernel runs one MQ coder per LANE (64 blocks per wavefront).
#pragma once
#include <stdint.h>
#include "../../include/roger_amd.h"

namespace rg_t1 {

// MQ probability table (T.800 Table C.2): Qe | NMPS << 16 | NLPS << 22 | SWITCH << 28.
#define MQE(q, nm, nl, sw) ((uint32_t)(q) | ((uint32_t)(nm) << 16) | ((uint32_t)(nl) << 22) | ((uint32_t)(sw) << 28))
RG_CONST uint32_t MQT[47] = {
    MQE(0x5601, 1, 1, 1),   MQE(0x3401, 2, 6, 0),   MQE(0x1801, 3, 9, 0),   MQE(0x0AC1, 4, 12, 0),
    MQE(0x0521, 5, 29, 0),  MQE(0x0221, 38, 33, 0), MQE(0x5601, 7, 6, 1),   MQE(0x5401, 8, 14, 0),
    MQE(0x4801, 9, 14, 0),  MQE(0x3801, 10, 14, 0), MQE(0x3001, 11, 17, 0), MQE(0x2401, 12, 18, 0),
    MQE(0x1C01, 13, 20, 0), MQE(0x1601, 29, 21, 0), MQE(0x5601, 15, 14, 1), MQE(0x5401, 16, 14, 0),
    MQE(0x5101, 17, 15, 0), MQE(0x4801, 18, 16, 0), MQE(0x3801, 19, 17, 0), MQE(0x3401, 20, 18, 0),
    MQE(0x3001, 21, 19, 0), MQE(0x2801, 22, 19, 0), MQE(0x2401, 23, 20, 0), MQE(0x2201, 24, 21, 0),
    MQE(0x1C01, 25, 22, 0), MQE(0x1801, 26, 23, 0), MQE(0x1601, 27, 24, 0), MQE(0x1401, 28, 25, 0),
    MQE(0x1201, 29, 26, 0), MQE(0x1101, 30, 27, 0), MQE(0x0AC1, 31, 28, 0), MQE(0x09C1, 32, 29, 0),
    MQE(0x08A1, 33, 30, 0), MQE(0x0521, 34, 31, 0), MQE(0x0441, 35, 32, 0), MQE(0x02A1, 36, 33, 0),
    MQE(0x0221, 37, 34, 0), MQE(0x0141, 38, 35, 0), MQE(0x0111, 39, 36, 0), MQE(0x0085, 40, 37, 0),
    MQE(0x0049, 41, 38, 0), MQE(0x0025, 42, 39, 0), MQE(0x0015, 43, 40, 0), MQE(0x0009, 44, 41, 0),
    MQE(0x0005, 45, 42, 0), MQE(0x0001, 45, 43, 0), MQE(0x5601, 46, 46, 0)};
#undef MQE

constexpr int CTX_RL = 17, CTX_UNI = 18;

// ZC context (Table D.1), same rule as the oracle.
RG_DEV int zc_ctx(int orient, int h, int v, int d)
{
    if (orient == 1) { int t = h; h = v; v = t; }
    if (orient == 3) {
        int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv >= 1 ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv == 1 ? 4 : 3);
        return hv >= 2 ? 2 : hv;
    }
    if (h == 2) return 8;
    if (h == 1) return v >= 1 ? 7 : (d >= 1 ? 6 : 5);
    if (v == 2) return 4;
    if (v == 1) return 3;
    return d >= 2 ? 2 : d;
}

// Sign context (Table D.3): returns ctx << 1 | xorbit.
RG_DEV int sc_ctx(int H, int V)
{
    int x = 0;
    if (H < 0 || (H == 0 && V < 0)) { x = 1; H = -H; V = -V; }
    int c = (H == 1) ? (V == 1 ? 13 : (V == 0 ? 12 : 11)) : (V == 1 ? 10 : 9);
    return (c << 1) | x;
}

RG_DEV uint32_t win6(uint64_t m, int r0)
{
    // bits r0-1 .. r0+4 of m (bit 0 of the window = row r0-1)
    return r0 == 0 ? (uint32_t)(m << 1) & 0x3Fu : (uint32_t)(m >> (r0 - 1)) & 0x3Fu;
}

RG_DEV uint32_t b1(uint32_t w, int i) { return (w >> i) & 1u; }

// A lane's (context << 1 | bit) symbols of one stripe, oldest in the low
// byte of f0: at most 10 (cleanup: run-length, 2 uniform, sign, then 3 x
// (ZC, sign)).
struct SymRun {
    uint32_t f0 = 0, f1 = 0, f2 = 0;
    int n = 0;
};

// Byte k of spread(w) holds bits k, k+1, ... of w (w < 64): the four copies
// of w shifted left by 7k do not overlap, so the product is their OR, and
// bits 0..5-k of byte k are clean.  One 24-bit multiply turns a 6-bit row
// window into the four rows' neighbourhoods, one row per byte.
RG_DEV uint32_t spread(uint32_t w) { return wv_mul24(w, 0x00204081u); }
// bit k of a 4-bit row mask -> bit 0 of byte k
RG_DEV uint32_t spread4(uint32_t m) { return spread(m) & 0x01010101u; }

// Context tables of one code-block in workgroup memory (filled by
// t1_lut_fill for the block's orientation):
//   zc[256]: ZC context (Table D.1) of a neighbourhood pattern -- bits 0-2
//            left column rows i-1..i+1, 3-5 right column rows i-1..i+1,
//            6 the row above, 7 the row below (1 = significant);
//   sc[256]: SC context << 1 | XOR bit (Table D.3) of the (significant,
//            negative) pairs of the left (bits 0-1), right (2-3), above (4-5)
//            and below (6-7) neighbours.
constexpr int LUT_BYTES = 512;
// The context modelling also keeps, after the context tables, the byte
// selectors that pack a stripe's symbols in scan order: entry z | n << 4 (z:
// rows with a ZC or refinement symbol, n: rows with a sign symbol) holds the
// v_perm selectors (two words) that take, row by row, byte k of the first
// word (k) and then byte k of the sign word (4 + k); unused bytes select zero.
constexpr int SEL_BYTES = 256 * 8;
constexpr int CM_LDS_BYTES = LUT_BYTES + SEL_BYTES;

RG_DEV void t1_sel_fill(uint8_t *sel)
{
    const int lane = wv_lane();
    uint32_t *w = reinterpret_cast<uint32_t *>(sel);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x = lane * 4 + e, z = x & 15, n = x >> 4;
        uint64_t v = 0x0C0C0C0C0C0C0C0Cull;
        int j = 0;
        for (int k = 0; k < 4; ++k) {
            if ((z >> k) & 1) { v = (v & ~(0xFFull << (8 * j))) | ((uint64_t)k << (8 * j)); ++j; }
            if ((n >> k) & 1) { v = (v & ~(0xFFull << (8 * j))) | ((uint64_t)(4 + k) << (8 * j)); ++j; }
        }
        w[2 * x] = (uint32_t)v;
        w[2 * x + 1] = (uint32_t)(v >> 32);
    }
}


RG_DEV void t1_lut_fill(uint8_t *lut, int orient)
{
    const int lane = wv_lane();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x = lane * 4 + e;
        const int h = ((x >> 1) & 1) + ((x >> 4) & 1);
        const int v = ((x >> 6) & 1) + ((x >> 7) & 1);
        const int d = (x & 1) + ((x >> 2) & 1) + ((x >> 3) & 1) + ((x >> 5) & 1);
        lut[x] = (uint8_t)zc_ctx(orient, h, v, d);
        int c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = ((x >> (2 * k)) & 1) ? (((x >> (2 * k + 1)) & 1) ? -1 : 1) : 0;
        int H = c[0] + c[1], V = c[2] + c[3];
        H = H > 1 ? 1 : (H < -1 ? -1 : H);
        V = V > 1 ? 1 : (V < -1 ? -1 : V);
        lut[256 + x] = (uint8_t)sc_ctx(H, V);
    }
    wv_sync();
}

// ZC and sign symbols of the four rows of a stripe, one row per byte, in a
// significance-propagation or cleanup pass.  Windows (bit 0 = the row above
// the stripe): sw own column before the pass, Lw the left column's final
// state, rw the right column before the pass, N the rows of this column that
// become significant (so the row above row k is significant at its turn iff
// bit k of sw | N << 1); negw/neglw/negrw the sign windows; bb bit k = row k's
// magnitude bit.  Table indices as t1_lut_fill builds them: ZC bits 0-2 left
// rows k-1..k+1, 3-5 right rows k-1..k+1, 6 above, 7 below; sign bits 0-1
// left (significant, negative), 2-3 right, 4-5 above, 6-7 below.  Zs byte k =
// ZC context << 1 | bit; Ss byte k = sign context << 1 | (sign XOR bit), built
// only when want_sc (a wave-uniform flag: some lane has a new significance).
RG_DEV void zc_sc_syms(const uint8_t *lut, uint32_t N, uint32_t sw, uint32_t Lw, uint32_t rw, uint32_t negw,
                       uint32_t neglw, uint32_t negrw, uint32_t bb, bool want_sc, uint32_t &Zs, uint32_t &Ss)
{
    const uint32_t mL = spread(Lw), mR = spread(rw), mC = spread(sw | (N << 1)), mS = spread(sw);
    const uint32_t zi = (mL & 0x07070707u) | ((mR << 3) & 0x38383838u) | ((mC << 6) & 0x40404040u) |
                        ((mS << 5) & 0x80808080u);
    const uint32_t z0 = lut[zi & 0xFFu], z1 = lut[(zi >> 8) & 0xFFu], z2 = lut[(zi >> 16) & 0xFFu],
                   z3 = lut[zi >> 24];
    Zs = ((z0 | (z1 << 8) | (z2 << 16) | (z3 << 24)) << 1) | spread4(bb);
    Ss = 0;
    if (want_sc) {
        const uint32_t mN = spread(negw);
        const uint32_t si = ((mL >> 1) & 0x01010101u) | (spread(neglw) & 0x02020202u) | ((mR << 1) & 0x04040404u) |
                            ((spread(negrw) << 2) & 0x08080808u) | ((mC << 4) & 0x10101010u) |
                            ((mS << 4) & 0x40404040u) | ((mN << 5) & 0xA0A0A0A0u);
        const uint8_t *sc = lut + 256;
        const uint32_t s0 = sc[si & 0xFFu], s1 = sc[(si >> 8) & 0xFFu], s2 = sc[(si >> 16) & 0xFFu], s3 = sc[si >> 24];
        Ss = (s0 | (s1 << 8) | (s2 << 16) | (s3 << 24)) ^ ((mN >> 1) & 0x01010101u);
    }
}

// A stripe's symbols in scan order: row by row, Zs byte k if bit k of Z,
// then Ss byte k if bit k of Nm (sel: the selector table).
RG_DEV SymRun pack_run(const uint8_t *sel, uint32_t Z, uint32_t Nm, uint32_t Zs, uint32_t Ss)
{
    const uint32_t *w = reinterpret_cast<const uint32_t *>(sel) + 2 * (Z | (Nm << 4));
    SymRun r;
    r.f0 = wv_perm(Ss, Zs, w[0]);
    r.f1 = wv_perm(Ss, Zs, w[1]);
    r.n = __builtin_popcount(Z) + __builtin_popcount(Nm);
    return r;
}

struct MQEnc {
    uint32_t A, C, B;
    int CT, bp;       // bp: index of the pending byte B (-1: the fake byte before the codeword)
    int cap, overflow;
    uint8_t *out;
};

RG_DEV void mq_store(MQEnc &m, int lane)
{
    if (m.bp >= 0) {
        if (m.bp < m.cap) m.out[m.bp] = (uint8_t)m.B;   // B is uniform: every lane stores the same byte
        else m.overflow = 1;
    }
}

// BYTEOUT (C.2.7).  B always holds one byte: a carry out of C (bit 27) is
// added to B and then cleared from C before the next byte is taken.
RG_DEV void mq_byteout(MQEnc &m, int lane)
{
    if (m.B == 0xFF) {
        mq_store(m, lane); m.bp++; m.B = m.C >> 20; m.C &= 0xFFFFF; m.CT = 7;
    } else if (m.C < 0x8000000u) {
        mq_store(m, lane); m.bp++; m.B = m.C >> 19; m.C &= 0x7FFFF; m.CT = 8;
    } else {
        m.B++;
        m.C &= 0x7FFFFFF;
        if (m.B == 0xFF) {
            mq_store(m, lane); m.bp++; m.B = m.C >> 20; m.C &= 0xFFFFF; m.CT = 7;
        } else {
            mq_store(m, lane); m.bp++; m.B = m.C >> 19; m.C &= 0x7FFFF; m.CT = 8;
        }
    }
}

RG_DEV void mq_renorm(MQEnc &m, int lane)
{
    // shift until bit 15 of A is set, emitting a byte whenever CT runs out
    int n = wv_clz(m.A) - 16;
    while (n > 0) {
        int s = n < m.CT ? n : m.CT;
        m.A <<= s; m.C <<= s; m.CT -= s; n -= s;
        if (m.CT == 0) mq_byteout(m, lane);
    }
}

// ENCODE (C.2.2-C.2.6) as wave-uniform scalar code.  The 19 context states
// (I << 1 | MPS) live in lanes 0..18 of the VGPR `ctxv` and the probability
// table in lanes 0..46 of `mqtv`; readlane with a uniform index (and a lane
// select to write a state back) replaces every memory access of the coder.
RG_DEV void mq_encode(MQEnc &m, int &ctxv, int mqtv, int cx, int d, int lane)
{
    const int st = wv_readlane(ctxv, cx);
    const int mps = st & 1;
    const uint32_t t = (uint32_t)wv_readlane(mqtv, st >> 1);
    const uint32_t qe = t & 0xFFFF;
    m.A -= qe;
    if (d == mps) {
        if (m.A & 0x8000) {
            m.C += qe;
            return;
        }
        if (m.A < qe) m.A = qe; else m.C += qe;
        ctxv = lane == cx ? (int)((((t >> 16) & 63) << 1) | mps) : ctxv;
    } else {
        if (m.A < qe) m.C += qe; else m.A = qe;
        const int nm = mps ^ (int)(t >> 28);
        ctxv = lane == cx ? (int)((((t >> 22) & 63) << 1) | nm) : ctxv;
    }
    mq_renorm(m, lane);
}

// Codes one stripe's symbols in scan order: columns (lanes) left to right,
// each lane's own run oldest first.  Lanes without symbols are skipped via
// the ballot mask; a lane's run is read with three readlanes.
RG_DEV void mq_stripe(MQEnc &m, int &ctxv, int mqtv, const SymRun &r, int lane)
{
    uint64_t mask = wv_ballot(r.n > 0);
    while (mask) {
        const int j = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int n = wv_readlane(r.n, j);
        const uint32_t x0 = (uint32_t)wv_readlane((int)r.f0, j), x1 = (uint32_t)wv_readlane((int)r.f1, j),
                       x2 = (uint32_t)wv_readlane((int)r.f2, j);
        for (int t = 0; t < n; ++t) {
            const uint32_t w = t < 4 ? x0 : (t < 8 ? x1 : x2);
            const int sym = (int)((w >> ((t & 3) * 8)) & 0xFFu);
            mq_encode(m, ctxv, mqtv, sym >> 1, sym & 1, lane);
        }
    }
}

// Code-block styles (T.800 Table A.19) the coders implement.
constexpr int STYLE_BYPASS = 0x01;    // raw SPP/MRP passes from the fifth bit-plane (with TERMALL)
constexpr int STYLE_RESET = 0x02;     // contexts reset after every coding pass
constexpr int STYLE_TERMALL = 0x04;   // codeword terminated after every pass
constexpr int MAXP = 96;              // per-block pass-array stride (>= 3*32-2)

// Under BYPASS, coding pass k (0 = the first cleanup pass) is raw-coded when
// it is a significance-propagation or refinement pass of the fifth or a later
// bit-plane: k >= 10 and k mod 3 != 0.
RG_DEV bool pass_is_raw(int style, int k) { return (style & STYLE_BYPASS) && k >= 10 && k % 3 != 0; }

// One code-block job: coefficients (row pitch), size, orientation, output slot.
struct T1Job {
    const int32_t *src;
    int pitch, w, h, orient, cap;
    uint8_t *out;
    rg_cb_result *res;
    uint8_t *sym;   // split path: this block's symbol stream (symcap bytes, 16-byte aligned)
    int symcap;
    int style;      // split path: code-block style bits (STYLE_*); BYPASS changes raw signs
    i32x4 *cm;      // split path: {symbols, passes, bit-planes, overflow}
    int32_t *poff;  // split path (may be null): stream offset where each pass starts (MAXP)
    uint8_t *seg;   // parallel style: this block's segment arena (seg_bytes(symcap) bytes)
    int32_t *seglen;// parallel style: codeword segment length of each pass (MAXP)
    u32x4 *cms;     // two-phase context modelling: 64 per-lane state words kept between the phases
    uint64_t *sppm = nullptr;  // split path (may be null): per bit-plane p, word p*64+lane = the rows of
                               // column `lane` that became significant in plane p's significance pass
};

constexpr int RD_PLANES = 32;   // bit-planes per block in the significance-pass mask records

// Context modelling of coding passes [qbeg, qend) only (split path).  A phase
// ends before a significance-propagation pass (qend = 1 mod 3), where the
// per-lane state is just the significance and refinement masks: they, the
// stream position and the overflow flag carry over (job.cms, job.cm).
struct CMPhase {
    int qbeg = 0, qend = 1 << 30;
};

// Segment arena of the parallel style: pass p, whose symbols start at stream
// offset s0, codes into [seg_at(s0, p), +2*(symbols)+16): an MQ step emits at
// most 15 bits and a flush at most 5 bytes, so neighbouring passes never
// overlap and every segment starts 16-byte aligned.
RG_DEV int seg_at(int s0, int p) { return ((2 * s0 + 15) & ~15) + 32 * p; }
inline int seg_bytes(int symcap) { return 2 * symcap + 32 * MAXP + 32; }

// Exclusive prefix sum over the wavefront of a per-lane count n and the wave
// total (a DPP scan; lane 63's inclusive sum is the total).
RG_DEV int wave_excl_count(int n, int &total)
{
    const int incl = wv_incl_scan(n);
    total = wv_readlane(incl, 63);
    return incl - n;
}

// Split path: appends one stripe's symbols (scan order) to the block's
// stream, each lane's run at its exclusive prefix count (so every lane's
// range lies inside the stripe's [pos, pos + total), inside the stream).
// Byte t of the run goes out in the t-th store; lanes whose run is shorter
// aim it outside the stream, where the sink's range check drops it (no
// exec-mask branch).
RG_DEV void emit_run(const SymRun &r, const T1Job &job, int &pos, int &ovf)
{
    int total;
    const int off = pos + wave_excl_count(r.n, total);
    if (total == 0) return;
    if (pos + total > job.symcap) { ovf = 1; return; }
    const ByteSink snk = wv_byte_sink(job.sym, job.symcap);
#pragma unroll
    for (int t = 0; t < 12; ++t) {
        const bool has = t < r.n;
        if (!wv_any(has)) break;
        const uint32_t w = t < 4 ? r.f0 : (t < 8 ? r.f1 : r.f2);
        wv_sink_store(snk, has ? off : SINK_DROP, t, w >> (8 * (t & 3)));
    }
    pos += total;
}

// Encodes one code-block with the calling wavefront (64 lanes).  SPLIT:
// context modelling only, the symbols go to job.sym for the lane MQ coder.
// lut: LUT_BYTES of workgroup memory for the block's context tables.
template <bool SPLIT>
RG_DEV void t1_block(const T1Job &job, uint8_t *lut, CMPhase ph = CMPhase())
{
    const int lane = wv_lane();
    const int w = job.w, h = job.h, orient = job.orient, pitch = job.pitch;
    const int32_t *src = job.src;

    // ---- load column `lane`: magnitudes in registers, sign as a row mask ----
    uint32_t mag[64];
    uint64_t NEG = 0;
    uint32_t orm = 0;
    const bool colok = lane < w;
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        int v = (colok && r < h) ? src[(size_t)r * pitch + lane] : 0;
        uint32_t m = (uint32_t)(v < 0 ? -v : v);
        mag[r] = m;
        NEG |= (uint64_t)(v < 0) << r;
        orm |= m;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) orm |= wv_shfl_xor(orm, o);
    // the butterfly leaves the same value in every lane; saying so keeps the
    // whole bit-plane loop (and the MQ coder) in scalar registers
    const int nbp = wv_uniform(orm ? 32 - wv_clz(orm) : 0);
    // (block results are stored by every lane -- the same value to the same
    // address -- so no statement of the block body depends on the lane index
    // and a caller's loop around it stays wave-uniform)
    if (nbp == 0) {
        if (SPLIT) *job.cm = mk_i32x4(0, 0, 0, 0);
        else *job.res = rg_cb_result{0, 0, 0, 0};
        return;
    }
    const int npass = 3 * nbp - 2;
    if (SPLIT && ph.qbeg >= npass) return;   // a later phase with nothing left: the record is final
    t1_sel_fill(lut + LUT_BYTES);
    t1_lut_fill(lut, orient);   // (ends with the wavefront barrier both fills need)
    int spos = 0, sovf = 0;   // split path: stream position / overflow (uniform)
    int pidx = 0;             // split path: coding pass index

    // initial context states (Table D.7): ZC ctx 0 -> state 4, RL -> 3, UNIFORM -> 46, others 0
    int ctxv = lane == 0 ? (4 << 1) : lane == CTX_RL ? (3 << 1) : lane == CTX_UNI ? (46 << 1) : 0;
    const int mqtv = lane < 47 ? (int)MQT[lane] : 0;

    const uint64_t rowmask = h >= 64 ? ~0ull : ((1ull << h) - 1);
    const uint64_t NEGL = wv_shfl_up(NEG, 1);
    const uint64_t NEGR = wv_shfl_down(NEG, 1);
    const uint64_t negl = lane == 0 ? 0 : NEGL;
    const uint64_t negr = lane == 63 ? 0 : NEGR;

    MQEnc m;
    m.A = 0x8000; m.C = 0; m.B = 0; m.CT = 12; m.bp = -1;
    m.cap = job.cap; m.overflow = 0;
    m.out = job.out;

    uint64_t S = 0, V = 0, R = 0;
    const int nstripes = (h + 3) >> 2;

    // a later phase resumes at pass qbeg (bit-plane (qbeg + 2) / 3 from the top,
    // pass type (qbeg - 1) mod 3) from the state the previous phase left
    int pstart = nbp - 1, t0 = 2;
    if (SPLIT && ph.qbeg > 0) {
        const i32x4 c = *job.cm;
        spos = wv_uniform(c.x);
        sovf = wv_uniform(c.w);
        const u32x4 st = wv_load16(reinterpret_cast<const uint8_t *>(job.cms + lane));
        S = ((uint64_t)st.y << 32) | st.x;
        R = ((uint64_t)st.w << 32) | st.z;
        pidx = ph.qbeg;
        pstart = nbp - 1 - (ph.qbeg + 2) / 3;
        t0 = (ph.qbeg - 1) % 3;
    }

    for (int p = pstart; p >= 0; --p) {
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int r = 0; r < 32; ++r) lo |= ((mag[r] >> p) & 1u) << r;
#pragma unroll
        for (int r = 0; r < 32; ++r) hi |= ((mag[r + 32] >> p) & 1u) << r;
        const uint64_t B = ((uint64_t)hi << 32) | lo;

        for (int pass = p == pstart ? t0 : 0; pass < 3; ++pass) {
            if (SPLIT && job.poff) job.poff[pidx] = spos;   // uniform value, stored by every lane
            if (SPLIT && pidx >= ph.qend) {
                // end of this phase: keep the state for the next one (V is clear
                // after a cleanup pass); the record holds the stream position
                wv_store16(reinterpret_cast<uint8_t *>(job.cms + lane), (uint32_t)S, (uint32_t)(S >> 32), (uint32_t)R,
                           (uint32_t)(R >> 32));
                *job.cm = mk_i32x4(spos, npass, nbp, sovf);
                return;
            }
            ++pidx;
            const uint64_t S0 = S;
            // under BYPASS a raw pass (this pass is pidx - 1) sends its decisions
            // and signs as bits, without contexts
            const bool rawp = SPLIT && pass != 2 && pass_is_raw(job.style, pidx - 1);
            for (int s = 0; s < nstripes; ++s) {
                const int r0 = 4 * s;
                const uint32_t vrow = colok ? (uint32_t)(rowmask >> r0) & 0xFu : 0u;
                const uint32_t sw = win6(S, r0);
                const uint32_t lw = wv_from_left(sw), rw = wv_from_right(sw);
                const uint32_t bb = (win6(B, r0) >> 1) & 0xFu;   // bit k: row k's magnitude bit
                const uint32_t own = ~(sw >> 1) & vrow & 0xFu;     // valid rows not yet significant
                // Every pass reduces to: Zm rows with a ZC / refinement / raw-bit
                // symbol (bytes of Zs), Nm rows with a sign symbol after it (bytes
                // of Ss), and a cleanup run-length prefix of plen bytes.
                uint32_t Zm, Nm = 0, Zs, Ss = 0, pre = 0;
                int plen = 0;

                if (pass == 0) {
                    // ---- significance propagation ----
                    // Row k joins the pass if a neighbour is significant at its turn:
                    // before the pass (own column above and below, rows k-1..k+1 of
                    // the right column and of the left one), or newly significant in
                    // it (the row above; the left column, coded earlier).  If no
                    // candidate of any column has one of the former, the stripe has
                    // no pass symbols.
                    const uint32_t hpre = (sw | (sw >> 2) | rw | (rw >> 1) | (rw >> 2)) & 0xFu;
                    if (!wv_any(((hpre | lw | (lw >> 1) | (lw >> 2)) & own) != 0)) continue;
                    // joining through the row above: it joined and its bit is 1
                    const uint32_t q = own & (bb << 1);
                    uint32_t P = 0, N = 0, NL = 0;
                    // monotone fix-point over the left-column dependency: converges within
                    // 65 rounds (a chain across <= 64 lanes).  Another round runs only when
                    // the left column's newly significant samples border a row of this
                    // column that is not yet in the pass (the only way membership grows).
                    for (int round = 0; round < 80; ++round) {
                        const uint32_t Lr = lw | (NL << 1);
                        P = own & (hpre | Lr | (Lr >> 1) | (Lr >> 2));
                        P |= q & (P << 1);   // a newly significant row pulls in the one below,
                        P |= q & (P << 1);   // at most three times down the stripe
                        P |= q & (P << 1);
                        N = P & bb;
                        const uint32_t NLn = wv_from_left(N);
                        const uint32_t grown = NLn & ~NL;
                        NL = NLn;
                        if (!wv_any(((grown | (grown << 1) | (grown >> 1)) & own & ~P) != 0)) break;
                    }
                    const uint32_t negw = win6(NEG, r0);
                    if (rawp) {
                        Zs = spread4(bb);
                        Ss = (spread(negw) >> 1) & 0x01010101u;
                    } else {
                        // the left column's final state for the contexts
                        zc_sc_syms(lut, N, sw, lw | (NL << 1), rw, negw, win6(negl, r0), win6(negr, r0), bb,
                                   wv_any(N != 0), Zs, Ss);
                    }
                    Zm = P;
                    Nm = N;
                    S |= (uint64_t)N << r0;
                    V |= (uint64_t)P << r0;
                } else if (pass == 1) {
                    // ---- magnitude refinement (no intra-pass dependency) ----
                    const uint32_t vw = (uint32_t)(V >> r0) & 0xFu;
                    const uint32_t rf = (uint32_t)(R >> r0) & 0xFu;
                    Zm = (sw >> 1) & ~vw & vrow & 0xFu;
                    if (!wv_any(Zm != 0)) continue;
                    if (rawp) {
                        Zs = spread4(bb);
                    } else {
                        // context 16 after the first refinement, else 15 with a
                        // significant neighbour, else 14
                        const uint32_t t = lw | rw;
                        const uint32_t nb = (sw | (sw >> 2) | t | (t >> 1) | (t >> 2)) & ~rf & 0xFu;
                        Zs = ((0x0E0E0E0Eu + spread4(nb) + (spread4(rf) << 1)) << 1) | spread4(bb);
                    }
                    R |= (uint64_t)Zm << r0;
                } else {
                    // ---- cleanup ----
                    const uint32_t vw = (uint32_t)(V >> r0) & 0xFu;
                    const uint32_t Cm = own & ~vw;
                    // every sample of the stripe significant or visited: no cleanup
                    // symbols (most stripes of the low bit-planes)
                    if (!wv_any(Cm != 0)) continue;
                    const uint32_t N = Cm & bb;   // newly significant rows are known up front
                    const uint32_t NL = wv_from_left(N);
                    const uint32_t Lw = lw | (NL << 1);
                    const bool full = r0 + 4 <= h;
                    const bool rl = full && Cm == 0xFu && (Lw & 0x3Fu) == 0 && (rw & 0x3Fu) == 0 && (sw & 0x21u) == 0;
                    // run-length mode: (RL, 0) for an all-zero column stripe, else
                    // (RL, 1), the first 1's row as two UNIFORM bits, its sign, and
                    // ZC coding of the rows below it
                    const int kf = __builtin_ctz(N | 0x10u);
                    Zm = rl ? ((0x1Eu << kf) & 0xFu) : Cm;
                    Nm = N;
                    if (rl) {
                        pre = N ? ((2u * CTX_RL + 1) | ((2u * CTX_UNI + (uint32_t)(kf >> 1)) << 8) |
                                 ((2u * CTX_UNI + (uint32_t)(kf & 1)) << 16))
                              : 2u * CTX_RL;
                        plen = N ? 3 : 1;
                    }
                    Zs = 0;
                    if (wv_any((Zm | Nm) != 0))
                        zc_sc_syms(lut, N, sw, Lw, rw, win6(NEG, r0), win6(negl, r0), win6(negr, r0), bb,
                                   wv_any(N != 0), Zs, Ss);
                    S |= (uint64_t)N << r0;
                }

                SymRun run = pack_run(lut + LUT_BYTES, Zm, Nm, Zs, Ss);
                if (pass == 2) {
                    // the run-length prefix goes in front (plen 0, 1 or 3 bytes)
                    const int sh = 8 * plen;
                    const uint64_t body = ((uint64_t)run.f1 << 32) | run.f0;
                    const uint64_t lo = body << sh;
                    run.f2 = (uint32_t)(((uint64_t)run.f1 << sh) >> 32);
                    run.f0 = (uint32_t)lo | pre;
                    run.f1 = (uint32_t)(lo >> 32);
                    run.n += plen;
                }

                // ---- MQ-code this stripe's symbols in scan order (or hand them on) ----
                if constexpr (SPLIT) emit_run(run, job, spos, sovf);
                else mq_stripe(m, ctxv, mqtv, run, lane);
            }
            if (pass == 2) V = 0;
            if (SPLIT && pass == 0 && job.sppm) job.sppm[p * 64 + lane] = S & ~S0;
        }
    }

    if constexpr (SPLIT) {
        *job.cm = mk_i32x4(spos, npass, nbp, sovf);
        return;
    }
    // ---- FLUSH (C.2.9) ----
    {
        uint32_t tmp = m.C + m.A;
        m.C |= 0xFFFF;
        if (m.C >= tmp) m.C -= 0x8000;
        m.C <<= m.CT; mq_byteout(m, lane);
        m.C <<= m.CT; mq_byteout(m, lane);
        int len;
        if (m.B != 0xFF) { mq_store(m, lane); len = m.bp + 1; }
        else len = m.bp;
        if (len > job.cap) m.overflow = 1;
        *job.res = rg_cb_result{len, 3 * nbp - 2, nbp, m.overflow ? 1 : 0};
    }
}

// ---- split path, second half: one coder per lane ----
//
// Symbols arrive 16 at a time from unconditional (index-clamped) 16-byte
// loads issued ahead of use, so the memory latency stays off a chain's
// dependency path.  Output bytes go to a per-lane ring in workgroup memory
// (an LDS store per byte, whose waits never hold up a global load) and leave
// for global memory in 16-byte groups; output regions are 16-byte aligned.

// ---- lane MQ coder, pipelined form ----
//
// One MQ chain per lane.  A context's state lives in workgroup memory as its
// probability-table entry (Qe | NMPS << 16 | NLPS << 22 | SWITCH << 28) with
// the MPS in bit 31, one word per (context, lane), so a symbol reads its
// context with one LDS load and no table lookup, and writes the context back
// with one store.  A chain's serial dependency runs through that entry: the
// next state's table entry is one more LDS read.  So a step issues, before
// its interval arithmetic, both that read and the read of the NEXT symbol's
// context (forwarded from the store when the next symbol has the same
// context, since the store is issued later), and renormalisation and BYTEOUT
// overlap the reads.  Output bytes go to a per-lane 64-byte ring in workgroup
// memory with one unconditional store per BYTEOUT (a step without BYTEOUT
// rewrites the pending byte's own slot with its current value), so a step
// has no store branch; completed 16-byte groups leave for global memory once
// per 16-symbol chunk (ring_flush).  A chunk adds at most 48 bytes (three
// BYTEOUTs per step at most) to fewer than 16 unflushed ones, so the ring
// never overwrites an unflushed byte.  Chunks wholly inside a lane's symbol
// range run without per-symbol range tests.
constexpr int RING_STRIDE = 68;          // bytes per lane: 17 dwords, so lanes start in distinct LDS banks
constexpr int RING_BYTES = 64 * RING_STRIDE;
constexpr int MQ_CTX = 19;
// workgroup memory of the lane MQ coders: the rings, then the context words
// (context cx of lane l at word cx * 64 + l, so a wave's accesses hit 64
// distinct banks)
constexpr int MQ_LDS_BYTES = RING_BYTES + MQ_CTX * 64 * 4;

// Initial context entries (Table D.7: context 0 in state 4, RL (17) in 3,
// UNIFORM (18) in 46, all others in 0; MPS 0) for this lane (ce = its word
// of context 0).
// A lane coder's context entry: Qe | MPS << 16 | state index << 17.  The
// transition table (MQ_TT_WORDS of workgroup memory) holds at 2 * state + lps
// the entry of the state a renormalising decision moves to, with bit 16 set
// where that transition switches the MPS, so the entry after a renormalising
// step is TT[((e >> 16) & 0x7F) ^ d] ^ (e & 0x10000) (d the decision).
constexpr int MQ_TT_WORDS = 96;
RG_DEV uint32_t ce_entry(int st) { return (MQT[st] & 0xFFFFu) | ((uint32_t)st << 17); }
RG_DEV uint32_t tt_entry(int k)
{
    const uint32_t t = MQT[k >> 1], l = (uint32_t)k & 1u;
    const int j = (int)((t >> (l ? 22 : 16)) & 63u);
    return ce_entry(j) | ((l & (t >> 28)) << 16);
}
// Fills the transition table (every lane of the workgroup calls it; a
// barrier must follow).
RG_DEV void mq_tt_fill(uint32_t *tab)
{
    const int lane = wv_lane();
    if (wv_wave() == 0) {
        tab[lane] = tt_entry(lane);
        if (lane < 94 - 64) tab[64 + lane] = tt_entry(64 + lane);
    }
}

RG_DEV void ce_init(uint32_t *ce)
{
#pragma unroll
    for (int cx = 0; cx < MQ_CTX; ++cx) ce[64 * cx] = ce_entry(cx == 0 ? 4 : (cx == CTX_RL ? 3 : (cx == CTX_UNI ? 46 : 0)));
}

struct RingMQ {
    uint32_t A, C, B;
    int CT, bp;    // bp: position of the pending byte B (-1: the byte before the codeword)
    int fl;        // bytes [0, fl) are in global memory
    int ovf;
};

// BYTEOUT (C.2.7) when `need`, as predicated code: a carry into B, then 8 or
// (after 0xFF, bit stuffing) 7 bits move out of C.  B's slot is written in
// either case.
RG_DEV void ring_byteout(RingMQ &m, bool need, uint8_t *rb)
{
    const bool carry = need && m.B != 0xFF && m.C >= 0x8000000u;
    const uint32_t b2 = m.B + (carry ? 1u : 0u);
    const uint32_t c = carry ? (m.C & 0x7FFFFFFu) : m.C;
    const bool big = b2 == 0xFF;
    rb[m.bp & 63] = (uint8_t)b2;   // bp = -1 (no byte yet) lands in slot 63, rewritten before use
    m.bp += need ? 1 : 0;
    m.B = need ? (big ? (c >> 20) : (c >> 19)) : m.B;
    m.C = need ? (c & (big ? 0xFFFFFu : 0x7FFFFu)) : m.C;
    m.CT = need ? (big ? 7 : 8) : m.CT;
}

// Interval side of ENCODE (C.2.2-C.2.5) for decision d in a context with
// entry e (tnew: its transition-table word for d).  A is kept as A << 16, so
// the renormalisation shift is a plain count of leading zeros and "A - Qe
// below 0x8000" is the sign bit.  A is updated and renormalised, *enew is the
// context's entry after the symbol.  Returns the symbol's record for the code
// side: the amount added to C (Qe when the upper interval is kept, else 0) <<
// 16 | the renormalisation shift.
RG_DEV uint32_t mq_interval(uint32_t &A, uint32_t d, uint32_t e, uint32_t tnew, uint32_t &enew)
{
    const uint32_t qe32 = e << 16;
    const bool lps = ((e >> 16) ^ d) & 1u;
    const uint32_t a = A - qe32;
    const bool renorm = lps || (int32_t)a >= 0;
    // C.2.4 / C.2.5: the interval kept is the upper one (C += Qe) for an MPS
    // unless A < Qe (conditional exchange), and the reverse for an LPS
    const bool small = a < qe32;
    const bool upper = !renorm || (lps ? small : !small);
    A = upper ? a : qe32;
    enew = renorm ? (tnew ^ (e & 0x10000u)) : e;
    const uint32_t n = (uint32_t)__builtin_clz(A);   // A != 0
    A <<= n;
    return (upper ? qe32 : 0u) | n;
}

// Transition-table word index of decision d in a context with entry e.
RG_DEV uint32_t tt_index(uint32_t d, uint32_t e) { return ((e >> 16) & 0x7Fu) ^ d; }

// RENORME's C side (C.2.6): n shifts with BYTEOUT whenever CT reaches 0.  C
// takes min(n, CT), the predicated BYTEOUT, then the rest -- fewer than the 7
// or 8 bits of the new CT, except after an LPS in a very skewed state (n up
// to 15), which the loop finishes (entered only when a lane needs it; a
// second predicated BYTEOUT on every step instead measured slower).
RG_DEV void ring_renorm(RingMQ &m, int n, uint8_t *rb)
{
    const int sft = n < m.CT ? n : m.CT;
    m.C <<= sft; m.CT -= sft; n -= sft;
    ring_byteout(m, m.CT == 0, rb);
    while (n > 0 && n >= m.CT) {
        m.C <<= m.CT; n -= m.CT; m.CT = 0;
        ring_byteout(m, true, rb);
    }
    m.C <<= n; m.CT -= n;
}

// Code side of ENCODE for one symbol's record (mq_interval): C takes the
// addition, then RENORME's shifts with their BYTEOUTs.
RG_DEV void mq_code(RingMQ &m, uint32_t rec, uint8_t *rb)
{
    m.C += rec >> 16;
    ring_renorm(m, (int)(rec & 0xFFu), rb);
}

// mq_code for a shift of at most 7 (so at most one BYTEOUT: after one, CT is
// 7 or 8 and fewer than 7 shifts remain), with the pending byte's class taken
// before C is known: the C chain is add-shift, the carry bit, the next byte's
// width, the mask and the remaining shift.  Same state updates as mq_code.
RG_DEV void mq_code7(RingMQ &m, uint32_t rec, uint8_t *rb)
{
    const int n = (int)(rec & 0xFFu);
    const bool need = n >= m.CT;                     // CT reaches 0 inside this shift
    const int sft = need ? m.CT : n;
    const uint32_t isff = m.B == 0xFFu ? 1u : 0u, isfe = m.B == 0xFEu ? 1u : 0u;
    const uint32_t cyok = need && !isff ? 1u : 0u;   // a carry bit would go into B
    const uint32_t x = (m.C + (rec >> 16)) << sft;
    const uint32_t cy = (x >> 27) & cyok;
    // after B = 0xFF (as it was, or by this carry) the next byte takes 7 bits
    const uint32_t big = isff | (isfe & cy);
    rb[m.bp & 63] = (uint8_t)(m.B + cy);   // the pending byte's slot (rewritten before use when !need)
    const int rem = n - sft;
    if (need) {
        m.bp += 1;
        m.B = (x - (cy << 27)) >> (19 + big);
        m.C = (x & (0x7FFFFu | (big << 19))) << rem;
        m.CT = 8 - (int)big - rem;
    } else {
        m.C = x;
        m.CT -= n;
    }
}

// Moves the ring's completed 16-byte groups [fl, bp) to out (bytes at or
// beyond cap only set the overflow flag).
RG_DEV void ring_flush(RingMQ &m, const uint8_t *rb, uint8_t *out, int cap)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const bool g = m.bp - m.fl >= 16;
        if (!wv_any(g)) break;
        if (g) {
            const uint32_t *w = reinterpret_cast<const uint32_t *>(rb + (m.fl & 63));
            if (m.fl + 16 <= cap) wv_store16(out + m.fl, w[0], w[1], w[2], w[3]);
            else m.ovf = 1;
            m.fl += 16;
        }
    }
}

// ---- the coder as two chains ----
//
// ENCODE is two serial chains.  The interval side (context entries, the A
// register, the interval kept and the renormalisation shift) never reads the
// code register, and the code side (C, CT, BYTEOUT and the output bytes)
// needs from each symbol only its record from mq_interval.  ring_chain runs
// both sides in one wavefront (ROLE_BOTH: the records of a 16-symbol chunk in
// registers), or one side per wavefront of a two-wavefront workgroup, with the
// same chain in lane l of both (ROLE_INTERVAL in wavefront 0, ROLE_CODE in
// wavefront 1): each wavefront then runs only its side's instructions, one
// chunk apart, and a chain's time per symbol drops from the sum of the two
// sides to the longer one.  The records of chunk i pass through workgroup
// memory (buffer i & 1, 16 x 64 words); one barrier per chunk orders the
// writes of chunk i before its reads and the reads of chunk i - 1 before the
// buffer's reuse.  The final A follows the last chunk the same way.
constexpr int ROLE_BOTH = 0, ROLE_INTERVAL = 1, ROLE_CODE = 2;
constexpr int MQ_HAND_WORDS = 2 * 16 * 64 + 64;
// workgroup memory of the two-wavefront form: MQ_LDS_BYTES, then the records
constexpr int MQ2_LDS_BYTES = MQ_LDS_BYTES + MQ_HAND_WORDS * 4;

// Interval side of the chunk w4 (16 symbols from symbol `base`) into rec:
// the pipelined form when the whole chunk lies in [s0, s1) (`full`),
// otherwise symbol by symbol with records 0 (no change on the code side)
// outside the range.
RG_DEV void mq_interval_chunk(const uint32_t *tab, uint32_t *ce, uint32_t &A, const uint32_t (&w4)[4], bool full,
                              bool any, int base, int s0, int s1, uint32_t (&rec)[16])
{
    if (full) {
        uint32_t sy = w4[0] & 0xFFu;
        uint32_t ci = 64 * (sy >> 1);   // the context's word in ce
        uint32_t e = ce[ci];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t d = sy & 1u;
            const uint32_t tnew = tab[tt_index(d, e)];
            uint32_t sy2 = 0, ci2 = 0, e2 = 0;
            if (k < 15) {   // the next symbol's context, read before this one's store
                sy2 = (w4[(k + 1) >> 2] >> (8 * ((k + 1) & 3))) & 0xFFu;
                ci2 = 64 * (sy2 >> 1);
                e2 = ce[ci2];
            }
            wv_sched_fence();
            uint32_t enew;
            rec[k] = mq_interval(A, d, e, tnew, enew);
            ce[ci] = enew;
            if (k < 15) e = ci2 == ci ? enew : e2;
            sy = sy2;
            ci = ci2;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        rec[k] = 0;
        if (any && base + k >= s0 && base + k < s1) {
            const uint32_t sy = (w4[k >> 2] >> (8 * (k & 3))) & 0xFFu;
            const uint32_t cx = sy >> 1, d = sy & 1u;
            const uint32_t e = ce[64 * cx];
            uint32_t enew;
            rec[k] = mq_interval(A, d, e, tab[tt_index(d, e)], enew);
            ce[64 * cx] = enew;
        }
    }
}

// Codes symbols [s0, s1) of the 16-byte-aligned stream `sym` as one MQ
// codeword (INITENC, ENCODE per symbol, FLUSH C.2.9) into `out` (16-byte
// aligned, capacity cap), from the initial context states, in role ROLE (the
// two-wavefront roles are called by both wavefronts, each with its role).
// Every lane of a wavefront calls it; idle lanes pass active = false.
// Returns the codeword length (0 in ROLE_INTERVAL); ovf is set if it did not
// fit (out then holds its first complete 16-byte groups).  `tab`: the
// probability table in workgroup memory; `ring`: MQ_LDS_BYTES (ROLE_BOTH) or
// MQ2_LDS_BYTES of workgroup memory (4-byte aligned).
template <int ROLE>
RG_DEV int ring_chain(const uint32_t *tab, uint8_t *ring, bool active, const uint8_t *sym, int s0, int s1,
                      uint8_t *out, int cap, bool &ovf)
{
    constexpr bool IV = ROLE != ROLE_CODE, CD = ROLE != ROLE_INTERVAL;
    const int lane = wv_lane();
    uint8_t *rb = ring + lane * RING_STRIDE;
    uint32_t *ce = reinterpret_cast<uint32_t *>(ring + RING_BYTES) + lane;
    uint32_t *hand = reinterpret_cast<uint32_t *>(ring + MQ_LDS_BYTES) + lane;
    const int c0 = active ? s0 >> 4 : 0, c1 = active ? (s1 + 15) >> 4 : 0;
    const int nch = c1 - c0;
    int nmax = nch;
    for (int o = 32; o >= 1; o >>= 1) {
        const int t = wv_shfl_xor(nmax, o);
        nmax = nmax > t ? nmax : t;
    }
    uint32_t A = 0x8000u << 16;   // interval side: A << 16
    RingMQ m;
    m.A = 0x8000; m.C = 0; m.B = 0; m.CT = 12; m.bp = -1; m.fl = 0; m.ovf = 0;
    const int last = nch > 0 ? c1 - 1 : c0;
    u32x4 nxt = mk_u32x4(0, 0, 0, 0);
    if (IV) {
        ce_init(ce);
        nxt = wv_load16(sym + 16 * c0);
    }
    for (int i = 0; i < nmax; ++i) {
        uint32_t rec[16];
        if (IV) {
            const u32x4 cur = nxt;
            const int cn = c0 + i + 1;
            nxt = wv_load16(sym + 16 * (cn < last ? cn : last));
            const uint32_t w4[4] = {cur.x, cur.y, cur.z, cur.w};
            const int base = 16 * (c0 + i);
            const bool full = i < nch && base >= s0 && base + 16 <= s1;
            mq_interval_chunk(tab, ce, A, w4, full, i < nch, base, s0, s1, rec);
        }
        if (ROLE == ROLE_INTERVAL) {
#pragma unroll
            for (int k = 0; k < 16; ++k) hand[((i & 1) * 16 + k) * 64] = rec[k];
            wg_sync();
        }
        if (ROLE == ROLE_CODE) {
            wg_sync();
#pragma unroll
            for (int k = 0; k < 16; ++k) rec[k] = hand[((i & 1) * 16 + k) * 64];
        }
        if (CD) {
            // shifts of 8 or more (an LPS in a very skewed state) are rare: a
            // chunk without one in any lane takes the short form
            uint32_t big = 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) big |= rec[k];
            if (wv_any((big & 0x8u) != 0)) {
#pragma unroll
                for (int k = 0; k < 16; ++k) mq_code(m, rec[k], rb);
            } else {
#pragma unroll
                for (int k = 0; k < 16; ++k) mq_code7(m, rec[k], rb);
            }
            ring_flush(m, rb, out, cap);
        }
    }
    A >>= 16;
    if (ROLE == ROLE_INTERVAL) {
        hand[2 * 16 * 64] = A;
        wg_sync();
        ovf = false;
        return 0;
    }
    if (ROLE == ROLE_CODE) {
        wg_sync();
        A = hand[2 * 16 * 64];
    }
    // FLUSH (C.2.9); idle lanes run it on their untouched coder, unused
    m.A = A;
    const uint32_t tmp = m.C + m.A;
    m.C |= 0xFFFF;
    if (m.C >= tmp) m.C -= 0x8000;
    m.C <<= m.CT; ring_byteout(m, true, rb);
    m.C <<= m.CT; ring_byteout(m, true, rb);
    int len = m.bp;
    if (m.B != 0xFF) {
        rb[m.bp & 63] = (uint8_t)m.B;
        len = m.bp + 1;
    }
    m.bp = len;
    ring_flush(m, rb, out, cap);
    if (active && (len & 15) != 0 && (len & ~15) < cap) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(rb + ((len & ~15) & 63));
        wv_store16(out + (len & ~15), w[0], w[1], w[2], w[3]);
    }
    ovf = m.ovf || len > cap;
    return len;
}

// The coder in the workgroup's form: ROLE_BOTH with one wavefront (SPLIT
// false), else the role of this wavefront of two.
template <bool SPLIT>
RG_DEV int ring_chain_wg(const uint32_t *tab, uint8_t *ring, bool active, const uint8_t *sym, int s0, int s1,
                         uint8_t *out, int cap, bool &ovf)
{
    if (!SPLIT) return ring_chain<ROLE_BOTH>(tab, ring, active, sym, s0, s1, out, cap, ovf);
    if (wv_wave() == 0) return ring_chain<ROLE_INTERVAL>(tab, ring, active, sym, s0, s1, out, cap, ovf);
    return ring_chain<ROLE_CODE>(tab, ring, active, sym, s0, s1, out, cap, ovf);
}

// MQ-codes one block per lane from the streams the context-modelling pass
// wrote.  Block `i` (i >= njobs: idle lane) is either jobs[i] (QUEUE) or
// entry i of the flat arrays.  `tab` is MQ_TT_WORDS of workgroup-shared memory,
// `ring` MQ_LDS_BYTES (MQ2_LDS_BYTES with SPLIT: two wavefronts, each lane's
// block coded by lane l of both, ring_chain_wg) of it; the code side writes
// the results.
template <bool QUEUE, bool SPLIT = false>
RG_DEV void t1_mq_lanes(uint32_t *tab, uint8_t *ring, const T1Job *jobs, int i, int njobs, const uint8_t *syms,
                        int symcap, const i32x4 *cm, uint8_t *slots, int slot, rg_cb_result *res)
{
    const int lane = wv_lane();
    mq_tt_fill(tab);
    wg_sync();
    const bool active = i < njobs;
    i32x4 c = mk_i32x4(0, 0, 0, 0);
    // idle lanes read block 0's stream (always allocated) and store nothing
    const uint8_t *src = QUEUE ? jobs[0].sym : syms;
    uint8_t *out = nullptr;
    rg_cb_result *dst = nullptr;
    int cap = 0;
    if (active) {
        if (QUEUE) {
            const T1Job &j = jobs[i];
            c = *j.cm; src = j.sym; out = j.out; cap = j.cap; dst = j.res;
        } else {
            c = cm[i]; src = syms + (size_t)i * symcap; out = slots + (size_t)i * slot; cap = slot; dst = res + i;
        }
    }
    const bool coded = active && c.z > 0;   // a block without bit-planes has no codeword
    bool ovf;
    const int len = ring_chain_wg<SPLIT>(tab, ring, coded, src, 0, c.x, out, cap, ovf);
    if (!active || (SPLIT && wv_wave() == 0)) return;
    if (!coded) {
        *dst = rg_cb_result{0, 0, 0, 0};
        return;
    }
    *dst = rg_cb_result{len, c.y, c.z, (ovf || c.w) ? 1 : 0};
}

// ---- parallel style (RESET | TERMALL): one MQ chain per (code-block, pass) ----
//
// With both flags every coding pass starts from the initial context states
// and a freshly initialised coder and ends in a flush, so the passes of a
// block are independent.  A lane codes pass p of one block: symbols
// [poff[p], poff[p+1]) of its stream into the block's segment arena, and
// records the segment length; t1_pack_block then concatenates the segments.
// `len_out` < 0 marks a segment that did not fit its arena region.
// With SPLIT the workgroup's two wavefronts code the chain together
// (ring_chain_wg; `ring` MQ2_LDS_BYTES) and the code side writes the length.
template <bool SPLIT = false>
RG_DEV void t1_mq_pass_lane(uint32_t *tab, uint8_t *ring, bool active, const uint8_t *sym, int s0, int s1,
                            uint8_t *seg, int cap, int32_t *len_out)
{
    const int lane = wv_lane();
    mq_tt_fill(tab);
    wg_sync();
    bool ovf;
    const int len = ring_chain_wg<SPLIT>(tab, ring, active, sym, s0, s1, seg, cap, ovf);
    if (active && !(SPLIT && wv_wave() == 0)) *len_out = ovf ? -1 : len;
}

// Raw (bypass) pass of the parallel style with BYPASS (T.800 D.6): the
// decision bits of symbols [s0, s1) go out as they are, MSB first, with 7
// bits in the byte after an 0xFF (its MSB a stuffed 0); an unfinished byte,
// including the 7-bit byte owed after a final 0xFF, is padded with 0,1,0,...
// so the segment never ends in 0xFF.  Bits are taken 16 symbols at a time
// (the low bit of each symbol byte gathered with one multiply per word) into
// a bit accumulator; each step then completes up to three bytes (16 + 7
// pending bits, >= 7 bits a byte) with selects, writing every candidate byte
// to the lane's ring slot unconditionally (an unfinished one is rewritten
// before it is flushed).  `ring`: RING_BYTES of workgroup memory.
RG_DEV void t1_raw_pass_lane(uint8_t *ring, bool active, const uint8_t *sym, int s0, int s1, uint8_t *seg, int cap,
                             int32_t *len_out)
{
    uint8_t *rb = ring + wv_lane() * RING_STRIDE;
    const int c0 = active ? s0 >> 4 : 0, c1 = active ? (s1 + 15) >> 4 : 0;
    int nch = c1 - c0, nmax = nch;
    for (int o = 32; o >= 1; o >>= 1) {
        const int t = wv_shfl_xor(nmax, o);
        nmax = nmax > t ? nmax : t;
    }
    RingMQ m;
    m.A = 0; m.C = 0; m.B = 0; m.CT = 0; m.bp = 0; m.fl = 0; m.ovf = 0;
    uint32_t acc = 0;   // the low nacc bits are pending, oldest most significant
    int nacc = 0, need = 8;   // need: bits of the next byte (7 after an 0xFF)
    const int last = nch > 0 ? c1 - 1 : c0;
    // a step is short next to a load from HBM (the stream was written by the
    // context-modelling kernel and is not cached), so chunk c is requested
    // RAW_AHEAD steps before it is coded; the loop is unrolled by RAW_AHEAD
    // and each slot reloaded right after its chunk's bits are taken
    constexpr int RAW_AHEAD = 4;
    u32x4 pre[RAW_AHEAD];
#pragma unroll
    for (int a = 0; a < RAW_AHEAD; ++a) pre[a] = wv_load16(sym + 16 * (c0 + a < last ? c0 + a : last));
    for (int i = 0; i < nmax; i += RAW_AHEAD) {
#pragma unroll
        for (int a = 0; a < RAW_AHEAD; ++a) {
            const int step = i + a;
            // symbol k of the chunk -> bit 15 - k (first symbol most significant)
            const uint32_t g0 = ((pre[a].x & 0x01010101u) * 0x08040201u) >> 24;
            const uint32_t g1 = ((pre[a].y & 0x01010101u) * 0x08040201u) >> 24;
            const uint32_t g2 = ((pre[a].z & 0x01010101u) * 0x08040201u) >> 24;
            const uint32_t g3 = ((pre[a].w & 0x01010101u) * 0x08040201u) >> 24;
            const uint32_t b16 = (g0 << 12) | (g1 << 8) | (g2 << 4) | g3;
            const int cn = c0 + step + RAW_AHEAD;
            wv_sched_fence();   // the slot's old chunk is dead before its reload is issued
            pre[a] = wv_load16(sym + 16 * (cn < last ? cn : last));
            if (step < nmax) {
                const int base = 16 * (c0 + step);
                const int lo = s0 > base ? s0 - base : 0;
                const int hi = s1 - base < 16 ? s1 - base : 16;
                const int nb = step < nch && hi > lo ? hi - lo : 0;
                const uint32_t bits = (b16 >> (16 - hi < 16 ? 16 - hi : 0)) & ((1u << nb) - 1u);
                acc = (acc << nb) | bits;
                nacc += nb;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const bool emit = nacc >= need;
                    const uint32_t byte = (acc >> (emit ? nacc - need : 0)) & ((1u << need) - 1u);
                    rb[m.bp & 63] = (uint8_t)byte;
                    nacc -= emit ? need : 0;
                    acc &= (1u << nacc) - 1u;
                    m.bp += emit ? 1 : 0;
                    need = emit ? (byte == 0xFFu ? 7 : 8) : need;
                }
            }
        }
        // at most 4 x 3 new bytes since the last flush left < 16 pending
        ring_flush(m, rb, seg, cap);
    }
    // termination: pad the byte being filled (need - nacc free bits) with 0,1,0,...
    const int padn = need - nacc;
    if (padn < 8) {
        rb[m.bp & 63] = (uint8_t)(((acc << padn) | (0x55u >> (8 - padn))) & 0xFFu);
        m.bp++;
    }
    const int len = m.bp;
    ring_flush(m, rb, seg, cap);
    if (!active) return;
    if ((len & 15) != 0 && (len & ~15) < cap) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(rb + ((len & ~15) & 63));
        wv_store16(seg + (len & ~15), w[0], w[1], w[2], w[3]);
    }
    *len_out = (m.ovf || len > cap) ? -1 : len;
}

// Coding pass q of a block (grid row q): its pass index, symbol range and
// arena region; false if the block has fewer than q+1 passes.  (Counting rows
// from each block's last pass instead, so that a row holds the same bit-plane
// in every block, measured 12% slower for the 4096x4096 RGB workload.)
RG_DEV bool pass_range(const i32x4 &c, const int32_t *poff, int q, int &p, int &s0, int &s1, int &at, int &cap)
{
    if (q >= c.y) return false;
    p = q;
    s0 = poff[p];
    s1 = p + 1 < c.y ? poff[p + 1] : c.x;
    at = seg_at(s0, p);
    cap = 2 * (s1 - s0) + 16;
    return true;
}

// Workgroup memory of t1_pack_block: the output offset of every pass
// (MAXP + 1 words) and its segment's arena offset (MAXP words).
constexpr int PACK_WORDS = 2 * MAXP + 1;

// One wavefront per block: concatenates the pass segments into the block's
// output slot and writes the block result.  The passes' lengths are read at
// once and prefix-summed across the lanes, so the copy does not wait on a
// chain of per-pass loads: lane l copies output bytes l, l + 64, ... (four
// loads in flight), finding each byte's pass with a cursor that only moves
// forward.  Passes from the first overflowed segment (length < 0) on are
// dropped and the block flagged.  `sh`: PACK_WORDS words of workgroup memory.
RG_DEV void t1_pack_block(const i32x4 &c, const int32_t *poff, const uint8_t *seg, const int32_t *seglen,
                          uint8_t *out, int cap, rg_cb_result *res, int32_t *sh)
{
    static_assert(MAXP <= 128, "two passes per lane");
    const int lane = wv_lane();
    const int np = c.y;
    int32_t *pos = sh, *soff = sh + MAXP + 1;
    const int q1 = lane + 64;
    const int l0 = lane < np ? seglen[lane] : 0, l1 = q1 < np ? seglen[q1] : 0;
    const uint64_t bad0 = wv_ballot(l0 < 0), bad1 = wv_ballot(l1 < 0);
    const int pf = bad0 ? __builtin_ctzll(bad0) : (bad1 ? 64 + __builtin_ctzll(bad1) : np);
    const int n0 = lane < pf ? l0 : 0, n1 = q1 < pf ? l1 : 0;
    int t0, t1;
    const int e0 = wave_excl_count(n0, t0);
    const int e1 = wave_excl_count(n1, t1);
    const int total = t0 + t1;
    pos[lane] = e0;
    soff[lane] = lane < pf ? seg_at(poff[lane], lane) : 0;
    if (q1 < MAXP) {
        pos[q1] = t0 + e1;
        soff[q1] = q1 < pf ? seg_at(poff[q1], q1) : 0;
    }
    if (lane == 0) pos[MAXP] = total;
    wv_sync();
    const int lim = total < cap ? total : cap;
    int p = 0;   // this lane's current pass: pos[p] <= k < pos[p + 1] (k only grows)
    // Whole 16-byte output chunks first (16-byte aligned slot): a chunk inside
    // one pass is five aligned source dwords byte-aligned into four (segments
    // start 16-byte aligned with room after their end, so the reads past the
    // segment stay inside the arena); one across a pass boundary is gathered
    // bytewise.
    const int n16 = ((uintptr_t)out & 15) == 0 ? lim >> 4 : 0;
    constexpr int P16 = 4;   // chunks per lane in flight
    for (int j0 = 0; j0 < n16; j0 += 64 * P16) {
        uint32_t v[P16][4];
#pragma unroll
        for (int u = 0; u < P16; ++u) {
            const int j = j0 + 64 * u + lane, k = 16 * j;
#pragma unroll
            for (int i = 0; i < 4; ++i) v[u][i] = 0;
            if (j < n16) {
                while (pos[p + 1] <= k) ++p;
                if (k + 15 < pos[p + 1]) {
                    const int so = soff[p] + (k - pos[p]);
                    const uint32_t *w = reinterpret_cast<const uint32_t *>(seg + (so & ~3));
                    uint32_t a[5];
#pragma unroll
                    for (int i = 0; i < 5; ++i) a[i] = w[i];
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        v[u][i] = (uint32_t)((((uint64_t)a[i + 1] << 32) | a[i]) >> (8 * (so & 3)));
                } else {
                    // every source offset first (workgroup memory only), then
                    // the 16 byte loads together
                    int q = p, src[16];
#pragma unroll
                    for (int b = 0; b < 16; ++b) {
                        while (pos[q + 1] <= k + b) ++q;
                        src[b] = soff[q] + (k + b - pos[q]);
                    }
                    uint32_t by[16];
#pragma unroll
                    for (int b = 0; b < 16; ++b) by[b] = seg[src[b]];
#pragma unroll
                    for (int b = 0; b < 16; ++b) v[u][b >> 2] |= by[b] << (8 * (b & 3));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < P16; ++u) {
            const int j = j0 + 64 * u + lane;
            if (j < n16) {
                *reinterpret_cast<u32x4 *>(out + 16 * j) = mk_u32x4(v[u][0], v[u][1], v[u][2], v[u][3]);
            }
        }
    }
    // Then whole output dwords (4-byte aligned slot): a dword inside one pass
    // is two aligned source dwords funnel-shifted; one across a pass boundary
    // is gathered bytewise.
    const int nd = ((uintptr_t)out & 3) == 0 ? lim >> 2 : 0;
    for (int j0 = 4 * n16; j0 < nd; j0 += 256) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + 64 * u + lane, k = 4 * j;
            v[u] = 0;
            if (j < nd) {
                while (pos[p + 1] <= k) ++p;
                if (k + 3 < pos[p + 1]) {
                    const int so = soff[p] + (k - pos[p]);
                    const uint32_t *w = reinterpret_cast<const uint32_t *>(seg + (so & ~3));
                    v[u] = (uint32_t)((((uint64_t)w[1] << 32) | w[0]) >> (8 * (so & 3)));
                } else {
                    int q = p, src[4];
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        while (pos[q + 1] <= k + b) ++q;
                        src[b] = soff[q] + (k + b - pos[q]);
                    }
#pragma unroll
                    for (int b = 0; b < 4; ++b) v[u] |= (uint32_t)seg[src[b]] << (8 * b);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = j0 + 64 * u + lane;
            if (j < nd) reinterpret_cast<uint32_t *>(out)[j] = v[u];
        }
    }
    // the remaining bytes (all of them if the slot is not 4-byte aligned)
    for (int k0 = 4 * nd; k0 < lim; k0 += 256) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + 64 * u + lane;
            v[u] = 0;
            if (k < lim) {
                while (pos[p + 1] <= k) ++p;
                v[u] = seg[soff[p] + (k - pos[p])];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + 64 * u + lane;
            if (k < lim) out[k] = (uint8_t)v[u];
        }
    }
    const bool ovf = c.w || pf < np || total > cap;
    if (lane == 0) *res = rg_cb_result{c.z == 0 ? 0 : total, c.y, c.z, ovf ? 1 : 0};
    wv_sync();
}

// ---- parallel styles, sorted form: chains ordered longest first ----
//
// A (block, pass) chain's coding time is its symbol count, and a wavefront
// takes as long as its longest lane, so lanes are grouped by length: chains
// are bucketed by length (32-symbol buckets, longest first) into one list of
// MQ chains followed by one list of raw chains (a device counting sort), and
// wavefront w codes list entries [64w, 64w + 64).  The first wavefronts then
// carry the longest chains and start first.  Which lane codes a chain does
// not change its bytes.  Workspace (int32): CH_HIST histogram words (MQ
// buckets, then raw buckets), as many cursors, the two list lengths, then
// the list of chain ids (block * MAXP + pass).
constexpr int CH_BUCKETS = 256, CH_HIST = 2 * CH_BUCKETS, CH_COUNTS = 2 * CH_HIST, CH_LIST = CH_COUNTS + 16;

// Bucket of a chain: 0 = longest.
RG_DEV int chain_bucket(int len)
{
    const int k = len >> 5;
    return CH_BUCKETS - 1 - (k < CH_BUCKETS - 1 ? k : CH_BUCKETS - 1);
}

// Histogram key of chain (block c/poff, pass q), or -1 if the block has no
// pass q or q is outside [qlo, qhi).  Raw chains (BYPASS) take the second
// half of the keys.
RG_DEV int chain_key(const i32x4 &c, const int32_t *poff, int q, int style, int qlo = 0, int qhi = MAXP)
{
    int p, s0, s1, at, cap;
    if (q < qlo || q >= qhi || !pass_range(c, poff, q, p, s0, s1, at, cap)) return -1;
    return (pass_is_raw(style, q) ? CH_BUCKETS : 0) + chain_bucket(s1 - s0);
}

// Chain coder of the sorted form for list entry j of a wavefront whose
// entries are [base, end) (end <= base: idle lane); `raw` is uniform.
// J(b) returns block b's job record.  SPLIT: MQ chains coded by the two
// wavefronts of the workgroup together (t1_mq_pass_lane).
template <class J, bool SPLIT = false>
RG_DEV void t1_chain_lane(uint32_t *tab, uint8_t *ring, bool raw, const int32_t *list, int base, int end,
                          const J &job)
{
    const int j = base + wv_lane();
    bool active = j < end;
    const int id = list[active ? j : base];
    const T1Job &jb = job(id / MAXP);
    const int q = id % MAXP;
    int p = 0, s0 = 0, s1 = 0, at = 0, cap = 0;
    if (active) active = pass_range(*jb.cm, jb.poff, q, p, s0, s1, at, cap);
    if (raw)
        t1_raw_pass_lane(ring, active, jb.sym, s0, s1, jb.seg + at, cap, jb.seglen + p);
    else
        t1_mq_pass_lane<SPLIT>(tab, ring, active, jb.sym, s0, s1, jb.seg + at, cap, jb.seglen + p);
}

}  // namespace rg_t1
