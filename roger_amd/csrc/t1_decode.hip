// EBCOT tier-1 decoder (T.800 Annex D, MQ decoder of Annex C.3).
//
// The reference has no decoder at all (its OCLDecoder only runs the DWT
// stage, src/OCLDecoder.cpp:35-38, and even that is never called).  Design:
// one wavefront per code-block, lane j owns column j with its state as 64-bit
// row masks (significance, sign, visited, refined, decoded bit-plane) and its
// 64 magnitudes in registers (shift-accumulated MSB first, one bit-plane at a
// time).  Decoding is inherently sequential in scan order, so the coder and
// the context formation of the significance-propagation and cleanup passes
// run as wave-uniform scalar code: per stripe the wave walks the columns,
// pulling the 6-row windows of columns j-1, j, j+1 with readlane and writing
// a column's new state back with one lane-select.  The magnitude-refinement
// pass has no intra-pass dependency: its contexts are formed per lane in
// parallel and only the MQ decoding is sequential.  The codeword is staged
// in LDS once; the 19 context states and the 47-entry probability table sit
// in VGPR lanes.
#include "rg_internal.h"

namespace {

#define MQE(q, nm, nl, sw) ((uint32_t)(q) | ((uint32_t)(nm) << 16) | ((uint32_t)(nl) << 22) | ((uint32_t)(sw) << 28))
__constant__ uint32_t MQT_D[47] = {
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

constexpr int CTX_RL = 17, CTX_UNI = 18;
constexpr int DEC_LDS = 12288;   // codeword bytes staged per block (64x64 slot = 3 B/sample)

__device__ __forceinline__ int zc_ctx_d(int orient, int h, int v, int d)
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

__device__ __forceinline__ uint32_t bit_(uint32_t w, int i) { return (w >> i) & 1u; }

__device__ __forceinline__ uint32_t win6d(uint64_t m, int r0)
{
    return r0 == 0 ? (uint32_t)(m << 1) & 0x3Fu : (uint32_t)(m >> (r0 - 1)) & 0x3Fu;
}

struct MQDec {
    uint32_t A, C;
    int CT, bp, len;
    const uint8_t *buf;  // LDS copy of the codeword
};

__device__ __forceinline__ uint32_t mqd_byte(const MQDec &m, int i)
{
    return i < m.len ? (uint32_t)__builtin_amdgcn_readfirstlane((int)m.buf[i]) : 0xFFu;
}

__device__ __forceinline__ void mqd_bytein(MQDec &m)
{
    const uint32_t b = mqd_byte(m, m.bp), b1 = mqd_byte(m, m.bp + 1);
    if (b == 0xFF) {
        if (b1 > 0x8F) { m.C += 0xFF00; m.CT = 8; }
        else { m.bp++; m.C += b1 << 9; m.CT = 7; }
    } else {
        m.bp++; m.C += b1 << 8; m.CT = 8;
    }
}

__device__ __forceinline__ int mqd_decode(MQDec &m, int &ctxv, int mqtv, int cx, int lane)
{
    const int st = __builtin_amdgcn_readlane(ctxv, cx);
    const int mps = st & 1;
    const uint32_t t = (uint32_t)__builtin_amdgcn_readlane(mqtv, st >> 1);
    const uint32_t qe = t & 0xFFFF;
    int d, ns;
    m.A -= qe;
    if ((m.C >> 16) < qe) {
        if (m.A < qe) { m.A = qe; d = mps; ns = (int)(((t >> 16) & 63) << 1) | mps; }
        else { m.A = qe; d = 1 - mps; ns = (int)(((t >> 22) & 63) << 1) | (mps ^ (int)(t >> 28)); }
    } else {
        m.C -= qe << 16;
        if (m.A & 0x8000) return mps;
        if (m.A < qe) { d = 1 - mps; ns = (int)(((t >> 22) & 63) << 1) | (mps ^ (int)(t >> 28)); }
        else { d = mps; ns = (int)(((t >> 16) & 63) << 1) | mps; }
    }
    ctxv = lane == cx ? ns : ctxv;
    do {
        if (m.CT == 0) mqd_bytein(m);
        m.A <<= 1; m.C <<= 1; m.CT--;
    } while (!(m.A & 0x8000));
    return d;
}

// Sign context and XOR bit (Table D.3) from contribution sums.
__device__ __forceinline__ int sc_of(int H, int V, int &x)
{
    H = H > 1 ? 1 : (H < -1 ? -1 : H);
    V = V > 1 ? 1 : (V < -1 ? -1 : V);
    x = 0;
    if (H < 0 || (H == 0 && V < 0)) { x = 1; H = -H; V = -V; }
    return (H == 1) ? (V == 1 ? 13 : (V == 0 ? 12 : 11)) : (V == 1 ? 10 : 9);
}

__device__ __forceinline__ int contrib(uint32_t sig, uint32_t neg, int i)
{
    return bit_(sig, i) ? (bit_(neg, i) ? -1 : 1) : 0;
}

// Decodes the sign of window row i of the current column; returns 1 if negative.
__device__ __forceinline__ int decode_sign(MQDec &m, int &ctxv, int mqtv, int lane, int i, uint32_t Lw, uint32_t Lneg,
                                           uint32_t Cw, uint32_t Cneg, uint32_t Rw, uint32_t Rneg)
{
    const int H = contrib(Lw, Lneg, i) + contrib(Rw, Rneg, i);
    const int V = contrib(Cw, Cneg, i - 1) + contrib(Cw, Cneg, i + 1);
    int x;
    const int cx = sc_of(H, V, x);
    return mqd_decode(m, ctxv, mqtv, cx, lane) ^ x;
}

__device__ __forceinline__ int zc_at_d(int orient, int i, uint32_t Lw, uint32_t Cw, uint32_t Rw)
{
    const int h = (int)(bit_(Lw, i) + bit_(Rw, i));
    const int v = (int)(bit_(Cw, i - 1) + bit_(Cw, i + 1));
    const int d = (int)(bit_(Lw, i - 1) + bit_(Lw, i + 1) + bit_(Rw, i - 1) + bit_(Rw, i + 1));
    return zc_ctx_d(orient, h, v, d);
}

struct DecJob {
    const uint8_t *data;
    int len, npasses, nbp, w, h, orient, pitch;
    int32_t *dst;
    int style;              // code-block style: BYPASS (0x01) / RESET (0x02) / TERMALL (0x04) bits
    const int32_t *seglen;  // TERMALL: codeword segment length of each pass
};

// Raw (bypass) segment reader (T.800 D.6): bits MSB first, 7 in the byte
// after an 0xFF; past the segment end it reads 0xFF bytes, as the MQ decoder
// does.  Reuses MQDec: C holds the current byte, CT its unread bits.
__device__ __forceinline__ void raw_start(MQDec &m, const uint8_t *lds, int off, int len)
{
    m.buf = lds + off; m.len = len; m.bp = 0; m.C = 0; m.CT = 0; m.A = 0;
}

__device__ __forceinline__ int raw_bit(MQDec &m)
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

// One decision of the current pass: raw bit (bypass pass) or MQ symbol.
__device__ __forceinline__ int dec_bit(MQDec &m, bool raw, int &ctxv, int mqtv, int cx, int lane)
{
    return raw ? raw_bit(m) : mqd_decode(m, ctxv, mqtv, cx, lane);
}

// INITDEC (C.3.5) on the codeword segment [off, off + len) of the LDS copy.
__device__ __forceinline__ void mqd_start(MQDec &m, const uint8_t *lds, int off, int len)
{
    m.buf = lds + off; m.len = len; m.bp = 0;
    m.C = mqd_byte(m, 0) << 16;
    mqd_bytein(m);
    m.C <<= 7; m.CT -= 7; m.A = 0x8000;
}

__device__ void t1_decode_block(const DecJob &job, uint8_t *lds)
{
    const int lane = threadIdx.x;
    const int w = job.w, h = job.h, orient = job.orient, nbp = job.nbp;
    const int len = job.len < DEC_LDS ? job.len : DEC_LDS;
    for (int i = lane; i < len; i += 64) lds[i] = job.data[i];
    __syncthreads();

    uint32_t mag[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) mag[r] = 0;
    uint64_t S = 0, NEG = 0, V = 0, R = 0;
    const uint64_t rowmask = h >= 64 ? ~0ull : ((1ull << h) - 1);
    const int nstripes = (h + 3) >> 2;
    const int ctx0 = lane == 0 ? (4 << 1) : lane == CTX_RL ? (3 << 1) : lane == CTX_UNI ? (46 << 1) : 0;
    int ctxv = ctx0;
    const int mqtv = lane < 47 ? (int)MQT_D[lane] : 0;
    const bool termall = (job.style & 0x04) && job.seglen;

    MQDec m;
    int seg_off = 0;
    mqd_start(m, lds, 0, termall ? min(job.seglen[0], len) : len);

    int passes = 0;
    for (int p = nbp - 1; p >= 0; --p) {
        uint64_t B = 0;  // bit p of this lane's column
        for (int pass = (p == nbp - 1) ? 2 : 0; pass < 3 && passes < job.npasses; ++pass, ++passes) {
            // BYPASS (with TERMALL): SPP/MRP passes from index 10 on are raw segments
            const bool raw = termall && (job.style & 0x01) && passes >= 10 && pass < 2;
            if (passes > 0) {
                if (job.style & 0x02) ctxv = ctx0;      // RESET: initial context states
                if (termall) {                          // TERMALL: next codeword segment
                    seg_off += job.seglen[passes - 1];
                    const int rest = len - seg_off;
                    const int sl = job.seglen[passes];
                    const int o = seg_off < len ? seg_off : len, n = rest <= 0 ? 0 : (sl < rest ? sl : rest);
                    if (raw) raw_start(m, lds, o, n);
                    else mqd_start(m, lds, o, n);
                }
            }
            for (int s = 0; s < nstripes; ++s) {
                const int r0 = 4 * s;
                const uint32_t vrow = (uint32_t)(rowmask >> r0) & 0xFu;
                const uint32_t swv = win6d(S, r0), negv = win6d(NEG, r0);
                if (pass == 1) {
                    // ---- magnitude refinement: contexts in parallel, bits in order ----
                    uint32_t lw = __shfl_up(swv, 1, 64), rw = __shfl_down(swv, 1, 64);
                    if (lane == 0) lw = 0;
                    if (lane + 1 >= w) rw = 0;
                    const uint32_t vw = (uint32_t)(V >> r0) & 0xFu, rf = (uint32_t)(R >> r0) & 0xFu;
                    const uint32_t M = lane < w ? ((swv >> 1) & ~vw & vrow) : 0u;
                    uint32_t cxs = 0;  // 4 contexts x 5 bits
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int i = k + 1;
                        const uint32_t nb = bit_(swv, i - 1) | bit_(swv, i + 1) | ((lw >> (i - 1)) & 7u) |
                                            ((rw >> (i - 1)) & 7u);
                        const uint32_t cx = bit_(rf, k) ? 16u : (nb ? 15u : 14u);
                        cxs |= cx << (5 * k);
                    }
                    uint64_t mask = __ballot(M != 0);
                    while (mask) {
                        const int j = __builtin_ctzll(mask);
                        mask &= mask - 1;
                        const uint32_t Mj = (uint32_t)__builtin_amdgcn_readlane((int)M, j);
                        const uint32_t cj = (uint32_t)__builtin_amdgcn_readlane((int)cxs, j);
                        uint32_t bits = 0;
                        for (int k = 0; k < 4; ++k)
                            if (bit_(Mj, k) && dec_bit(m, raw, ctxv, mqtv, (int)((cj >> (5 * k)) & 31u), lane))
                                bits |= 1u << k;
                        if (lane == j) B |= (uint64_t)bits << r0;
                    }
                    R |= (uint64_t)M << r0;
                    continue;
                }
                // ---- significance propagation (pass 0) / cleanup (pass 2): sequential columns ----
                const uint32_t vwv = (uint32_t)(V >> r0) & 0xFu;
                const bool full = r0 + 4 <= h;
                uint32_t Lw = 0, Lneg = 0;
                uint32_t Cw = (uint32_t)__builtin_amdgcn_readlane((int)swv, 0);
                uint32_t Cneg = (uint32_t)__builtin_amdgcn_readlane((int)negv, 0);
                for (int j = 0; j < w; ++j) {
                    uint32_t Rw = 0, Rneg = 0;
                    if (j + 1 < w) {
                        Rw = (uint32_t)__builtin_amdgcn_readlane((int)swv, j + 1);
                        Rneg = (uint32_t)__builtin_amdgcn_readlane((int)negv, j + 1);
                    }
                    uint32_t vis = 0, nsig = 0;
                    if (pass == 0) {
                        if (((Lw | Cw | Rw) & 0x3Fu) != 0) {
                            for (int k = 0; k < 4; ++k) {
                                const int i = k + 1;
                                if (!bit_(vrow, k) || bit_(Cw, i)) continue;
                                const uint32_t nb = bit_(Cw, i - 1) | bit_(Cw, i + 1) | ((Lw >> (i - 1)) & 7u) |
                                                    ((Rw >> (i - 1)) & 7u);
                                if (!nb) continue;
                                vis |= 1u << k;
                                if (dec_bit(m, raw, ctxv, mqtv, zc_at_d(orient, i, Lw, Cw, Rw), lane)) {
                                    const int neg = raw ? raw_bit(m)
                                                        : decode_sign(m, ctxv, mqtv, lane, i, Lw, Lneg, Cw, Cneg, Rw,
                                                                      Rneg);
                                    Cw |= 1u << i;
                                    Cneg |= (uint32_t)neg << i;
                                    nsig |= 1u << k;
                                }
                            }
                        }
                    } else {
                        const uint32_t vj = (uint32_t)__builtin_amdgcn_readlane((int)vwv, j);
                        const uint32_t cand = ~(Cw >> 1) & ~vj & vrow & 0xFu;
                        int k0 = 0;
                        bool done = cand == 0;
                        if (!done && full && cand == 0xFu && ((Lw | Rw) & 0x3Fu) == 0 && (Cw & 0x21u) == 0) {
                            if (!mqd_decode(m, ctxv, mqtv, CTX_RL, lane)) {
                                done = true;
                            } else {
                                int k = mqd_decode(m, ctxv, mqtv, CTX_UNI, lane) << 1;
                                k |= mqd_decode(m, ctxv, mqtv, CTX_UNI, lane);
                                const int neg = decode_sign(m, ctxv, mqtv, lane, k + 1, Lw, Lneg, Cw, Cneg, Rw, Rneg);
                                Cw |= 1u << (k + 1);
                                Cneg |= (uint32_t)neg << (k + 1);
                                nsig |= 1u << k;
                                k0 = k + 1;
                            }
                        }
                        if (!done) {
                            for (int k = k0; k < 4; ++k) {
                                if (!bit_(cand, k)) continue;
                                const int i = k + 1;
                                if (mqd_decode(m, ctxv, mqtv, zc_at_d(orient, i, Lw, Cw, Rw), lane)) {
                                    const int neg = decode_sign(m, ctxv, mqtv, lane, i, Lw, Lneg, Cw, Cneg, Rw, Rneg);
                                    Cw |= 1u << i;
                                    Cneg |= (uint32_t)neg << i;
                                    nsig |= 1u << k;
                                }
                            }
                        }
                    }
                    if (vis | nsig) {
                        if (lane == j) {
                            V |= (uint64_t)vis << r0;
                            S |= (uint64_t)nsig << r0;
                            NEG |= (uint64_t)((Cneg >> 1) & nsig) << r0;
                            B |= (uint64_t)nsig << r0;
                        }
                    }
                    Lw = Cw; Lneg = Cneg;
                    Cw = Rw; Cneg = Rneg;
                }
            }
            if (pass == 2) V = 0;
        }
        // fold bit-plane p into the magnitudes (MSB first)
        uint32_t lo = (uint32_t)B, hi = (uint32_t)(B >> 32);
#pragma unroll
        for (int r = 0; r < 32; ++r) mag[r] = (mag[r] << 1) | ((lo >> r) & 1u);
#pragma unroll
        for (int r = 0; r < 32; ++r) mag[r + 32] = (mag[r + 32] << 1) | ((hi >> r) & 1u);
    }
    if (lane < w) {
#pragma unroll
        for (int r = 0; r < 64; ++r) {
            if (r < h) {
                const int v = (int)mag[r];
                job.dst[(size_t)r * job.pitch + lane] = ((NEG >> r) & 1) ? -v : v;
            }
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) t1_decode_kernel(const uint8_t *__restrict__ data, const long *__restrict__ offs,
                                                       const rg_cb_result *__restrict__ res,
                                                       const rg_cb_desc *__restrict__ desc, int32_t *__restrict__ q,
                                                       int pitch, size_t plane_stride, int style,
                                                       const int32_t *__restrict__ seglens, int maxp)
{
    __shared__ uint8_t lds[DEC_LDS];
    const rg_cb_desc d = desc[blockIdx.x];
    const rg_cb_result r = res[blockIdx.x];
    DecJob job;
    job.data = data + offs[blockIdx.x];
    job.len = r.len; job.npasses = r.npasses; job.nbp = r.nbp;
    job.w = d.w; job.h = d.h; job.orient = d.orient; job.pitch = pitch;
    job.dst = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
    job.style = style;
    job.seglen = seglens ? seglens + (size_t)blockIdx.x * maxp : nullptr;
    t1_decode_block(job, lds);
}

}  // namespace

int rg_launch_t1_decode(hipStream_t s, const uint8_t *data, const long *offs, const rg_cb_result *res,
                        const rg_cb_desc *desc, int ncb, int32_t *q, int pitch, size_t plane_stride, int style,
                        const int32_t *seglens)
{
    if (ncb <= 0) return RG_SUCCESS;
    t1_decode_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(data, offs, res, desc, q, pitch, plane_stride, style,
                                                              seglens, RG_MAX_PASSES);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}
