// EBCOT tier-1 decoder kernels.  The reference has no decoder at all (its
// OCLDecoder only runs the DWT stage, src/OCLDecoder.cpp:35-38, and even that
// is never called).
//
// Default: the wavefront decoder (t1_dec_core.h), one 64-thread workgroup per
// code-block whose decisions run as scalar code.  RG_T1_DEC=lane selects the
// lane decoder (t1_dec_lane.h): one code-block per lane with the block's state
// in workgroup memory, then a lane-per-column epilogue that folds the decoded
// bit-planes into the coefficients.  Both are bit-exact; the lane decoder
// moves the decisions onto the vector unit but its lanes diverge on every
// sample, and 4096x4096 RGB tier-1 decode takes 65 ms (bypass style) and
// 97 ms (standard style) against 31 and 61 ms with the wavefront decoder.
#include <cstdlib>
#include <cstring>
#include "rg_internal.h"
#include "wave_hip.h"
#include "t1_dec_core.h"
#include "t1_dec_lane.h"

using namespace rg_t1;

namespace {

// Six waves per SIMD (80 VGPRs, 136 bytes per lane spilled, against 121
// VGPRs and four waves): the scalar-bound decode interleaves more blocks per
// SIMD; 4096x4096 RGB tier-1 decode 32.1 -> 31.2 ms (bypass style).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6, 8))) t1_decode_kernel(const uint8_t *__restrict__ data, const long *__restrict__ offs,
                                                       const rg_cb_result *__restrict__ res,
                                                       const rg_cb_desc *__restrict__ desc, int32_t *__restrict__ q,
                                                       int pitch, size_t plane_stride, int style,
                                                       const int32_t *__restrict__ seglens, int maxp)
{
    __shared__ uint8_t lut[LUT_BYTES];
    __shared__ uint32_t rawbuf[DEC_RAW_WORDS];
    const rg_cb_desc d = desc[blockIdx.x];
    const rg_cb_result r = res[blockIdx.x];
    DecJob job;
    job.data = data + offs[blockIdx.x];
    job.len = r.len; job.npasses = r.npasses; job.nbp = r.nbp;
    job.w = d.w; job.h = d.h; job.orient = d.orient; job.pitch = pitch;
    job.dst = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
    job.style = style;
    job.seglen = seglens ? seglens + (size_t)blockIdx.x * maxp : nullptr;
    t1_dec_block(job, lut, rawbuf);
}

// ---------------- lane decoder: one code-block per lane ----------------
// Block state in workgroup memory, interleaved by lane so that the lanes'
// accesses to the same field and column fall on consecutive addresses.
template <int LANES>
struct LdsMem {
    uint64_t *st;          // [5][64][LANES]: S, N, V, R, B column masks
    uint8_t *cx;           // [19][LANES]: context states
    const uint8_t *lutp;   // the block's orientation's ZC / SC tables (512 bytes)
    const uint32_t *mq;    // MQ probability table
    int lane;
    RG_DEV uint64_t &f(int k, int c) { return st[(k * 64 + c) * LANES + lane]; }
    RG_DEV uint64_t &S(int c) { return f(0, c); }
    RG_DEV uint64_t &N(int c) { return f(1, c); }
    RG_DEV uint64_t &V(int c) { return f(2, c); }
    RG_DEV uint64_t &R(int c) { return f(3, c); }
    RG_DEV uint64_t &B(int c) { return f(4, c); }
    RG_DEV int ctx(int k) const { return cx[k * LANES + lane]; }
    RG_DEV void set_ctx(int k, int v) { cx[k * LANES + lane] = (uint8_t)v; }
    RG_DEV uint32_t mqt(int k) const { return mq[k]; }
    RG_DEV uint32_t lut(int x) const { return lutp[x]; }
};

// Codeword bytes through a 48-byte window of three aligned 16-byte chunks in
// registers.  Reads never go more than one byte backwards, so the window
// slides by a chunk once an offset reaches its last chunk, and the chunk
// after that is loaded then, 16 bytes (tens of symbols) before it is needed.
// A jump past the window (the rest of a segment left unread) reloads it.
// Only chunks holding a byte of the codeword are loaded (an aligned 16-byte
// chunk never crosses a page); get(i) is only called for i below the
// codeword length.
struct ChunkRd {
    const uint8_t *ab;     // codeword start aligned down to 16 bytes
    int skew, end;         // codeword start - ab; codeword end - ab
    int wo;                // window start (offset from ab, multiple of 16)
    uint4 w0, w1, w2;
    RG_DEV uint4 chunk(int off) const
    {
        if (off >= end) return make_uint4(~0u, ~0u, ~0u, ~0u);
        return *reinterpret_cast<const uint4 *>(ab + off);
    }
    RG_DEV void init(const uint8_t *base, int n)
    {
        ab = reinterpret_cast<const uint8_t *>((uintptr_t)base & ~(uintptr_t)15);
        skew = (int)(base - ab);
        end = skew + n;
        wo = 0;
        w0 = chunk(0);
        w1 = chunk(16);
        w2 = chunk(32);
    }
    static RG_DEV uint32_t pick(const uint4 &v, int o)
    {
        const int d = o >> 2;
        const uint32_t x = d == 0 ? v.x : (d == 1 ? v.y : (d == 2 ? v.z : v.w));
        return (x >> (8 * (o & 3))) & 0xFFu;
    }
    RG_DEV uint32_t get(int i)
    {
        const int a = i + skew;
        if (a >= wo + 32) {
            if (a < wo + 48) {
                wo += 16;
                w0 = w1;
                w1 = w2;
                w2 = chunk(wo + 32);
            } else {   // keep one chunk behind a, so a re-read of a - 1 stays in the window
                wo = (a & ~15) >= 16 ? (a & ~15) - 16 : 0;
                w0 = chunk(wo);
                w1 = chunk(wo + 16);
                w2 = chunk(wo + 32);
            }
        }
        const int o = a - wo;
        if (o < 16) return pick(w0, o);
        if (o < 32) return pick(w1, o - 16);
        return pick(w2, o - 32);
    }
};

constexpr int DEC_PLANES = 32;   // bit-plane slots per block in the workspace (nbp <= 31)

template <int LANES>
__global__ void __launch_bounds__(64) t1_decode_lane_kernel(const uint8_t *__restrict__ data,
                                                            const long *__restrict__ offs,
                                                            const rg_cb_result *__restrict__ res,
                                                            const rg_cb_desc *__restrict__ desc, int ncb, int style,
                                                            const int32_t *__restrict__ seglens, int maxp,
                                                            uint64_t *__restrict__ planes, uint64_t *__restrict__ negs)
{
    __shared__ uint8_t lut[4 * 512];
    __shared__ uint32_t mq[48];
    __shared__ uint64_t st[5 * 64 * LANES];
    __shared__ uint8_t cx[19 * LANES];
    for (int x = threadIdx.x; x < 4 * 512; x += 64) lut[x] = (uint8_t)lut_entry(x >> 9, x & 511);
    if (threadIdx.x < 47) mq[threadIdx.x] = MQT[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x;
    const int k = blockIdx.x * LANES + lane;
    if (lane >= LANES || k >= ncb) return;
    const rg_cb_desc d = desc[k];
    const rg_cb_result r = res[k];
    LaneJob job;
    // plane slots bound the bit-planes (T.800: at most 31 here; the callers check)
    job.len = r.len; job.npasses = r.npasses; job.nbp = r.nbp < 0 ? 0 : (r.nbp > DEC_PLANES - 1 ? DEC_PLANES - 1 : r.nbp);
    job.w = d.w; job.h = d.h; job.orient = d.orient;
    job.style = style;
    job.seglen = seglens ? seglens + (size_t)k * maxp : nullptr;
    job.planes = planes + (size_t)k * DEC_PLANES * 64;
    job.negs = negs + (size_t)k * 64;
    LdsMem<LANES> mem{st, cx, lut + 512 * (d.orient & 3), mq, lane};
    ChunkRd rd;
    rd.init(data + offs[k], r.len);
    t1_dec_lane(job, mem, rd);
}

// Epilogue: one wavefront per block, lane = column: the bit-planes folded
// MSB first into the magnitudes, signed, written as row runs.
__global__ void __launch_bounds__(64) t1_decode_fold_kernel(const rg_cb_result *__restrict__ res,
                                                            const rg_cb_desc *__restrict__ desc,
                                                            const uint64_t *__restrict__ planes,
                                                            const uint64_t *__restrict__ negs, int32_t *__restrict__ q,
                                                            int pitch, size_t plane_stride)
{
    const rg_cb_desc d = desc[blockIdx.x];
    const int nbp0 = res[blockIdx.x].nbp;
    const int nbp = nbp0 < 0 ? 0 : (nbp0 > DEC_PLANES - 1 ? DEC_PLANES - 1 : nbp0);
    const int c = threadIdx.x;
    if (c >= d.w) return;
    const uint64_t *pl = planes + (size_t)blockIdx.x * DEC_PLANES * 64 + c;
    uint32_t mag[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) mag[r] = 0;
    for (int p = nbp - 1; p >= 0; --p) {
        const uint64_t b = pl[p * 64];
        const uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
#pragma unroll
        for (int r = 0; r < 32; ++r) mag[r] = (mag[r] << 1) | ((lo >> r) & 1u);
#pragma unroll
        for (int r = 0; r < 32; ++r) mag[r + 32] = (mag[r + 32] << 1) | ((hi >> r) & 1u);
    }
    const uint64_t ng = negs[(size_t)blockIdx.x * 64 + c];
    int32_t *dst = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0 + c;
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        if (r < d.h) {
            const int v = (int)mag[r];
            dst[(size_t)r * pitch] = ((ng >> r) & 1) ? -v : v;
        }
    }
}

// Decoder selection, read on every launch: the wavefront decoder unless
// RG_T1_DEC=lane; RG_T1_DEC_LANES=8|12|16 code-blocks per lane-decoder
// workgroup (default 12: 33 KB of workgroup memory each; 8 and 16 measured
// 4 and 8% slower).
bool dec_lane()
{
    const char *e = getenv("RG_T1_DEC");
    return e && strcmp(e, "lane") == 0;
}

int dec_lanes()
{
    const char *e = getenv("RG_T1_DEC_LANES");
    const int n = e ? atoi(e) : 12;
    return (n == 8 || n == 16) ? n : 12;
}

}  // namespace

size_t rg_t1_decode_ws_bytes(int ncb) { return dec_lane() ? (size_t)ncb * (DEC_PLANES + 1) * 64 * 8 : 0; }

int rg_launch_t1_decode(hipStream_t s, const uint8_t *data, const long *offs, const rg_cb_result *res,
                        const rg_cb_desc *desc, int ncb, int32_t *q, int pitch, size_t plane_stride, int style,
                        const int32_t *seglens, void *ws)
{
    if (ncb <= 0) return RG_SUCCESS;
    if (!dec_lane()) {
        t1_decode_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(data, offs, res, desc, q, pitch, plane_stride,
                                                                  style, seglens, RG_MAX_PASSES);
        RG_LAUNCH_CHECK();
        return RG_SUCCESS;
    }
    void *w = ws;
    if (!w) RG_HIP(hipMallocAsync(&w, rg_t1_decode_ws_bytes(ncb), s));
    uint64_t *planes = reinterpret_cast<uint64_t *>(w), *negs = planes + (size_t)ncb * DEC_PLANES * 64;
    const int L = dec_lanes();
    const dim3 g((unsigned)((ncb + L - 1) / L)), b(64);
    if (L == 8)
        t1_decode_lane_kernel<8><<<g, b, 0, s>>>(data, offs, res, desc, ncb, style, seglens, RG_MAX_PASSES, planes,
                                                 negs);
    else if (L == 16)
        t1_decode_lane_kernel<16><<<g, b, 0, s>>>(data, offs, res, desc, ncb, style, seglens, RG_MAX_PASSES, planes,
                                                  negs);
    else
        t1_decode_lane_kernel<12><<<g, b, 0, s>>>(data, offs, res, desc, ncb, style, seglens, RG_MAX_PASSES, planes,
                                                  negs);
    RG_LAUNCH_CHECK();
    t1_decode_fold_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(res, desc, planes, negs, q, pitch, plane_stride);
    RG_LAUNCH_CHECK();
    if (!ws) RG_HIP(hipFreeAsync(w, s));
    return RG_SUCCESS;
}
