// Multi-level 2-D DWT: reversible 5/3 (int32) and irreversible 9/7 (FP32),
// forward with fused quantisation and inverse with fused dequantisation.
//
// Replaces src/CL/ocldwt53.cl:167-292, src/CL/ocldwt97.cl:196-420,487-717 and
// the (never-working) inverses src/CL/ocldwt53rev.cl, src/CL/ocldwt97rev.cl,
// driven per level as in src/OCLDWTForward.cpp:124-137.  Arithmetic follows
// T.800 Annex F exactly as the CPU oracle does: VER_SD (columns) then HOR_SD
// (rows) per level, whole-sample symmetric extension, a length-1 signal
// passes through, 9/7 low band x 1/K and high band x K, every 9/7 lifting
// step x[i] = fma(c, x[i-1] + x[i+1], x[i]).
//
// Kernel design (gfx950), forward: one 192-thread (three-wavefront)
// workgroup per 128x16 tile of the level's input region.
//   1. vertical pass straight from HBM: work item (column) loads its 16
//      rows plus a 4-row (9/7) / 2-row (5/3) halo with row-contiguous loads
//      (all 24 in flight), lifts in registers and stores the 16 results into
//      an LDS tile (16 rows x 136 columns, one pad word per 16 columns so
//      the row pass's 16-column segments fall on different banks);
//   2. horizontal pass from LDS: work item (row, 16-column segment) lifts in
//      registers, then writes its 8 low + 8 high outputs back de-interleaved;
//   3. each row goes out as 64 low + 64 high consecutive values (256-byte
//      runs per wavefront store) to their Mallat positions, quantised to int32
//      for the high bands, LL to the next level's buffer, with non-temporal
//      stores: nothing reads the outputs before they would leave the L2, and
//      streaming them past it lifts a read+write stream on this part from
//      about 6.0 to 7.3 TB/s (measured copy kernels).
// Short tiles keep every load of a workgroup in one round and give 12K+
// workgroups per 4096x4096 level; the vertical halo (8 rows per 16) is read
// again by the neighbouring tile, which the XCD-aware tile order places on
// the same L2.  Inverse: one 512-thread workgroup per 64x64 tile; the tile
// and its halo are first staged into LDS, interleaved and dequantised, with
// every load of a thread issued before any is used: interior tiles read each
// band row in aligned 16-byte quads (four samples of one band per load; per
// sample, the per-element address work and dequantisation made the staging
// the bottleneck), border tiles sample by sample with the band pointer
// chosen per lane; then the horizontal synthesis (T.800 2D_SR: rows first)
// and the vertical one, whose column items store 64-wide row runs
// (non-temporal).
//
// The native encoder's first level also takes the colour stage
// (fwd1_colour_kernel): the workgroup reads its tile of the interleaved
// uint8 image, level shifts and colour transforms each pixel once into LDS
// and lifts from there, so the FP32 colour planes never go to HBM (3 B/px
// read instead of 3 B/px + 12 B/px written + 12 B/px read again).  The
// decoder's last inverse level likewise ends in the inverse colour stage
// (inv1_colour_kernel) and writes the uint8 image itself.
#include <hip/hip_fp16.h>
#include <type_traits>
#include "colour_px.h"
#include "rg_internal.h"

namespace {

constexpr int SEG = 16;                    // samples a work item lifts in one run
constexpr int FTW = RG_DWT_FWD_TW, FTH = RG_DWT_FWD_TH, FWG = 192;   // forward tile and workgroup
constexpr int ITW = RG_DWT_INV_TW, ITH = RG_DWT_INV_TH, IWG = 512;   // inverse tile and workgroup

// LDS column of tile column c: one pad word after every 16 columns
__host__ __device__ constexpr int lcol(int c) { return c + (c >> 4); }

constexpr float A97 = (float)-1.586134342059924;
constexpr float B97 = (float)-0.052980118572961;
constexpr float G97 = (float)0.882911075530934;
constexpr float D97 = (float)0.443506852043971;
constexpr float K97 = (float)1.230174104914001;
constexpr float IK97 = (float)(1.0 / 1.230174104914001);

template <int F> struct Filt;
template <> struct Filt<RG_FILTER_53> { static constexpr int H = 2; using T = int; };
template <> struct Filt<RG_FILTER_97> { static constexpr int H = 4; using T = float; };

// ---- 1-D lifting on a register window e[0 .. SEG+2H) whose element 0 is at
//      an even signal position; outputs valid at [H, H+SEG). ----
template <int N>
__device__ __forceinline__ void fwd53(int (&e)[N])
{
#pragma unroll
    for (int j = 1; j <= N - 3; j += 2) e[j] -= (e[j - 1] + e[j + 1]) >> 1;
#pragma unroll
    for (int j = 2; j <= N - 4; j += 2) e[j] += (e[j - 1] + e[j + 1] + 2) >> 2;
}

template <int N>
__device__ __forceinline__ void fwd97(float (&e)[N], bool scale)
{
#pragma unroll
    for (int j = 1; j <= N - 3; j += 2) e[j] = __builtin_fmaf(A97, e[j - 1] + e[j + 1], e[j]);
#pragma unroll
    for (int j = 2; j <= N - 4; j += 2) e[j] = __builtin_fmaf(B97, e[j - 1] + e[j + 1], e[j]);
#pragma unroll
    for (int j = 3; j <= N - 5; j += 2) e[j] = __builtin_fmaf(G97, e[j - 1] + e[j + 1], e[j]);
#pragma unroll
    for (int j = 4; j <= N - 6; j += 2) e[j] = __builtin_fmaf(D97, e[j - 1] + e[j + 1], e[j]);
    if (scale) {
#pragma unroll
        for (int j = 4; j < N - 4; ++j) e[j] = (j & 1) ? e[j] * K97 : e[j] * IK97;
    }
}

template <int N>
__device__ __forceinline__ void inv53(int (&e)[N])
{
#pragma unroll
    for (int j = 2; j <= N - 2; j += 2) e[j] -= (e[j - 1] + e[j + 1] + 2) >> 2;
#pragma unroll
    for (int j = 3; j <= N - 3; j += 2) e[j] += (e[j - 1] + e[j + 1]) >> 1;
}

template <int N>
__device__ __forceinline__ void inv97(float (&e)[N], bool active)
{
    if (!active) return;
#pragma unroll
    for (int j = 0; j < N; ++j) e[j] = (j & 1) ? e[j] * IK97 : e[j] * K97;
#pragma unroll
    for (int j = 2; j <= N - 2; j += 2) e[j] = __builtin_fmaf(-D97, e[j - 1] + e[j + 1], e[j]);
#pragma unroll
    for (int j = 3; j <= N - 3; j += 2) e[j] = __builtin_fmaf(-G97, e[j - 1] + e[j + 1], e[j]);
#pragma unroll
    for (int j = 4; j <= N - 4; j += 2) e[j] = __builtin_fmaf(-B97, e[j - 1] + e[j + 1], e[j]);
#pragma unroll
    for (int j = 5; j <= N - 5; j += 2) e[j] = __builtin_fmaf(-A97, e[j - 1] + e[j + 1], e[j]);
}

// ---- element access helpers ----
template <int F>
__device__ __forceinline__ typename Filt<F>::T load_in(const void *p, long idx, int half)
{
    if constexpr (F == RG_FILTER_53) {
        return reinterpret_cast<const int *>(p)[idx];
    } else {
        if (half) return __half2float(reinterpret_cast<const __half *>(p)[idx]);
        return reinterpret_cast<const float *>(p)[idx];
    }
}

__device__ __forceinline__ int quant(float v, float inv)
{
    int m = (int)(fabsf(v) * inv);
    return v < 0.0f ? -m : m;
}

// sign(q) (|q| + 1/2) s for q != 0, written as (q + copysign(1/2, q)) s: the
// conversion, the sum and the product of IEEE arithmetic round the same for
// q and -q, so both forms give the same bits, with fewer instructions here
__device__ __forceinline__ float dequant(int q, float s)
{
    const float m = (float)q;
    return q == 0 ? 0.0f : (m + __builtin_copysignf(0.5f, m)) * s;
}

// ============================ forward ============================
template <int TW, int TH> constexpr int fwd_lds_bytes() { return TH * (lcol(TW + 8) + 1) * 4; }
template <int TW, int TH> constexpr int inv_lds_bytes() { return (TH + 8) * (lcol(TW + 8) + 1) * 4; }
constexpr int FWD_LDS_BYTES = fwd_lds_bytes<FTW, FTH>();   // largest (9/7) halo
constexpr int INV_LDS_BYTES = inv_lds_bytes<ITW, ITH>();

// Symmetric-extension index: one reflection then a clamp.  Exact for every
// position an output inside the signal depends on (overshoot <= halo < n);
// positions further out only feed outputs outside the signal, and the clamp
// keeps their loads in bounds.  Levels shorter than halo+1 use rg_mirror.
__device__ __forceinline__ int mirror1(int i, int n)
{
    i = i < 0 ? -i : (i >= n ? 2 * n - 2 - i : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// Streaming (non-temporal) stores of the level outputs.
__device__ __forceinline__ void st_nt(int *p, int v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(__half *p, float v)
{
    __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v)), reinterpret_cast<unsigned short *>(p));
}

// Output sample v of tile row y (level coordinates), column xo of its half
// band (high: odd source column): LL to the next level's buffer, the rest
// quantised (9/7) or as is into the Mallat plane.
template <int F>
__device__ __forceinline__ void fwd_put(const rg_fwd_level_args &a, int plane, char *qbase, int y, int xo, bool high,
                                        int wl, int hl, typename Filt<F>::T v)
{
    const bool lowrow = !(y & 1);
    if (lowrow && !high && a.ll_out) {
        const long off = (long)plane * a.ll_ps + (long)(y >> 1) * a.ll_pitch + xo;
        if constexpr (F == RG_FILTER_53) st_nt(reinterpret_cast<int *>(a.ll_out) + off, v);
        else if (a.ll_half) st_nt(reinterpret_cast<__half *>(a.ll_out) + off, v);
        else st_nt(reinterpret_cast<float *>(a.ll_out) + off, v);
        return;
    }
    const long off = (long)(lowrow ? (y >> 1) : hl + (y >> 1)) * a.qp + (high ? wl + xo : xo);
    if constexpr (F == RG_FILTER_53) {
        st_nt(reinterpret_cast<int *>(qbase) + off, v);
    } else if (a.q_float) {
        st_nt(reinterpret_cast<float *>(qbase) + off, v);
    } else {
        // sign(v) * floor(|v| * inv) == truncation of v * inv (IEEE products are sign-symmetric)
        const float inv = lowrow ? (high ? a.inv_hl : a.inv_ll) : (high ? a.inv_hh : a.inv_lh);
        st_nt(reinterpret_cast<int *>(qbase) + off, (int)(v * inv));
    }
}

// One TW x TH tile (bx, by) of plane `plane`; `lds` holds fwd_lds_bytes.
template <int F, int TW, int TH, int WGS>
__device__ __forceinline__ void fwd_tile(const rg_fwd_level_args &a, int plane, int bx, int by, char *lds)
{
    using T = typename Filt<F>::T;
    constexpr int H = Filt<F>::H;
    constexpr int CW = TW + 2 * H;            // tile columns with halo
    constexpr int LP = lcol(CW) + 1;          // LDS row pitch
    constexpr int N = SEG + 2 * H;            // register window
    constexpr int VS = TH / SEG, HS = TW / SEG;
    constexpr int RS = WGS / TW;              // rows per store round
    static_assert(TH % SEG == 0 && TW % SEG == 0 && TH * HS <= WGS && RS >= 1, "tile / workgroup shape");
    T *tile = reinterpret_cast<T *>(lds);

    const int x0 = bx * TW, y0 = by * TH;
    const int cw = a.cw, ch = a.ch;
    const void *in = reinterpret_cast<const char *>(a.in) +
                     (long)plane * a.in_ps * (F == RG_FILTER_53 ? 4 : (a.in_half ? 2 : 4));
    const bool vlift = ch > 1, hlift = cw > 1;
    const bool simple = cw > H && ch > H;      // one reflection suffices
    // the tile and its halo lie inside the level: no extension, no bounds tests
    const bool interior = x0 >= H && y0 >= H && x0 + TW + H <= cw && y0 + TH + H <= ch;
    const bool small_plane = (long)ch * a.in_pitch * 4 < (1L << 31);   // byte offsets fit 32 bits

    // ---- vertical pass straight from HBM: (column c of CW, segment k) ----
    for (int item = threadIdx.x; item < CW * VS; item += WGS) {
        const int c = item % CW, k = item / CW;
        const int x = simple ? mirror1(x0 - H + c, cw) : rg_mirror(x0 - H + c, cw);
        const int ybase = y0 + k * SEG - H;
        T e[N];
        if (interior) {
            const long first = (long)ybase * a.in_pitch + x;
            if (F == RG_FILTER_97 && a.in_half) {
#pragma unroll
                for (int j = 0; j < N; ++j) e[j] = load_in<F>(in, first + (long)j * a.in_pitch, 1);
            } else if (small_plane) {
                // 32-bit byte offsets from the uniform plane base: one add per load
                const char *pb = reinterpret_cast<const char *>(in);
                const uint32_t b0 = (uint32_t)first * (uint32_t)sizeof(T), pitch_b = (uint32_t)a.in_pitch * sizeof(T);
#pragma unroll
                for (int j = 0; j < N; ++j) e[j] = *reinterpret_cast<const T *>(pb + (b0 + (uint32_t)j * pitch_b));
            } else {
                const T *ip = reinterpret_cast<const T *>(in) + first;
#pragma unroll
                for (int j = 0; j < N; ++j) e[j] = ip[(long)j * a.in_pitch];
            }
        } else if (simple) {
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = load_in<F>(in, (long)mirror1(ybase + j, ch) * a.in_pitch + x, a.in_half);
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = load_in<F>(in, (long)rg_mirror(ybase + j, ch) * a.in_pitch + x, a.in_half);
        }
        if (vlift) {
            if constexpr (F == RG_FILTER_53) fwd53(e); else fwd97(e, true);
        }
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[(k * SEG + j) * LP + lcol(c)] = e[H + j];
    }
    __syncthreads();

    // ---- horizontal pass from LDS: (segment kk, row rr), kk fastest ----
    T o[SEG];
    const int kk = threadIdx.x % HS, rr = threadIdx.x / HS;
    const bool rowt = threadIdx.x < TH * HS;
    if (rowt) {
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j) e[j] = tile[rr * LP + lcol(kk * SEG + j)];
        if (hlift) {
            if constexpr (F == RG_FILTER_53) fwd53(e); else fwd97(e, true);
        }
#pragma unroll
        for (int j = 0; j < SEG; ++j) o[j] = e[H + j];
    }
    __syncthreads();
    // de-interleave the row: low outputs to [0, TW/2), high to [TW/2, TW)
    if (rowt) {
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[rr * LP + lcol((j & 1) * (TW / 2) + kk * (SEG / 2) + (j >> 1))] = o[j];
    }
    __syncthreads();

    // ---- write out: thread = column c of the de-interleaved tile, rows r0, r0+RS, ... ----
    const int wl = (cw + 1) >> 1, hl = (ch + 1) >> 1;
    char *qbase = reinterpret_cast<char *>(a.q) + (long)plane * a.q_ps * 4;
    if (threadIdx.x >= RS * TW) return;
    const int c = threadIdx.x % TW, r0 = threadIdx.x / TW;
    const bool high = c >= TW / 2;
    const int xo = (x0 >> 1) + (c & (TW / 2 - 1));       // column inside its half band
    const int x = 2 * xo + (high ? 1 : 0);               // source column
    if (interior) {
        for (int r = r0; r < TH; r += RS) fwd_put<F>(a, plane, qbase, y0 + r, xo, high, wl, hl, tile[r * LP + lcol(c)]);
        return;
    }
    if (x >= cw) return;
    for (int r = r0; r < TH && y0 + r < ch; r += RS)
        fwd_put<F>(a, plane, qbase, y0 + r, xo, high, wl, hl, tile[r * LP + lcol(c)]);
}

// The common 9/7 levels (FP32 or FP16 input, FP32 or FP16 LL output or LL
// quantised with the other bands, quantised indices, both sides longer than
// the 4-sample halo): fwd_tile's passes with the formats fixed at compile
// time.  Its leaner code measured 8-10% faster on 4096x4096x3 level 1 with
// inputs not in any cache (92 -> 85 us).
template <int TW, int TH, int WGS, bool IN_HALF, bool LL_HALF>
__device__ __forceinline__ void fwd_tile_plain97(const rg_fwd_level_args &a, int plane, int bx, int by, char *lds)
{
    using T = float;
    constexpr int H = 4;
    constexpr int CW = TW + 2 * H;
    constexpr int LP = lcol(CW) + 1;
    constexpr int N = SEG + 2 * H;
    constexpr int VS = TH / SEG, HS = TW / SEG;
    constexpr int RS = WGS / TW;
    static_assert(TH % SEG == 0 && TW % SEG == 0 && TH * HS <= WGS && RS >= 1, "tile / workgroup shape");
    T *tile = reinterpret_cast<T *>(lds);
    const int x0 = bx * TW, y0 = by * TH;
    const int cw = a.cw, ch = a.ch;
    using In = typename std::conditional<IN_HALF, __half, float>::type;
    const In *in = reinterpret_cast<const In *>(a.in) + (long)plane * a.in_ps;
    const bool interior = x0 >= H && y0 >= H && x0 + TW + H <= cw && y0 + TH + H <= ch;
    for (int item = threadIdx.x; item < CW * VS; item += WGS) {
        // items [0, TW): the tile's own columns, so that a wavefront's loads
        // start on the tile's (aligned) first column; [TW, TW + H): the left
        // halo; [TW + H, CW): the right halo
        const int i = item % CW, k = item / CW;
        const int c = i < TW ? i + H : (i < TW + H ? i - TW : i);
        const int x = mirror1(x0 - H + c, cw);
        const int ybase = y0 + k * SEG - H;
        T e[N];
        if (interior) {
            const long first = (long)ybase * a.in_pitch + x;
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = (float)in[first + (long)j * a.in_pitch];
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = (float)in[(long)mirror1(ybase + j, ch) * a.in_pitch + x];
        }
        fwd97(e, true);
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[(k * SEG + j) * LP + lcol(c)] = e[H + j];
    }
    __syncthreads();
    T o[SEG];
    const int kk = threadIdx.x % HS, rr = threadIdx.x / HS;
    const bool rowt = threadIdx.x < TH * HS;
    if (rowt) {
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j) e[j] = tile[rr * LP + lcol(kk * SEG + j)];
        fwd97(e, true);
#pragma unroll
        for (int j = 0; j < SEG; ++j) o[j] = e[H + j];
    }
    __syncthreads();
    if (rowt) {
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[rr * LP + lcol((j & 1) * (TW / 2) + kk * (SEG / 2) + (j >> 1))] = o[j];
    }
    __syncthreads();
    const int wl = (cw + 1) >> 1, hl = (ch + 1) >> 1;
    int *qbase = reinterpret_cast<int *>(a.q) + (long)plane * a.q_ps;
    using Ll = typename std::conditional<LL_HALF, __half, float>::type;
    Ll *llbase = a.ll_out ? reinterpret_cast<Ll *>(a.ll_out) + (long)plane * a.ll_ps : nullptr;
    if (threadIdx.x >= RS * TW) return;
    const int c = threadIdx.x % TW, r0 = threadIdx.x / TW;
    const bool high = c >= TW / 2;
    const int xo = (x0 >> 1) + (c & (TW / 2 - 1));
    const int x = 2 * xo + (high ? 1 : 0);
    for (int r = r0; r < TH; r += RS) {
        const int y = y0 + r;
        if (!interior && (y >= ch || x >= cw)) continue;
        const T v = tile[r * LP + lcol(c)];
        const bool lowrow = !(y & 1);
        if (lowrow && !high && a.ll_out) {
            st_nt(llbase + (long)(y >> 1) * a.ll_pitch + xo, v);
            continue;
        }
        // sign(v) * floor(|v| * inv) == truncation of v * inv (IEEE products are sign-symmetric)
        const float inv = lowrow ? (high ? a.inv_hl : a.inv_ll) : (high ? a.inv_hh : a.inv_lh);
        const long off = (long)(lowrow ? (y >> 1) : hl + (y >> 1)) * a.qp + (high ? wl + xo : xo);
        st_nt(qbase + off, (int)(v * inv));
    }
}

// ==================== first level fused with the colour stage ====================
// One TW x SEG tile of the interleaved uint8 image, every component:
//   1. staging: items (row, 4-pixel quad) over the tile's rows with the
//      vertical halo (mirrored rows loaded in their place) and columns
//      [x0 - 4, x0 + TW + 4) clipped to the image (quads start 4-pixel
//      aligned, so a quad is NC aligned dwords); every load of a thread is
//      issued before any is used; each pixel is level shifted and colour
//      transformed once (rg_px::px_fwd, the colour kernels' arithmetic) and
//      its components go to NC float / int planes in LDS; the mirrored halo
//      columns at the image sides lie inside that span;
//   2. column items (component, column): read the column from the staged
//      plane, lift it in registers, and -- after a barrier, all reads done --
//      write the SEG results over the staging as component k's tile;
//   3. row items (component, row, 16-column segment) lift and de-interleave
//      as fwd_tile does;
//   4. stores: thread = one column of the de-interleaved tile and one row of
//      every RS, over every component; a wavefront stores 64 consecutive
//      outputs (non-temporal).
// The level kernels are bound by VALU issue about as much as by HBM (lifting
// alone is ~27 operations per sample and component with the halos), so the
// items are laid out for full wavefronts and nothing is converted twice.
// LDS: NC x (SEG + 2H) x LPS words.
template <int TW> constexpr int fwd1c_lps() { return lcol(TW + 8) + 1; }
template <int F, int NC, int TW, int TH> constexpr int fwd1c_lds_bytes()
{
    return NC * (TH + 2 * Filt<F>::H) * fwd1c_lps<TW>() * 4;
}

template <int F, int NC, int MODE, int TW, int TH, int WGS>
__device__ __forceinline__ void fwd1_colour_tile(const rg_fwd1c_args &a, int bx, int by, char *lds)
{
    using T = typename Filt<F>::T;
    constexpr int H = Filt<F>::H;
    constexpr int CW = TW + 2 * H, RH = TH + 2 * H;
    constexpr int SW = TW + 8;                           // staged columns: image columns x0 - 4 .. x0 + TW + 4
    constexpr int LPS = fwd1c_lps<TW>();                 // staged and tile row pitch (words)
    constexpr int N = SEG + 2 * H;                       // a row item's window
    constexpr int NV = TH + 2 * H;                       // a column item's window
    constexpr int HS = TW / SEG;
    constexpr int RS = WGS / TW;                         // rows per store round
    static_assert(TW % 64 == 0 && WGS % TW == 0 && TH % RS == 0 && TH % SEG == 0 && NC * CW <= WGS,
                  "fused tile / workgroup shape");
    T *fst = reinterpret_cast<T *>(lds);                 // component k, staged row r at (k * RH + r) * LPS
    T *tile = fst;                                       // then component k, row j at (k * TH + j) * LPS
    const int x0 = bx * TW, y0 = by * TH;
    const int w = a.w, h = a.h;
    const int xs = max(x0 - 4, 0), xe = min(x0 + TW + 4, w);   // staged span (both multiples of 4)
    const long rowb = (long)w * NC;

    // 1. staging: quads (row r, quad q), image column x0 - 4 + 4q; item i of
    //    round t is threadIdx.x + t * WGS, its (r, q) stepped from round 0's
    constexpr int NQ = SW / 4, NIT = RH * NQ, NI = (NIT + WGS - 1) / WGS;
    constexpr int DR = WGS / NQ, DQ = WGS % NQ;
    int rq[NI][2];
    rq[0][0] = threadIdx.x / NQ;
    rq[0][1] = threadIdx.x - rq[0][0] * NQ;
#pragma unroll
    for (int t = 1; t < NI; ++t) {
        const int q = rq[t - 1][1] + DQ;
        rq[t][0] = rq[t - 1][0] + DR + (q >= NQ ? 1 : 0);
        rq[t][1] = q >= NQ ? q - NQ : q;
    }
    // byte offsets fit 32 bits for images below 2 GiB (the launcher checks)
    const uint32_t rowb32 = (uint32_t)rowb;
    uint32_t wq[NI][NC];
#pragma unroll
    for (int t = 0; t < NI; ++t) {
        const int i = threadIdx.x + t * WGS;
        const int r = rq[t][0], q = rq[t][1];
        const int x = x0 - 4 + 4 * q;
#pragma unroll
        for (int k = 0; k < NC; ++k) wq[t][k] = 0;
        if (i < NIT && x >= xs && x < xe) {
            const uint32_t off = (uint32_t)mirror1(y0 - H + r, h) * rowb32 + (uint32_t)(x * NC);
            const uint32_t *p = reinterpret_cast<const uint32_t *>(a.src + off);
#pragma unroll
            for (int k = 0; k < NC; ++k) wq[t][k] = p[k];
        }
    }
#pragma unroll
    for (int t = 0; t < NI; ++t) {
        const int i = threadIdx.x + t * WGS;
        const int r = rq[t][0], q = rq[t][1];
        const int x = x0 - 4 + 4 * q;
        if (i < NIT && x >= xs && x < xe) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                T o[NC];
                if constexpr (F == RG_FILTER_97) {
                    // byte to float in one conversion, then the level shift (exact)
                    float v[NC];
#pragma unroll
                    for (int k = 0; k < NC; ++k)
                        v[k] = (float)((wq[t][(NC * j + k) >> 2] >> (((NC * j + k) & 3) * 8)) & 0xFFu) - 128.0f;
                    rg_px::px_fwd_f<NC, MODE>(v, o);
                } else {
                    int v[NC];
#pragma unroll
                    for (int k = 0; k < NC; ++k)
                        v[k] = (int)((wq[t][(NC * j + k) >> 2] >> (((NC * j + k) & 3) * 8)) & 0xFFu) - 128;
                    rg_px::px_fwd<NC, MODE, false>(v, o);
                }
#pragma unroll
                for (int k = 0; k < NC; ++k) fst[(k * RH + r) * LPS + lcol(4 * q + j)] = o[k];
            }
        }
    }
    __syncthreads();

    // 2. columns: items (component, column), component slowest; one round
    const int k2 = threadIdx.x / CW, c2 = threadIdx.x - k2 * CW;
    const bool colt = threadIdx.x < NC * CW;
    T e[NV];
    if (colt) {
        // every column an output inside the image needs mirrors into [xs, xe);
        // the clamp only keeps the reads of the other columns in the span
        const int x = min(max(mirror1(x0 - H + c2, w), xs), xe - 1);
        const T *col = fst + k2 * RH * LPS + lcol(x - (x0 - 4));
#pragma unroll
        for (int r = 0; r < NV; ++r) e[r] = col[r * LPS];
        if constexpr (F == RG_FILTER_53) fwd53(e); else fwd97(e, true);
    }
    __syncthreads();
    if (colt) {
#pragma unroll
        for (int j = 0; j < TH; ++j) tile[(k2 * TH + j) * LPS + lcol(c2)] = e[H + j];
    }
    __syncthreads();

    // 3. rows: items (component x row, segment), segment fastest
    constexpr int NROW = NC * TH * HS;
    constexpr int RR = (NROW + WGS - 1) / WGS;
    T o[RR][SEG];
#pragma unroll
    for (int u2 = 0; u2 < RR; ++u2) {
        const int it = threadIdx.x + u2 * WGS;
        if (it < NROW) {
            const int kk = it % HS, rr = it / HS;
            T er[N];
#pragma unroll
            for (int j = 0; j < N; ++j) er[j] = tile[rr * LPS + lcol(kk * SEG + j)];
            if constexpr (F == RG_FILTER_53) fwd53(er); else fwd97(er, true);
#pragma unroll
            for (int j = 0; j < SEG; ++j) o[u2][j] = er[H + j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u2 = 0; u2 < RR; ++u2) {
        const int it = threadIdx.x + u2 * WGS;
        if (it < NROW) {
            const int kk = it % HS, rr = it / HS;
#pragma unroll
            for (int j = 0; j < SEG; ++j)
                tile[rr * LPS + lcol((j & 1) * (TW / 2) + kk * (SEG / 2) + (j >> 1))] = o[u2][j];
        }
    }
    __syncthreads();

    // 4. stores.  Wavefront w takes half band (w & 1) of the de-interleaved
    //    tile and rows of one parity ((w >> 1) & 1; y0 is even), so its band,
    //    quantiser and destination are uniform; its lanes cover LPH = TW / 2
    //    columns of HPW = 64 / LPH rows, so a store instruction writes one or
    //    two runs of LPH consecutive outputs.  Round j covers RPR rows of each
    //    half band.
    constexpr int LPH = TW / 2, HPW = 64 / LPH, NW = WGS / 64;
    constexpr int RPR = (NW / 2) * HPW;
    static_assert(NW % 4 == 0 && TH % RPR == 0, "store mapping");
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int half = wv & 1, grp = wv >> 1;
    const int par = grp & 1, gsub = grp >> 1;        // row parity; group among the waves of that parity
    const int cc = lane % LPH, rsub = lane / LPH;
    const int wl = (w + 1) >> 1, hl = (h + 1) >> 1;
    const int xo = (x0 >> 1) + cc;
    const bool col_in = 2 * xo + half < w;
    const bool lowrow = par == 0;
    const bool to_ll = lowrow && !half && a.ll_out;
    const long pitch = to_ll ? a.ll_pitch : a.qp;
    const float inv = lowrow ? (half ? a.inv_hl : a.inv_ll) : (half ? a.inv_hh : a.inv_lh);
    // output row of round j: (y0 + r) / 2 with r = par + 2 * (gsub * HPW + rsub + j * RPR / 2)
    const int orow0 = (y0 >> 1) + gsub * HPW + rsub;
    const long first = to_ll ? (long)orow0 * pitch + xo
                             : (long)(lowrow ? orow0 : hl + orow0) * pitch + (half ? wl + xo : xo);
    // every read of the tile first (all inside the LDS tile, so none needs a
    // guard), then the stores: one branch on the destination for the whole
    // wavefront, and no per-store bounds test on tiles inside the image
    constexpr int NJ = TH / RPR;
    T v[NC][NJ];
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
            v[k][j] = tile[(k * TH + par + 2 * (gsub * HPW + rsub + j * (RPR / 2))) * LPS + lcol(half * (TW / 2) + cc)];
    const bool inner = x0 + TW <= w && y0 + TH <= h;
    const int rlim = h - y0 - par - 2 * (gsub * HPW + rsub);   // round j stores while j * RPR < rlim
    const long step = (long)(RPR / 2) * pitch;
    auto put = [&](auto guarded) {
        constexpr bool G = decltype(guarded)::value;
        if (to_ll) {
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                T *llk = reinterpret_cast<T *>(a.ll_out) + k * a.ll_ps + first;
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                    if (!G || (col_in && j * RPR < rlim)) st_nt(llk + j * step, v[k][j]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < NC; ++k) {
                int32_t *qk = a.q + k * a.q_ps + first;
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    if (G && !(col_in && j * RPR < rlim)) continue;
                    if constexpr (F == RG_FILTER_53) {
                        st_nt(qk + j * step, v[k][j]);
                    } else {
                        // sign(v) * floor(|v| * inv) == truncation of v * inv (IEEE products are sign-symmetric)
                        st_nt(qk + j * step, (int)(v[k][j] * inv));
                    }
                }
            }
        }
    };
    if (inner) put(std::false_type{});
    else put(std::true_type{});
}

// ============================ inverse ============================
// One TW x TH tile of the level region: stage the tile and its halo (rows
// y0-H .. y0+TH+H, columns x0-H .. x0+TW+H, symmetric extension) from the
// four bands, dequantised, into LDS in interleaved order; then rows, then
// columns.
// PLAIN: the common 9/7 level (LL from its FP32 buffer, quantised bands,
// both sides longer than the halo), its format choices fixed at compile time.
template <int F, int TW, int TH, int WGS, bool PLAIN = false>
__device__ __forceinline__ void inv_tile(const rg_inv_level_args &a, int plane, int bx, int by, char *lds)
{
    using T = typename Filt<F>::T;
    constexpr int H = Filt<F>::H;
    constexpr int CW = TW + 2 * H, RH = TH + 2 * H;
    constexpr int LP = lcol(CW) + 1;
    constexpr int N = SEG + 2 * H;
    constexpr int HS = TW / SEG, VS = TH / SEG;
    constexpr int NI = (CW * RH + WGS - 1) / WGS;   // staged samples per thread
    static_assert(RH * HS <= WGS, "one round of row items");
    T *tile = reinterpret_cast<T *>(lds);

    const int x0 = bx * TW, y0 = by * TH;
    const int cw = a.cw, ch = a.ch;
    const int wl = (cw + 1) >> 1, hl = (ch + 1) >> 1;
    const bool vlift = PLAIN || ch > 1, hlift = PLAIN || cw > 1;
    const bool simple = PLAIN || (cw > H && ch > H);
    const bool ll_is_q = PLAIN ? false : a.ll_is_q != 0, q_float = PLAIN ? false : a.q_float != 0;
    const int32_t *qb = reinterpret_cast<const int32_t *>(a.q) + (long)plane * a.q_ps;
    const int32_t *lb = ll_is_q ? qb : reinterpret_cast<const int32_t *>(a.ll_in) + (long)plane * a.ll_ps;

    // ---- stage: every load first, then convert ----
    // Interior tiles read each band row as aligned 16-byte quads: band indices
    // [bq, bq + 4*NQ), bq = x0/2 - 4, band index b landing in interleaved
    // column 2*(b - (x0-H)/2) + (high band); other tiles go sample by sample.
    const int32_t *lrow_base = ll_is_q ? qb : lb;
    const long lpitch = ll_is_q ? a.qp : a.ll_pitch;
    const bool quads = x0 >= 8 && y0 >= H && (x0 >> 1) + TW / 2 + 4 <= wl && y0 + TH + H <= ch &&
                       ((a.qp | lpitch | wl) & 3) == 0 && (((uintptr_t)qb | (uintptr_t)lrow_base) & 15) == 0;
    if (quads) {
        constexpr int NQ = (TW / 2 + 8) / 4;            // quads per band and row
        constexpr int NI4 = (2 * NQ * RH + WGS - 1) / WGS;
        int4 raw[NI4];
        float stp[NI4];                                  // 9/7: step of an index, 0 for a float word
        const int bq = (x0 >> 1) - 4;
#pragma unroll
        for (int i = 0; i < NI4; ++i) {
            const int item = threadIdx.x + i * WGS;
            const int it = item < 2 * NQ * RH ? item : 0;
            const int j = it % (2 * NQ), rr = it / (2 * NQ);
            const int high = j >= NQ, m4 = j - high * NQ;
            const int y = y0 - H + rr;
            const bool ly = !(y & 1);
            const int byy = y >> 1;
            const bool from_ll = ly && !high;
            const int32_t *p = from_ll ? lrow_base + (long)byy * lpitch + bq + 4 * m4
                                       : qb + (long)(ly ? byy : hl + byy) * a.qp + (high ? wl : 0) + bq + 4 * m4;
            raw[i] = *reinterpret_cast<const int4 *>(p);
            if constexpr (F == RG_FILTER_97)
                stp[i] = ((from_ll && !ll_is_q) || q_float) ? 0.0f
                                                                 : (ly ? (high ? a.s_hl : a.s_ll) : (high ? a.s_hh : a.s_lh));
        }
#pragma unroll
        for (int i = 0; i < NI4; ++i) {
            const int item = threadIdx.x + i * WGS;
            if (item >= 2 * NQ * RH) break;
            const int j = item % (2 * NQ), rr = item / (2 * NQ);
            const int high = j >= NQ, m4 = j - high * NQ;
            const int c0 = 2 * (4 * m4 + H / 2 - 4) + high;
            const int32_t v4[4] = {raw[i].x, raw[i].y, raw[i].z, raw[i].w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = c0 + 2 * u;
                if (c < 0 || c >= CW) continue;
                T v;
                if constexpr (F == RG_FILTER_53) v = v4[u];
                else v = stp[i] != 0.0f ? dequant(v4[u], stp[i]) : __int_as_float(v4[u]);
                tile[rr * LP + lcol(c)] = v;
            }
        }
    } else {
        int32_t raw[NI];
        float stp[NI];      // 9/7: dequantisation step of an index, 0 for a float word
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int item = threadIdx.x + i * WGS;
            const int it = item < CW * RH ? item : 0;
            const int c = it % CW, rr = it / CW;
            const int y = simple ? mirror1(y0 - H + rr, ch) : rg_mirror(y0 - H + rr, ch);
            const int x = simple ? mirror1(x0 - H + c, cw) : rg_mirror(x0 - H + c, cw);
            const bool ly = !(y & 1), lx = !(x & 1);
            const int byy = y >> 1, bxx = x >> 1;
            const bool from_ll = ly && lx && !ll_is_q;
            const int32_t *p = from_ll ? lb + (long)byy * a.ll_pitch + bxx
                                       : qb + (long)(ly ? byy : hl + byy) * a.qp + (lx ? bxx : wl + bxx);
            raw[i] = *p;
            if constexpr (F == RG_FILTER_97)
                stp[i] = (from_ll || q_float) ? 0.0f : (ly ? (lx ? a.s_ll : a.s_hl) : (lx ? a.s_lh : a.s_hh));
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int item = threadIdx.x + i * WGS;
            if (item >= CW * RH) break;
            const int c = item % CW, rr = item / CW;
            T v;
            if constexpr (F == RG_FILTER_53) v = raw[i];
            else v = stp[i] != 0.0f ? dequant(raw[i], stp[i]) : __int_as_float(raw[i]);
            tile[rr * LP + lcol(c)] = v;
        }
    }
    __syncthreads();

    // ---- horizontal synthesis: (segment kk, row rr of RH), in place via registers ----
    const int kk = threadIdx.x % HS, rr = threadIdx.x / HS;
    const bool rowt = threadIdx.x < RH * HS;
    T o[SEG];
    if (rowt) {
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j) e[j] = tile[rr * LP + lcol(kk * SEG + j)];
        if constexpr (F == RG_FILTER_53) { if (hlift) inv53(e); }
        else inv97(e, hlift);
#pragma unroll
        for (int j = 0; j < SEG; ++j) o[j] = e[H + j];
    }
    __syncthreads();
    if (rowt) {
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[rr * LP + lcol(H + kk * SEG + j)] = o[j];
    }
    __syncthreads();

    // ---- vertical synthesis: (column c, segment k); row runs of TW samples ----
    T *obase = reinterpret_cast<T *>(reinterpret_cast<char *>(a.out) + (long)plane * a.out_ps * 4);
    for (int item = threadIdx.x; item < TW * VS; item += WGS) {
        const int c = item % TW, k = item / TW;
        const int x = x0 + c;
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j) e[j] = tile[(k * SEG + j) * LP + lcol(H + c)];
        if constexpr (F == RG_FILTER_53) { if (vlift) inv53(e); }
        else inv97(e, vlift);
        if (x >= cw) continue;
#pragma unroll
        for (int j = 0; j < SEG; ++j) {
            const int y = y0 + k * SEG + j;
            if (y >= ch) break;
            st_nt(obase + (long)y * a.out_pitch + x, e[H + j]);
        }
    }
}

// ==================== last inverse level fused with the inverse colour stage ====================
// One TW x TH tile of the image, every component: the level-1 bands of each
// component (dequantised) are staged interleaved in LDS exactly as inv_tile
// stages one plane, synthesised by rows then columns (items over all
// components at once), the columns' outputs written back over the staging,
// and the pixels colour transformed back (rg_px::px_inv, the inverse colour
// kernel's arithmetic) and stored as the interleaved uint8 image, four pixels
// (NC aligned dwords) per item.  The FP32 planes never go to HBM.
template <int F, int NC, int TW, int TH> constexpr int inv1c_lds_bytes()
{
    return NC * (TH + 2 * Filt<F>::H) * (lcol(TW + 2 * Filt<F>::H) + 1) * 4;
}

// Four pixels of row r (image row y, column x) from the synthesised tiles,
// colour transformed back, as NC aligned dwords of the uint8 image.
template <int F, int NC, int MODE, int TW, int H, int LP, int PS, class T>
__device__ __forceinline__ void inv1c_store_quad(const rg_inv1c_args &a, const T *tile, int r, int q4, int y, int x)
{
    uint32_t wd[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) wd[k] = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T v[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) v[k] = tile[k * PS + (H + r) * LP + lcol(H + 4 * q4 + i)];
        uint8_t ob[NC];
        rg_px::px_inv<NC, MODE, F == RG_FILTER_97>(v, ob);
#pragma unroll
        for (int k = 0; k < NC; ++k) wd[(NC * i + k) >> 2] |= (uint32_t)ob[k] << (((NC * i + k) & 3) * 8);
    }
    uint32_t *d32 = reinterpret_cast<uint32_t *>(a.dst + ((long)y * a.w + x) * NC);
#pragma unroll
    for (int k = 0; k < NC; ++k) __builtin_nontemporal_store(wd[k], d32 + k);
}

template <int F, int NC, int MODE, int TW, int TH, int WGS>
__device__ __forceinline__ void inv1_colour_tile(const rg_inv1c_args &a, int bx, int by, char *lds)
{
    using T = typename Filt<F>::T;
    constexpr int H = Filt<F>::H;
    constexpr int CW = TW + 2 * H, RH = TH + 2 * H;
    constexpr int LP = lcol(CW) + 1;
    constexpr int PS = RH * LP;                      // words per component
    constexpr int N = SEG + 2 * H;
    constexpr int HS = TW / SEG, VS = TH / SEG;
    static_assert(WGS % 64 == 0, "whole wavefronts");
    T *tile = reinterpret_cast<T *>(lds);
    const int x0 = bx * TW, y0 = by * TH;
    const int w = a.w, h = a.h;
    const int wl = (w + 1) >> 1, hl = (h + 1) >> 1;
    const bool ll_is_q = a.ll_is_q != 0;

    // ---- stage every component's bands, every load first ----
    const int32_t *qb0 = a.q;
    const int32_t *lb0 = ll_is_q ? a.q : reinterpret_cast<const int32_t *>(a.ll_in);
    const long lpitch = ll_is_q ? a.qp : a.ll_pitch, lps = ll_is_q ? a.q_ps : a.ll_ps;
    const bool quads = x0 >= 8 && y0 >= H && (x0 >> 1) + TW / 2 + 4 <= wl && y0 + TH + H <= h &&
                       ((a.qp | lpitch | wl | a.q_ps | lps) & 3) == 0 && (((uintptr_t)qb0 | (uintptr_t)lb0) & 15) == 0;
    if (quads) {
        constexpr int NQ = (TW / 2 + 8) / 4;            // quads per band and row
        constexpr int PI = 2 * NQ * RH;                 // items per component
        constexpr int NI4 = (NC * PI + WGS - 1) / WGS;
        int4 raw[NI4];
        float stp[NI4];                                  // 9/7: step of an index, 0 for an FP32 LL word
        const int bq = (x0 >> 1) - 4;
#pragma unroll
        for (int i = 0; i < NI4; ++i) {
            const int item = threadIdx.x + i * WGS;
            const int it = item < NC * PI ? item : 0;
            const int k = it / PI, r2 = it - k * PI;
            const int j = r2 % (2 * NQ), rr = r2 / (2 * NQ);
            const int high = j >= NQ, m4 = j - high * NQ;
            const int y = y0 - H + rr;
            const bool ly = !(y & 1);
            const int byy = y >> 1;
            const bool from_ll = ly && !high;
            const int32_t *p = from_ll ? lb0 + k * lps + (long)byy * lpitch + bq + 4 * m4
                                       : qb0 + k * a.q_ps + (long)(ly ? byy : hl + byy) * a.qp + (high ? wl : 0) + bq +
                                             4 * m4;
            raw[i] = *reinterpret_cast<const int4 *>(p);
            if constexpr (F == RG_FILTER_97)
                stp[i] = (from_ll && !ll_is_q) ? 0.0f : (ly ? (high ? a.s_hl : a.s_ll) : (high ? a.s_hh : a.s_lh));
        }
#pragma unroll
        for (int i = 0; i < NI4; ++i) {
            const int item = threadIdx.x + i * WGS;
            if (item >= NC * PI) break;
            const int k = item / PI, r2 = item - k * PI;
            const int j = r2 % (2 * NQ), rr = r2 / (2 * NQ);
            const int high = j >= NQ, m4 = j - high * NQ;
            const int c0 = 2 * (4 * m4 + H / 2 - 4) + high;
            const int32_t v4[4] = {raw[i].x, raw[i].y, raw[i].z, raw[i].w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int c = c0 + 2 * u;
                if (c < 0 || c >= CW) continue;
                T v;
                if constexpr (F == RG_FILTER_53) v = v4[u];
                else v = stp[i] != 0.0f ? dequant(v4[u], stp[i]) : __int_as_float(v4[u]);
                tile[k * PS + rr * LP + lcol(c)] = v;
            }
        }
    } else {
        constexpr int PI = CW * RH;
        constexpr int NI = (NC * PI + WGS - 1) / WGS;
        int32_t raw[NI];
        float stp[NI];
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int item = threadIdx.x + i * WGS;
            const int it = item < NC * PI ? item : 0;
            const int k = it / PI, r2 = it - k * PI;
            const int c = r2 % CW, rr = r2 / CW;
            const int y = mirror1(y0 - H + rr, h), x = mirror1(x0 - H + c, w);
            const bool ly = !(y & 1), lx = !(x & 1);
            const int byy = y >> 1, bxx = x >> 1;
            const bool from_ll = ly && lx && !ll_is_q;
            const int32_t *p = from_ll ? lb0 + k * lps + (long)byy * a.ll_pitch + bxx
                                       : qb0 + k * a.q_ps + (long)(ly ? byy : hl + byy) * a.qp + (lx ? bxx : wl + bxx);
            raw[i] = *p;
            if constexpr (F == RG_FILTER_97)
                stp[i] = from_ll ? 0.0f : (ly ? (lx ? a.s_ll : a.s_hl) : (lx ? a.s_lh : a.s_hh));
        }
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int item = threadIdx.x + i * WGS;
            if (item >= NC * PI) break;
            const int k = item / PI, r2 = item - k * PI;
            const int c = r2 % CW, rr = r2 / CW;
            T v;
            if constexpr (F == RG_FILTER_53) v = raw[i];
            else v = stp[i] != 0.0f ? dequant(raw[i], stp[i]) : __int_as_float(raw[i]);
            tile[k * PS + rr * LP + lcol(c)] = v;
        }
    }
    __syncthreads();

    // ---- horizontal synthesis: items (component, row of RH, segment) ----
    constexpr int NR = NC * RH * HS, RR = (NR + WGS - 1) / WGS;
    T o[RR][SEG];
#pragma unroll
    for (int u = 0; u < RR; ++u) {
        const int it = threadIdx.x + u * WGS;
        if (it < NR) {
            const int kk = it % HS, rk = it / HS;   // rk = component * RH + row
            T e[N];
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = tile[rk * LP + lcol(kk * SEG + j)];
            if constexpr (F == RG_FILTER_53) inv53(e); else inv97(e, true);
#pragma unroll
            for (int j = 0; j < SEG; ++j) o[u][j] = e[H + j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RR; ++u) {
        const int it = threadIdx.x + u * WGS;
        if (it < NR) {
            const int kk = it % HS, rk = it / HS;
#pragma unroll
            for (int j = 0; j < SEG; ++j) tile[rk * LP + lcol(H + kk * SEG + j)] = o[u][j];
        }
    }
    __syncthreads();

    // ---- vertical synthesis: items (component, column, segment); outputs
    //      written back over the staging (rows H .. H + TH) after every read ----
    constexpr int NV = NC * TW * VS, RV = (NV + WGS - 1) / WGS;
    T ov[RV][SEG];
#pragma unroll
    for (int u = 0; u < RV; ++u) {
        const int it = threadIdx.x + u * WGS;
        if (it < NV) {
            const int vc = it % TW, vs = (it / TW) % VS, vk = it / (TW * VS);
            T e[N];
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = tile[vk * PS + (vs * SEG + j) * LP + lcol(H + vc)];
            if constexpr (F == RG_FILTER_53) inv53(e); else inv97(e, true);
#pragma unroll
            for (int j = 0; j < SEG; ++j) ov[u][j] = e[H + j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RV; ++u) {
        const int it = threadIdx.x + u * WGS;
        if (it < NV) {
            const int vc = it % TW, vs = (it / TW) % VS, vk = it / (TW * VS);
#pragma unroll
            for (int j = 0; j < SEG; ++j) tile[vk * PS + (H + vs * SEG + j) * LP + lcol(H + vc)] = ov[u][j];
        }
    }
    __syncthreads();

    // ---- inverse colour: items (row, 4-pixel quad), NC aligned dwords out ----
    for (int it = threadIdx.x; it < TH * TW / 4; it += WGS) {
        const int r = it / (TW / 4), q4 = it % (TW / 4);
        const int y = y0 + r, x = x0 + 4 * q4;
        if (y >= h || x >= w) continue;
        inv1c_store_quad<F, NC, MODE, TW, H, LP, PS>(a, tile, r, q4, y, x);
    }
}


// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (each with its own L2), so in launch order neighbouring tiles land on
// different XCDs and every halo column is fetched from HBM again.  The
// bijection below gives XCD k (workgroups with linear id = k mod 8) the k-th
// contiguous run of tiles in raster order (plane, row, column): horizontal
// and vertical neighbours then share an L2.  Ids past the largest multiple
// of 8 keep their own tile.
constexpr int NXCD = 8;
__device__ __forceinline__ void xcd_tile(int &plane, int &bx, int &by)
{
    const int tx = gridDim.x, ty = gridDim.y;
    const int n = tx * ty * gridDim.z;
    const int lin = (blockIdx.z * ty + blockIdx.y) * tx + blockIdx.x;
    const int q = n / NXCD;
    const int t = lin < q * NXCD ? (lin % NXCD) * q + lin / NXCD : lin;
    plane = t / (tx * ty);
    by = (t / tx) % ty;
    bx = t % tx;
}

// PLAIN: 0 the general tile; 1-3 the plain 9/7 tile with FP32 input and LL
// (1), FP32 input and FP16 LL (2), FP16 input and LL (3)
template <int F, int PLAIN = 0>
__global__ void __launch_bounds__(FWG) fwd_level_kernel(rg_fwd_level_args a)
{
    __shared__ __attribute__((aligned(16))) char lds[FWD_LDS_BYTES];
    int plane, bx, by;
    xcd_tile(plane, bx, by);
    if constexpr (PLAIN == 1) fwd_tile_plain97<FTW, FTH, FWG, false, false>(a, plane, bx, by, lds);
    else if constexpr (PLAIN == 2) fwd_tile_plain97<FTW, FTH, FWG, false, true>(a, plane, bx, by, lds);
    else if constexpr (PLAIN == 3) fwd_tile_plain97<FTW, FTH, FWG, true, true>(a, plane, bx, by, lds);
    else fwd_tile<F, FTW, FTH, FWG>(a, plane, bx, by, lds);
}

// Fused colour + level 1, one workgroup per TW x TH tile of the image.
template <int F, int NC, int MODE, int TW, int TH, int WGS>
__global__ void __launch_bounds__(WGS) fwd1_colour_kernel(rg_fwd1c_args a)
{
    __shared__ __attribute__((aligned(16))) char lds[fwd1c_lds_bytes<F, NC, TW, TH>()];
    int plane, bx, by;
    xcd_tile(plane, bx, by);
    fwd1_colour_tile<F, NC, MODE, TW, TH, WGS>(a, bx, by, lds);
}

// Last inverse level with the inverse colour stage, one workgroup per
// TW x TH tile of the image.
template <int F, int NC, int MODE, int TW, int TH, int WGS>
__global__ void __launch_bounds__(WGS) inv1_colour_kernel(rg_inv1c_args a)
{
    __shared__ __attribute__((aligned(16))) char lds[inv1c_lds_bytes<F, NC, TW, TH>()];
    int plane, bx, by;
    xcd_tile(plane, bx, by);
    inv1_colour_tile<F, NC, MODE, TW, TH, WGS>(a, bx, by, lds);
}

template <int F, bool PLAIN = false>
__global__ void __launch_bounds__(IWG) inv_level_kernel(rg_inv_level_args a)
{
    __shared__ __attribute__((aligned(16))) char lds[INV_LDS_BYTES];
    int plane, bx, by;
    xcd_tile(plane, bx, by);
    inv_tile<F, ITW, ITH, IWG, PLAIN>(a, plane, bx, by, lds);
}

// Batched: one workgroup per job {image args index, plane, tile x, tile y};
// images of both filters share the launch (the filter is uniform per job).
__global__ void __launch_bounds__(FWG) fwd_level_batch_kernel(const rg_fwd_level_args *__restrict__ args,
                                                               const int4 *__restrict__ jobs)
{
    __shared__ __attribute__((aligned(16))) char lds[FWD_LDS_BYTES];
    const int4 j = jobs[blockIdx.x];
    const rg_fwd_level_args &a = args[j.x];
    if (a.filter == RG_FILTER_53) fwd_tile<RG_FILTER_53, FTW, FTH, FWG>(a, j.y, j.z, j.w, lds);
    else fwd_tile<RG_FILTER_97, FTW, FTH, FWG>(a, j.y, j.z, j.w, lds);
}

__global__ void __launch_bounds__(IWG) inv_level_batch_kernel(const rg_inv_level_args *__restrict__ args,
                                                               const int4 *__restrict__ jobs)
{
    __shared__ __attribute__((aligned(16))) char lds[INV_LDS_BYTES];
    const int4 j = jobs[blockIdx.x];
    const rg_inv_level_args &a = args[j.x];
    if (a.filter == RG_FILTER_53) inv_tile<RG_FILTER_53, ITW, ITH, IWG>(a, j.y, j.z, j.w, lds);
    else inv_tile<RG_FILTER_97, ITW, ITH, IWG>(a, j.y, j.z, j.w, lds);
}

}  // namespace

int rg_launch_fwd_level_batch(hipStream_t s, const rg_fwd_level_args *args, const int4 *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    fwd_level_batch_kernel<<<dim3((unsigned)njobs), dim3(FWG), 0, s>>>(args, jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_inv_level_batch(hipStream_t s, const rg_inv_level_args *args, const int4 *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    inv_level_batch_kernel<<<dim3((unsigned)njobs), dim3(IWG), 0, s>>>(args, jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_fwd53_level(hipStream_t s, const rg_fwd_level_args &a)
{
    dim3 grid((a.cw + FTW - 1) / FTW, (a.ch + FTH - 1) / FTH, a.nplanes);
    fwd_level_kernel<RG_FILTER_53><<<grid, dim3(FWG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_fwd97_level(hipStream_t s, const rg_fwd_level_args &a)
{
    dim3 grid((a.cw + FTW - 1) / FTW, (a.ch + FTH - 1) / FTH, a.nplanes);
    const bool plain = !a.q_float && a.cw > 4 && a.ch > 4;
    // the LL format only matters when the level writes an LL buffer
    const bool llh = a.ll_out && a.ll_half;
    if (plain && !a.in_half && !llh)
        fwd_level_kernel<RG_FILTER_97, 1><<<grid, dim3(FWG), 0, s>>>(a);
    else if (plain && !a.in_half)
        fwd_level_kernel<RG_FILTER_97, 2><<<grid, dim3(FWG), 0, s>>>(a);
    else if (plain && (llh || !a.ll_out))
        fwd_level_kernel<RG_FILTER_97, 3><<<grid, dim3(FWG), 0, s>>>(a);
    else
        fwd_level_kernel<RG_FILTER_97><<<grid, dim3(FWG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

bool rg_fwd1_colour_ok(int w, int h, int ncomp, int colour, int filter)
{
    const int halo = filter == RG_FILTER_97 ? 4 : 2;
    if (ncomp != 1 && ncomp != 3) return false;
    if (colour != RG_COLOUR_NONE && !(ncomp == 3 && ((colour == RG_COLOUR_ICT && filter == RG_FILTER_97) ||
                                                     (colour == RG_COLOUR_RCT && filter == RG_FILTER_53))))
        return false;
    return ((long)w * ncomp) % 4 == 0 && w > halo && h > halo && (long)w * h * ncomp < (1L << 31);
}

namespace {
// Fused tile shapes: 64 x 16 with 256 threads (22 KB of LDS for three
// components, 7 workgroups a CU), 64 x 32 with 256 threads (37 KB, 4 a CU;
// RG_DWT_FUSED_TH=32: fewer halo rows staged and lifted, 63 -> 70 us) or
// 128 x 16 with 512 threads (42 KB, 3 a CU; RG_DWT_FUSED_TW=128).  4096x4096
// RGB ICT + 9/7 with caches cold: 66 us against 85 us for the last.
template <int F, int NC, int MODE>
void launch_fwd1c(hipStream_t s, const rg_fwd1c_args &a)
{
    static const int tw = [] { const char *e = getenv("RG_DWT_FUSED_TW"); return e && atoi(e) == 128 ? 128 : 64; }();
    static const int th = [] { const char *e = getenv("RG_DWT_FUSED_TH"); return e && atoi(e) == 32 ? 32 : 16; }();
    if (tw == 128) {
        dim3 grid((a.w + 127) / 128, (a.h + SEG - 1) / SEG, 1);
        fwd1_colour_kernel<F, NC, MODE, 128, SEG, 512><<<grid, dim3(512), 0, s>>>(a);
    } else if (th == 16) {
        dim3 grid((a.w + 63) / 64, (a.h + 15) / 16, 1);
        fwd1_colour_kernel<F, NC, MODE, 64, 16, 256><<<grid, dim3(256), 0, s>>>(a);
    } else {
        dim3 grid((a.w + 63) / 64, (a.h + 31) / 32, 1);
        fwd1_colour_kernel<F, NC, MODE, 64, 32, 256><<<grid, dim3(256), 0, s>>>(a);
    }
}
}  // namespace

int rg_launch_fwd1_colour(hipStream_t s, const rg_fwd1c_args &a)
{
    RG_CHECK(rg_fwd1_colour_ok(a.w, a.h, a.ncomp, a.colour, a.filter),
             "fused colour + DWT: unsupported shape %dx%dx%d (colour %d, filter %d)", a.w, a.h, a.ncomp, a.colour,
             a.filter);
    RG_CHECK(((uintptr_t)a.src & 3) == 0, "fused colour + DWT: source must be 4-byte aligned");
    RG_CHECK((long)a.w * a.h * a.ncomp < (1L << 31), "fused colour + DWT: image of %dx%dx%d beyond 2 GiB", a.w, a.h,
             a.ncomp);
    if (a.filter == RG_FILTER_97) {
        if (a.ncomp == 1) launch_fwd1c<RG_FILTER_97, 1, 0>(s, a);
        else if (a.colour == RG_COLOUR_ICT) launch_fwd1c<RG_FILTER_97, 3, 2>(s, a);
        else launch_fwd1c<RG_FILTER_97, 3, 0>(s, a);
    } else {
        if (a.ncomp == 1) launch_fwd1c<RG_FILTER_53, 1, 0>(s, a);
        else if (a.colour == RG_COLOUR_RCT) launch_fwd1c<RG_FILTER_53, 3, 1>(s, a);
        else launch_fwd1c<RG_FILTER_53, 3, 0>(s, a);
    }
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

bool rg_inv1_colour_ok(int w, int h, int ncomp, int colour, int filter)
{
    return rg_fwd1_colour_ok(w, h, ncomp, colour, filter);
}

namespace {
template <int F, int NC, int MODE, int TW, int TH, int WGS>
void launch_inv1c_shape(hipStream_t s, const rg_inv1c_args &a)
{
    dim3 grid((a.w + TW - 1) / TW, (a.h + TH - 1) / TH, 1);
    inv1_colour_kernel<F, NC, MODE, TW, TH, WGS><<<grid, dim3(WGS), 0, s>>>(a);
}

// Fused inverse tile: 64 x 32 with 512 threads (37 KB of LDS for RGB, three
// workgroups a CU).  4096x4096 RGB, caches cold: 64 x 16 with 256 threads,
// 64 x 32 with 256 and 64 x 64 with 512 measured 4-16% slower.
template <int F, int NC, int MODE>
void launch_inv1c(hipStream_t s, const rg_inv1c_args &a)
{
    launch_inv1c_shape<F, NC, MODE, 64, 32, 512>(s, a);
}
}  // namespace

int rg_launch_inv1_colour(hipStream_t s, const rg_inv1c_args &a)
{
    RG_CHECK(rg_inv1_colour_ok(a.w, a.h, a.ncomp, a.colour, a.filter),
             "fused inverse DWT + colour: unsupported shape %dx%dx%d (colour %d, filter %d)", a.w, a.h, a.ncomp,
             a.colour, a.filter);
    RG_CHECK(((uintptr_t)a.dst & 3) == 0, "fused inverse DWT + colour: destination must be 4-byte aligned");
    if (a.filter == RG_FILTER_97) {
        if (a.ncomp == 1) launch_inv1c<RG_FILTER_97, 1, 0>(s, a);
        else if (a.colour == RG_COLOUR_ICT) launch_inv1c<RG_FILTER_97, 3, 2>(s, a);
        else launch_inv1c<RG_FILTER_97, 3, 0>(s, a);
    } else {
        if (a.ncomp == 1) launch_inv1c<RG_FILTER_53, 1, 0>(s, a);
        else if (a.colour == RG_COLOUR_RCT) launch_inv1c<RG_FILTER_53, 3, 1>(s, a);
        else launch_inv1c<RG_FILTER_53, 3, 0>(s, a);
    }
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_inv53_level(hipStream_t s, const rg_inv_level_args &a)
{
    dim3 grid((a.cw + ITW - 1) / ITW, (a.ch + ITH - 1) / ITH, a.nplanes);
    inv_level_kernel<RG_FILTER_53><<<grid, dim3(IWG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_inv97_level(hipStream_t s, const rg_inv_level_args &a)
{
    dim3 grid((a.cw + ITW - 1) / ITW, (a.ch + ITH - 1) / ITH, a.nplanes);
    if (!a.ll_is_q && !a.q_float && a.cw > 4 && a.ch > 4)
        inv_level_kernel<RG_FILTER_97, true><<<grid, dim3(IWG), 0, s>>>(a);
    else
        inv_level_kernel<RG_FILTER_97><<<grid, dim3(IWG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

// ---- zero-level transform: plain copy (5/3) or (de)quantisation (9/7) ----
namespace {
__global__ void copy_quant_kernel(int filter, const void *__restrict__ in, long n, float inv, int q_float,
                                  void *__restrict__ q)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (filter == RG_FILTER_53) reinterpret_cast<int *>(q)[i] = reinterpret_cast<const int *>(in)[i];
    else if (q_float) reinterpret_cast<float *>(q)[i] = reinterpret_cast<const float *>(in)[i];
    else reinterpret_cast<int *>(q)[i] = quant(reinterpret_cast<const float *>(in)[i], inv);
}

__global__ void copy_dequant_kernel(int filter, const void *__restrict__ q, long n, float step, int q_float,
                                    void *__restrict__ out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (filter == RG_FILTER_53) reinterpret_cast<int *>(out)[i] = reinterpret_cast<const int *>(q)[i];
    else if (q_float) reinterpret_cast<float *>(out)[i] = reinterpret_cast<const float *>(q)[i];
    else reinterpret_cast<float *>(out)[i] = dequant(reinterpret_cast<const int *>(q)[i], step);
}
}  // namespace

int rg_launch_copy_quant(hipStream_t s, int filter, const void *in, long n, int nplanes, float inv, int q_float,
                         void *q)
{
    long tot = n * nplanes;
    if (tot == 0) return RG_SUCCESS;
    copy_quant_kernel<<<dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s>>>(filter, in, tot, inv, q_float, q);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_copy_dequant(hipStream_t s, int filter, const void *q, long n, int nplanes, float step, int q_float,
                           void *out)
{
    long tot = n * nplanes;
    if (tot == 0) return RG_SUCCESS;
    copy_dequant_kernel<<<dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s>>>(filter, q, tot, step, q_float, out);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}
