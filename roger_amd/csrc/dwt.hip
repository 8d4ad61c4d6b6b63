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
// Kernel design (gfx950): one 320-thread (five-wavefront) workgroup per 64x64
// tile of the level's input region.  Forward:
//   1. vertical pass straight from HBM: work item (column, 16-row segment)
//      loads its segment plus a 4-row (9/7) / 2-row (5/3) halo with coalesced
//      row-contiguous loads, runs all lifting steps in registers, and stores
//      the 16 results into an LDS tile (64 rows x 72 columns);
//      -- the 72 (9/7) or 68 (5/3) halo'd columns x 4 segments are one round
//      of the 320 work items, so every load of the tile is issued at once;
//   2. one barrier;
//   3. horizontal pass from LDS: work item (row, 16-column segment) lifts in
//      registers and writes its 8 low + 8 high outputs straight to their
//      Mallat positions (quantised to int32 for the high bands, LL to the
//      next level's buffer), consecutive lanes covering consecutive columns.
// Each input sample is read from HBM about (72/64) x (88/64) times (halo,
// mostly served by L2) and each output written once.  Inverse kernels mirror
// this with the horizontal pass first (T.800 2D_SR).
#include <hip/hip_fp16.h>
#include "rg_internal.h"

namespace {

constexpr int TILE = 64;
constexpr int SEG = 16;
constexpr int NSEG = TILE / SEG;
constexpr int WG = 320;                    // >= (TILE + 2 * 4) * NSEG: one round of column segments
constexpr int ROWT = TILE * NSEG;          // work items of the row passes (the first 256 threads)
static_assert(WG >= (TILE + 8) * NSEG && WG >= ROWT, "one round of halo'd column segments");

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

__device__ __forceinline__ float dequant(int q, float s)
{
    float m = ((float)(q < 0 ? -q : q) + 0.5f) * s;
    return q == 0 ? 0.0f : (q < 0 ? -m : m);
}

// ============================ forward ============================
constexpr int FWD_LDS_BYTES = TILE * (TILE + 2 * 4 + 1) * 4;   // largest (9/7) tile
constexpr int INV_LDS_BYTES = (TILE + 2 * 4) * (TILE + 1) * 4;

// Symmetric-extension index: one reflection then a clamp.  Exact for every
// position an output inside the signal depends on (overshoot <= halo < n);
// positions further out only feed outputs outside the signal, and the clamp
// keeps their loads in bounds.  Levels shorter than halo+1 use rg_mirror.
__device__ __forceinline__ int mirror1(int i, int n)
{
    i = i < 0 ? -i : (i >= n ? 2 * n - 2 - i : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// One 64x64 tile (bx, by) of plane `plane`; `lds` holds FWD_LDS_BYTES.
template <int F>
__device__ __forceinline__ void fwd_tile(const rg_fwd_level_args &a, int plane, int bx, int by, char *lds)
{
    using T = typename Filt<F>::T;
    constexpr int H = Filt<F>::H;
    constexpr int TW = TILE + 2 * H;          // LDS columns (with halo)
    constexpr int LP = TW + 1;                // padded pitch
    constexpr int N = SEG + 2 * H;            // register window
    T *tile = reinterpret_cast<T *>(lds);

    const int x0 = bx * TILE, y0 = by * TILE;
    const int cw = a.cw, ch = a.ch;
    const void *in = reinterpret_cast<const char *>(a.in) +
                     (long)plane * a.in_ps * (F == RG_FILTER_53 ? 4 : (a.in_half ? 2 : 4));
    const bool vlift = ch > 1, hlift = cw > 1;
    const bool simple = cw > H && ch > H;      // one reflection suffices
    // the tile and its halo lie inside the level: no extension, no bounds tests
    const bool interior = x0 >= H && y0 >= H && x0 + TILE + H <= cw && y0 + TILE + H <= ch;

    // ---- vertical pass straight from HBM: (column c of TW, segment k) ----
    for (int item = threadIdx.x; item < TW * NSEG; item += blockDim.x) {
        const int c = item % TW, k = item / TW;
        const int x = simple ? mirror1(x0 - H + c, cw) : rg_mirror(x0 - H + c, cw);
        const int ybase = y0 + k * SEG - H;
        T e[N];
        if (interior) {
            const long first = (long)ybase * a.in_pitch + x;
#pragma unroll
            for (int j = 0; j < N; ++j) e[j] = load_in<F>(in, first + (long)j * a.in_pitch, a.in_half);
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
        for (int j = 0; j < SEG; ++j) tile[(k * SEG + j) * LP + c] = e[H + j];
    }
    __syncthreads();

    // ---- horizontal pass from LDS: (segment k, row r), k fastest ----
    T o[SEG];
    const int kk = threadIdx.x % NSEG, rr = threadIdx.x / NSEG;   // rr < TILE for the first ROWT threads
    const bool rowt = threadIdx.x < ROWT;
    if (rowt) {
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j) e[j] = tile[rr * LP + kk * SEG + j];
        if (hlift) {
            if constexpr (F == RG_FILTER_53) fwd53(e); else fwd97(e, true);
        }
#pragma unroll
        for (int j = 0; j < SEG; ++j) o[j] = e[H + j];
    }
    __syncthreads();
    // deinterleave the row into LDS: low outputs to [0, 32), high to [32, 64)
    if (rowt) {
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[rr * LP + (j & 1) * (TILE / 2) + kk * (SEG / 2) + (j >> 1)] = o[j];
    }
    __syncthreads();

    // ---- write out: each wavefront stores one row: 32 low + 32 high values ----
    const int wl = (cw + 1) >> 1, hl = (ch + 1) >> 1;
    char *qbase = reinterpret_cast<char *>(a.q) + (long)plane * a.q_ps * 4;
    if (interior) {
        // a thread keeps one column c and rows r0, r0+4, ...: its band, its
        // step and the parity of its rows never change (y0 is even)
        if (!rowt) return;
        const int c = threadIdx.x % TILE, r0 = threadIdx.x / TILE;   // r0 in 0..3
        const bool high = c >= TILE / 2, lowrow = !(r0 & 1);
        const int xo = (x0 >> 1) + (c & (TILE / 2 - 1));
        const int yh = (y0 + r0) >> 1;                                  // row inside the half band
        const T *src = tile + r0 * LP + c;
        if (lowrow && !high && a.ll_out) {
            const long off = (long)plane * a.ll_ps + (long)yh * a.ll_pitch + xo, step = 2L * a.ll_pitch;
#pragma unroll
            for (int i = 0; i < TILE / 4; ++i) {
                const T v = src[4 * i * LP];
                if constexpr (F == RG_FILTER_53) reinterpret_cast<int *>(a.ll_out)[off + i * step] = v;
                else if (a.ll_half) reinterpret_cast<__half *>(a.ll_out)[off + i * step] = __float2half_rn(v);
                else reinterpret_cast<float *>(a.ll_out)[off + i * step] = v;
            }
            return;
        }
        const long off = (long)(lowrow ? yh : hl + yh) * a.qp + (high ? wl + xo : xo), step = 2L * a.qp;
        if constexpr (F == RG_FILTER_53) {
#pragma unroll
            for (int i = 0; i < TILE / 4; ++i) reinterpret_cast<int *>(qbase)[off + i * step] = src[4 * i * LP];
        } else if (a.q_float) {
#pragma unroll
            for (int i = 0; i < TILE / 4; ++i) reinterpret_cast<float *>(qbase)[off + i * step] = src[4 * i * LP];
        } else {
            // sign(v) * floor(|v| * inv) == truncation of v * inv (IEEE products are sign-symmetric)
            const float inv = lowrow ? (high ? a.inv_hl : a.inv_ll) : (high ? a.inv_hh : a.inv_lh);
#pragma unroll
            for (int i = 0; i < TILE / 4; ++i)
                reinterpret_cast<int *>(qbase)[off + i * step] = (int)(src[4 * i * LP] * inv);
        }
        return;
    }
    for (int idx = threadIdx.x; idx < TILE * TILE; idx += blockDim.x) {
        const int r = idx / TILE, c = idx % TILE;
        const int y = y0 + r;
        const bool high = c >= TILE / 2;
        const int xo = (x0 >> 1) + (c & (TILE / 2 - 1));       // column inside its half band
        const int x = 2 * xo + (high ? 1 : 0);                 // source column
        if (y >= ch || x >= cw) continue;
        const T v = tile[r * LP + c];
        const bool lowrow = !(y & 1);
        if (lowrow && !high && a.ll_out) {
            const long off = (long)plane * a.ll_ps + (long)(y >> 1) * a.ll_pitch + xo;
            if constexpr (F == RG_FILTER_53) reinterpret_cast<int *>(a.ll_out)[off] = v;
            else if (a.ll_half) reinterpret_cast<__half *>(a.ll_out)[off] = __float2half_rn(v);
            else reinterpret_cast<float *>(a.ll_out)[off] = v;
            continue;
        }
        const long off = (long)(lowrow ? (y >> 1) : hl + (y >> 1)) * a.qp + (high ? wl + xo : xo);
        if constexpr (F == RG_FILTER_53) {
            reinterpret_cast<int *>(qbase)[off] = v;
        } else if (a.q_float) {
            reinterpret_cast<float *>(qbase)[off] = v;
        } else {
            const float inv = lowrow ? (high ? a.inv_hl : a.inv_ll) : (high ? a.inv_hh : a.inv_lh);
            reinterpret_cast<int *>(qbase)[off] = quant(v, inv);
        }
    }
}

// ============================ inverse ============================
template <int F>
__device__ __forceinline__ typename Filt<F>::T load_band(const rg_inv_level_args &a, int plane, int y, int x,
                                                          int wl, int hl)
{
    // sample (y, x) of the interleaved region, taken from its band
    const bool ly = !(y & 1), lx = !(x & 1);
    const int by = y >> 1, bx = x >> 1;
    if (ly && lx) {
        if (!a.ll_is_q) {
            long o = (long)plane * a.ll_ps + (long)by * a.ll_pitch + bx;
            if constexpr (F == RG_FILTER_53) return reinterpret_cast<const int *>(a.ll_in)[o];
            else return reinterpret_cast<const float *>(a.ll_in)[o];
        }
    }
    const int row = ly ? by : hl + by, col = lx ? bx : wl + bx;
    long o = (long)plane * a.q_ps + (long)row * a.qp + col;
    if constexpr (F == RG_FILTER_53) {
        return reinterpret_cast<const int *>(a.q)[o];
    } else {
        if (a.q_float) return reinterpret_cast<const float *>(a.q)[o];
        const float s = ly ? (lx ? a.s_ll : a.s_hl) : (lx ? a.s_lh : a.s_hh);
        return dequant(reinterpret_cast<const int *>(a.q)[o], s);
    }
}

template <int F>
__device__ __forceinline__ void inv_tile(const rg_inv_level_args &a, int plane, int bx, int by, char *lds)
{
    using T = typename Filt<F>::T;
    constexpr int H = Filt<F>::H;
    constexpr int TH = TILE + 2 * H;          // LDS rows (with halo)
    constexpr int LP = TILE + 1;
    constexpr int N = SEG + 2 * H;
    T *tile = reinterpret_cast<T *>(lds);

    const int x0 = bx * TILE, y0 = by * TILE;
    const int cw = a.cw, ch = a.ch;
    const int wl = (cw + 1) >> 1, hl = (ch + 1) >> 1;
    const bool vlift = ch > 1, hlift = cw > 1;

    // ---- horizontal synthesis: (segment k, row rr of TH), k fastest ----
    for (int item = threadIdx.x; item < TH * NSEG; item += blockDim.x) {
        const int k = item % NSEG, rr = item / NSEG;
        const bool simple = cw > H && ch > H;
        const int y = simple ? mirror1(y0 - H + rr, ch) : rg_mirror(y0 - H + rr, ch);
        const int xb = x0 + k * SEG - H;
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            e[j] = load_band<F>(a, plane, y, simple ? mirror1(xb + j, cw) : rg_mirror(xb + j, cw), wl, hl);
        if constexpr (F == RG_FILTER_53) { if (hlift) inv53(e); }
        else inv97(e, hlift);
#pragma unroll
        for (int j = 0; j < SEG; ++j) tile[rr * LP + k * SEG + j] = e[H + j];
    }
    __syncthreads();

    // ---- vertical synthesis: (column c, segment k) ----
    char *obase = reinterpret_cast<char *>(a.out) + (long)plane * a.out_ps * 4;
    for (int item = threadIdx.x; item < TILE * NSEG; item += blockDim.x) {
        const int c = item % TILE, k = item / TILE;
        const int x = x0 + c;
        if (x >= cw) continue;
        T e[N];
#pragma unroll
        for (int j = 0; j < N; ++j) e[j] = tile[(k * SEG + j) * LP + c];
        if constexpr (F == RG_FILTER_53) { if (vlift) inv53(e); }
        else inv97(e, vlift);
#pragma unroll
        for (int j = 0; j < SEG; ++j) {
            const int y = y0 + k * SEG + j;
            if (y >= ch) break;
            reinterpret_cast<T *>(obase)[(long)y * a.out_pitch + x] = e[H + j];
        }
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

template <int F>
__global__ void __launch_bounds__(WG) fwd_level_kernel(rg_fwd_level_args a)
{
    __shared__ __attribute__((aligned(16))) char lds[FWD_LDS_BYTES];
    int plane, bx, by;
    xcd_tile(plane, bx, by);
    fwd_tile<F>(a, plane, bx, by, lds);
}

template <int F>
__global__ void __launch_bounds__(WG) inv_level_kernel(rg_inv_level_args a)
{
    __shared__ __attribute__((aligned(16))) char lds[INV_LDS_BYTES];
    int plane, bx, by;
    xcd_tile(plane, bx, by);
    inv_tile<F>(a, plane, bx, by, lds);
}

// Batched: one workgroup per job {image args index, plane, tile x, tile y};
// images of both filters share the launch (the filter is uniform per job).
__global__ void __launch_bounds__(WG) fwd_level_batch_kernel(const rg_fwd_level_args *__restrict__ args,
                                                              const int4 *__restrict__ jobs)
{
    __shared__ __attribute__((aligned(16))) char lds[FWD_LDS_BYTES];
    const int4 j = jobs[blockIdx.x];
    const rg_fwd_level_args &a = args[j.x];
    if (a.filter == RG_FILTER_53) fwd_tile<RG_FILTER_53>(a, j.y, j.z, j.w, lds);
    else fwd_tile<RG_FILTER_97>(a, j.y, j.z, j.w, lds);
}

__global__ void __launch_bounds__(WG) inv_level_batch_kernel(const rg_inv_level_args *__restrict__ args,
                                                              const int4 *__restrict__ jobs)
{
    __shared__ __attribute__((aligned(16))) char lds[INV_LDS_BYTES];
    const int4 j = jobs[blockIdx.x];
    const rg_inv_level_args &a = args[j.x];
    if (a.filter == RG_FILTER_53) inv_tile<RG_FILTER_53>(a, j.y, j.z, j.w, lds);
    else inv_tile<RG_FILTER_97>(a, j.y, j.z, j.w, lds);
}

}  // namespace

int rg_launch_fwd_level_batch(hipStream_t s, const rg_fwd_level_args *args, const int4 *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    fwd_level_batch_kernel<<<dim3((unsigned)njobs), dim3(WG), 0, s>>>(args, jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_inv_level_batch(hipStream_t s, const rg_inv_level_args *args, const int4 *jobs, int njobs)
{
    if (njobs <= 0) return RG_SUCCESS;
    inv_level_batch_kernel<<<dim3((unsigned)njobs), dim3(WG), 0, s>>>(args, jobs);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_fwd53_level(hipStream_t s, const rg_fwd_level_args &a)
{
    dim3 grid((a.cw + TILE - 1) / TILE, (a.ch + TILE - 1) / TILE, a.nplanes);
    fwd_level_kernel<RG_FILTER_53><<<grid, dim3(WG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_fwd97_level(hipStream_t s, const rg_fwd_level_args &a)
{
    dim3 grid((a.cw + TILE - 1) / TILE, (a.ch + TILE - 1) / TILE, a.nplanes);
    fwd_level_kernel<RG_FILTER_97><<<grid, dim3(WG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_inv53_level(hipStream_t s, const rg_inv_level_args &a)
{
    dim3 grid((a.cw + TILE - 1) / TILE, (a.ch + TILE - 1) / TILE, a.nplanes);
    inv_level_kernel<RG_FILTER_53><<<grid, dim3(WG), 0, s>>>(a);
    RG_LAUNCH_CHECK();
    return RG_SUCCESS;
}

int rg_launch_inv97_level(hipStream_t s, const rg_inv_level_args &a)
{
    dim3 grid((a.cw + TILE - 1) / TILE, (a.ch + TILE - 1) / TILE, a.nplanes);
    inv_level_kernel<RG_FILTER_97><<<grid, dim3(WG), 0, s>>>(a);
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
