// Rate control by post-compression rate-distortion optimisation (EBCOT's
// PCRD-opt) for the parallel code-block styles, whose every coding pass is a
// terminated segment of known length, so a block can be cut after any pass.
//
// The reference has no rate control (its bit-plane coder writes no bytes,
// src/CL/oclbpc.cl:279-746); this is the published algorithm (Taubman, "High
// performance scalable image compression with EBCOT", IEEE TIP 2000, IV):
//
//   rd_dist_kernel   one wavefront per block, lane = column: the distortion
//                    reduction of every coding pass, in squared doubled index
//                    units (a decoder reconstructs the known bits plus half of
//                    the remaining interval; an insignificant sample is 0).
//                    Per bit-plane p the reductions follow from sums over the
//                    magnitudes' low bits g = m mod 2^p: a sample becoming
//                    significant adds 3*2^p*(2^p + 4g), a refinement of a 1
//                    bit 2^p*(4g - 2^p), of a 0 bit 2^p*(3*2^p - 4g) (at p = 0:
//                    4, 0 and 4).  Which of the plane's significance events
//                    fall in its significance-propagation pass (the rest are in
//                    its cleanup pass) comes from the context modeller's masks.
//   rd_hull_kernel   one thread per block: the upper convex hull of the
//                    points (R_n, w_b * D_n) over the pass ends n, each point
//                    with its slope to the previous one as a sortable key (the
//                    bit pattern of the non-negative double).
//   rd_hist_kernel / rd_pick_kernel
//                    eight rounds of radix selection on the 64-bit keys, eight
//                    bits a round: the largest key K with R(K) > target, where
//                    R(K) sums over all blocks the rate of the hull points with
//                    key >= K; the threshold is K + 1 (0 if everything fits).
//   rd_apply_kernel  one thread per block: keeps the hull points with key >=
//                    the threshold and cuts the block's result there.
//
// Workspace (all device memory, sized by rg_rd_ws_bytes): per-pass
// distortions [ncb][MAXP] int64; hull keys [MAXP][ncb] uint64, rates and pass
// counts [MAXP][ncb] int32 (point-major, so a thread per block reads
// neighbouring words), hull distortions [MAXP][ncb] int64, hull sizes [ncb];
// the 256-bin histogram and the selection state.
#include "rg_internal.h"
#include "wave_hip.h"
#include "t1_core.h"

using namespace rg_t1;

namespace {

constexpr int RD_MAX_NBP = 24;   // magnitudes below 2^24: every sum below fits int64

struct RdState {
    unsigned long long prefix, above, target, K, kept;
    unsigned int done, err;
};

struct RdWs {
    int64_t *dist;
    unsigned long long *hk;
    int32_t *hr, *hn, *cnt;
    int64_t *hd;
    unsigned long long *hist;
    RdState *st;
};

RdWs rd_ws(void *ws, int ncb)
{
    char *p = (char *)ws;
    RdWs w;
    const size_t n = (size_t)ncb * MAXP;
    w.dist = (int64_t *)p; p += n * 8;
    w.hk = (unsigned long long *)p; p += n * 8;
    w.hd = (int64_t *)p; p += n * 8;
    w.hr = (int32_t *)p; p += n * 4;
    w.hn = (int32_t *)p; p += n * 4;
    w.cnt = (int32_t *)p; p += ((size_t)ncb * 4 + 255) / 256 * 256;
    w.hist = (unsigned long long *)p; p += 256 * 8;
    w.st = (RdState *)p;
    return w;
}

RG_DEV int64_t wave_sum_i64(int64_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t lo = wv_shfl_xor((uint32_t)v, o), hi = wv_shfl_xor((uint32_t)((uint64_t)v >> 32), o);
        v += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    return v;
}

// Selection state for a new target (and, first, the error flag).
__global__ void rd_init_kernel(RdState *st, unsigned long long *hist, unsigned long long target, int first)
{
    const int t = threadIdx.x;
    hist[t] = 0;
    if (t == 0) {
        st->prefix = 0; st->above = 0; st->target = target; st->K = 0; st->kept = 0;
        st->done = 0;
        if (first) st->err = 0;
    }
}

// The sums over the low bits go through the block's bit-planes: with B_j the
// column's rows whose magnitude has bit j set (bit r = row r), the sum of
// g = m mod 2^p over a row set X is sum_{j<p} 2^j * popcount(X & B_j).  The
// planes are built once per block into workgroup memory (lane-private words,
// so no barrier), then each plane p reads its lower planes back.
__global__ void __launch_bounds__(64) rd_dist_kernel(const int32_t *__restrict__ q, int pitch, size_t plane_stride,
                                                     const rg_cb_desc *__restrict__ desc, const int4 *__restrict__ cm,
                                                     const uint64_t *__restrict__ sppm, int64_t *__restrict__ dist,
                                                     RdState *st)
{
    __shared__ uint64_t Bs[RD_MAX_NBP * 64];
    const int lane = wv_lane();
    const size_t b = blockIdx.x;
    const rg_cb_desc d = desc[b];
    const int nbp = wv_uniform(cm[b].z);
    int64_t *D = dist + b * MAXP;
    if (nbp == 0) return;
    if (nbp > RD_MAX_NBP) {
        if (lane == 0) atomicOr(&st->err, 1u);
        return;
    }
    {
        const int32_t *src = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
        const bool colok = lane < d.w;
        uint32_t mag[64];
#pragma unroll
        for (int r = 0; r < 64; ++r) {
            const int v = (colok && r < d.h) ? src[(size_t)r * pitch + lane] : 0;
            mag[r] = (uint32_t)(v < 0 ? -v : v);
        }
        for (int j = 0; j < nbp; ++j) {
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int r = 0; r < 32; ++r) lo |= ((mag[r] >> j) & 1u) << r;
#pragma unroll
            for (int r = 0; r < 32; ++r) hi |= ((mag[r + 32] >> j) & 1u) << r;
            Bs[j * 64 + lane] = ((uint64_t)hi << 32) | lo;
        }
    }
    const uint64_t *M = sppm + b * RD_PLANES * 64;
    uint64_t SIG = 0;   // rows significant above the current plane
    for (int p = nbp - 1; p >= 0; --p) {
        const uint64_t Bp = Bs[p * 64 + lane];
        const uint64_t nm = p == nbp - 1 ? 0 : M[p * 64 + lane];
        const uint64_t Q1 = Bp & ~SIG;                 // significant at p
        const uint64_t XS = Q1 & nm, XC = Q1 & ~nm;     // ... in the propagation / the cleanup pass
        const uint64_t X1 = Bp & SIG, X0 = SIG & ~Bp;   // refined with a 1 / a 0
        // g sums below 64 * 2^23: 32 bits suffice
        uint32_t gS = 0, gC = 0, g1 = 0, g0 = 0;
        for (int j = 0; j < p; ++j) {
            const uint64_t Bj = Bs[j * 64 + lane];
            gS += (uint32_t)__popcll(XS & Bj) << j;
            gC += (uint32_t)__popcll(XC & Bj) << j;
            g1 += (uint32_t)__popcll(X1 & Bj) << j;
            g0 += (uint32_t)__popcll(X0 & Bj) << j;
        }
        const int64_t cS = __popcll(XS), cC = __popcll(XC), c1 = __popcll(X1), c0 = __popcll(X0);
        int64_t dS, dC, dR;
        if (p == 0) {
            dS = 4 * cS;
            dC = 4 * cC;
            dR = 4 * c0;
        } else {
            const int64_t P = (int64_t)1 << p;
            dS = 3 * P * (cS * P + 4 * (int64_t)gS);
            dC = 3 * P * (cC * P + 4 * (int64_t)gC);
            dR = P * (4 * (int64_t)g1 - c1 * P + 3 * c0 * P - 4 * (int64_t)g0);
        }
        dS = wave_sum_i64(dS);
        dC = wave_sum_i64(dC);
        dR = wave_sum_i64(dR);
        if (lane == 0) {
            if (p == nbp - 1) {
                D[0] = dC;   // the first pass: cleanup of the top plane
            } else {
                const int k = 3 * (nbp - 1 - p) - 2;
                D[k] = dS;
                D[k + 1] = dR;
                D[k + 2] = dC;
            }
        }
        SIG |= Bp;
    }
}

RG_DEV unsigned long long key_of(double s) { return (unsigned long long)__double_as_longlong(s); }

__global__ void rd_hull_kernel(int ncb, const int4 *__restrict__ cm, const int32_t *__restrict__ seglens,
                               const double *__restrict__ weight, RdWs w)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ncb) return;
    const int npass = cm[b].y;
    const int32_t *pl = seglens + (size_t)b * MAXP;
    const int64_t *dd = w.dist + (size_t)b * MAXP;
    const double wb = weight[b];
    const size_t S = (size_t)ncb;
    int m = 0;
    long r = 0;
    int64_t d = 0;
    // the hull end (rate, distortion, key) stays in registers; memory is
    // read back only when a point is dropped
    long re = 0;
    int64_t de = 0;
    unsigned long long ke = 0;
    for (int n = 1; n <= npass; ++n) {
        r += pl[n - 1];
        d += dd[n - 1];
        for (;;) {
            if (d <= de) break;                                   // no gain over the hull end
            const double s = r > re ? (wb * (double)(d - de)) / (double)(r - re) : __builtin_huge_val();
            const unsigned long long k = key_of(s);
            if (m && k >= ke) {                                   // the hull end lies under the chord
                --m;
                re = m ? (long)w.hr[(m - 1) * S + b] : 0;
                de = m ? w.hd[(m - 1) * S + b] : 0;
                ke = m ? w.hk[(m - 1) * S + b] : 0;
                continue;
            }
            w.hn[m * S + b] = n;
            w.hr[m * S + b] = (int32_t)r;
            w.hd[m * S + b] = d;
            w.hk[m * S + b] = k;
            re = r; de = d; ke = k;
            ++m;
            break;
        }
    }
    w.cnt[b] = m;
}

// Round `round` (0..7) of the selection: histogram, over the points whose
// keys carry the prefix chosen so far, of the next eight key bits, weighted by
// each point's rate increment over its predecessor in its block.
__global__ void __launch_bounds__(256) rd_hist_kernel(int ncb, int round, RdWs w)
{
    __shared__ unsigned long long h[256];
    const int t = threadIdx.x;
    h[t] = 0;
    __syncthreads();
    const RdState *st = w.st;
    const int b = blockIdx.x * blockDim.x + t;
    if (!st->done && b < ncb) {
        const int shift = 56 - 8 * round;
        const unsigned long long hi = round == 0 ? 0ull : ~0ull << (shift + 8);
        const unsigned long long prefix = st->prefix;
        const size_t S = (size_t)ncb;
        const int c = w.cnt[b];
        int32_t prev = 0;
        for (int i = 0; i < c; ++i) {
            const unsigned long long k = w.hk[i * S + b];
            const int32_t rr = w.hr[i * S + b];
            if ((k & hi) == prefix) atomicAdd(&h[(k >> shift) & 255u], (unsigned long long)(rr - prev));
            prev = rr;
        }
    }
    __syncthreads();
    if (h[t]) atomicAdd(&w.hist[t], h[t]);
}

// One workgroup of 256: with S[d] the histogram's suffix sums, the digit is
// the largest d with above + S[d] > target (it exists: the previous round's
// choice guarantees above + S[0] > target).  Round 0 first checks whether
// everything fits (threshold 0).  Clears the histogram for the next round;
// the last round sets the threshold.
__global__ void __launch_bounds__(256) rd_pick_kernel(int round, RdWs w)
{
    __shared__ unsigned long long sc[256];
    __shared__ int ntrue;
    RdState *st = w.st;
    const int t = threadIdx.x;
    const unsigned done = st->done;
    sc[t] = w.hist[255 - t];   // reversed: an inclusive prefix scan gives suffix sums
    if (t == 0) ntrue = 0;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const unsigned long long x = t >= off ? sc[t - off] : 0ull;
        __syncthreads();
        sc[t] += x;
        __syncthreads();
    }
    // sc[255 - d] = S[d]
    const unsigned long long above = st->above, target = st->target;
    if (above + sc[255 - t] > target) atomicAdd(&ntrue, 1);
    __syncthreads();
    w.hist[t] = 0;
    if (t == 0 && !done) {
        if (round == 0 && sc[255] <= target) {   // the whole stream fits
            st->done = 1;
            st->K = 0;
        } else {
            const int dg = ntrue - 1;
            const int shift = 56 - 8 * round;
            st->prefix |= (unsigned long long)dg << shift;
            st->above = above + (dg < 255 ? sc[255 - (dg + 1)] : 0ull);
            if (round == 7) st->K = st->prefix + 1;
        }
    }
}

// Layer `layer` of nlayers: the block's pass count into layer_np (if
// given); the last layer also cuts the block's result.
__global__ void rd_apply_kernel(int ncb, RdWs w, rg_cb_result *__restrict__ res, int32_t *__restrict__ layer_np,
                                int nlayers, int layer)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ncb) return;
    const unsigned long long K = w.st->K;
    const size_t S = (size_t)ncb;
    const int c = w.cnt[b];
    int keep = 0;
    while (keep < c && w.hk[keep * S + b] >= K) ++keep;
    const int nt = keep ? w.hn[(keep - 1) * S + b] : 0;
    const int32_t rt = keep ? w.hr[(keep - 1) * S + b] : 0;
    if (layer_np) layer_np[(size_t)b * nlayers + layer] = nt;
    if (layer == nlayers - 1) {
        res[b].npasses = nt;
        res[b].len = rt;
    }
    if (rt) atomicAdd(&w.st->kept, (unsigned long long)rt);
}

}  // namespace

size_t rg_rd_ws_bytes(int ncb)
{
    const size_t n = (size_t)ncb * MAXP;
    return n * 8 * 3 + n * 4 * 2 + ((size_t)ncb * 4 + 255) / 256 * 256 + 256 * 8 + sizeof(RdState);
}

size_t rg_rd_sppm_bytes(int ncb) { return (size_t)ncb * RD_PLANES * 64 * 8; }

int rg_launch_rate_control(hipStream_t s, const int32_t *q, int pitch, size_t plane_stride, const rg_cb_desc *desc,
                           int ncb, const int4 *cm, const uint64_t *sppm, const int32_t *seglens,
                           const double *weights, int nlayers, const long *targets, void *ws, int32_t *layer_np,
                           rg_cb_result *res, long *kept_out, int *err_out)
{
    if (ncb <= 0) return RG_SUCCESS;
    RG_CHECK(nlayers >= 1 && targets, "rate control: no targets");
    for (int l = 0; l < nlayers; ++l)
        RG_CHECK(targets[l] >= 0 && (l == 0 || targets[l] >= targets[l - 1]),
                 "rate control: targets must be non-negative and non-decreasing");
    const RdWs w = rd_ws(ws, ncb);
    rd_init_kernel<<<1, 256, 0, s>>>(w.st, w.hist, (unsigned long long)targets[0], 1);
    RG_LAUNCH_CHECK();
    rd_dist_kernel<<<dim3((unsigned)ncb), dim3(64), 0, s>>>(q, pitch, plane_stride, desc, cm, sppm, w.dist, w.st);
    RG_LAUNCH_CHECK();
    const dim3 g((unsigned)((ncb + 255) / 256)), b(256);
    rd_hull_kernel<<<g, b, 0, s>>>(ncb, cm, seglens, weights, w);
    RG_LAUNCH_CHECK();
    for (int l = 0; l < nlayers; ++l) {
        if (l) {
            rd_init_kernel<<<1, 256, 0, s>>>(w.st, w.hist, (unsigned long long)targets[l], 0);
            RG_LAUNCH_CHECK();
        }
        for (int round = 0; round < 8; ++round) {
            rd_hist_kernel<<<g, b, 0, s>>>(ncb, round, w);
            RG_LAUNCH_CHECK();
            rd_pick_kernel<<<1, 256, 0, s>>>(round, w);
            RG_LAUNCH_CHECK();
        }
        rd_apply_kernel<<<g, b, 0, s>>>(ncb, w, res, layer_np, nlayers, l);
        RG_LAUNCH_CHECK();
    }
    if (kept_out || err_out) {
        RdState h;
        RG_HIP(hipMemcpyAsync(&h, w.st, sizeof h, hipMemcpyDeviceToHost, s));
        RG_HIP(hipStreamSynchronize(s));
        if (kept_out) *kept_out = (long)h.kept;
        if (err_out) *err_out = (int)h.err;
    }
    return RG_SUCCESS;
}
