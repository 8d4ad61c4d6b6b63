// C ABI of libroger_amd: contexts, device memory, stage entry points, the
// multi-level DWT drivers and the encoder / decoder objects that own their
// device buffers (the role of the reference's OCLMemoryManager,
// src/OCLMemoryManager.cpp:56-148: buffers are reallocated only when the
// image shape grows, so repeated runs keep everything resident in HBM).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <mutex>
#include <vector>
#include <algorithm>
#include "rg_internal.h"
#include "rg_t2.h"

// ------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

int rg_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *rg_last_error(void) { return g_err; }

extern "C" const char *rg_error_string(int code)
{
    switch (code) {
    case RG_SUCCESS: return "success";
    case RG_ERR_INVALID: return "invalid argument";
    case RG_ERR_DEVICE: return "device (HIP) error";
    case RG_ERR_OOM: return "out of memory";
    case RG_ERR_OVERFLOW: return "code-block output overflow";
    case RG_ERR_NOT_READY: return "no result available";
    default: return "unknown error";
    }
}

// ------------------------------------------------------------ static DWT scratch
// Scratch of rg_dwt_fwd_noalloc: the inter-level LL bands live in this
// file-scope device array instead of caller- or library-allocated memory.
// Every GPU the library runs on holds its own copy until the process ends;
// host code uses only the per-device address from hipGetSymbolAddress (the
// array's name is a host-side stand-in, not a device address).
#define RG_NOALLOC_SCRATCH_BYTES ((size_t)96 << 20)
static __device__ __align__(256) unsigned char scratch[RG_NOALLOC_SCRATCH_BYTES];
static constexpr int RG_MAX_DEVICES = 64;
static void *g_noalloc_scratch[RG_MAX_DEVICES];
static std::mutex g_noalloc_mu;

// Looks up this device's copy of the array once, with the device current
// (called from rg_context_create, the library's set-up call).
static int noalloc_scratch_setup(int device)
{
    RG_CHECK(device >= 0 && device < RG_MAX_DEVICES, "device %d beyond the library's limit of %d", device,
             RG_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_scratch[device]) return RG_SUCCESS;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return rg_fail(RG_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_noalloc_scratch[device] = p;
    return RG_SUCCESS;
}

// ------------------------------------------------------------ context
extern "C" int rg_context_create(int device, rg_context **out)
{
    RG_CHECK(out, "rg_context_create: null output");
    int n = 0;
    RG_HIP(hipGetDeviceCount(&n));
    RG_CHECK(device >= 0 && device < n, "rg_context_create: device %d out of range (%d devices)", device, n);
    RG_HIP(hipSetDevice(device));
    if (int rc = noalloc_scratch_setup(device)) return rc;
    rg_context *c = new rg_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_stop);
    if (e != hipSuccess) {
        delete c;
        return rg_fail(RG_ERR_DEVICE, "rg_context_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return RG_SUCCESS;
}

extern "C" int rg_context_destroy(rg_context *c)
{
    if (!c) return RG_SUCCESS;
    hipSetDevice(c->device);
    if (c->stream) { hipStreamSynchronize(c->stream); hipStreamDestroy(c->stream); }
    for (hipStream_t st : c->side)
        if (st) { hipStreamSynchronize(st); hipStreamDestroy(st); }
    if (c->upload) { hipStreamSynchronize(c->upload); hipStreamDestroy(c->upload); }
    if (c->ev_start) hipEventDestroy(c->ev_start);
    if (c->ev_stop) hipEventDestroy(c->ev_stop);
    delete c;
    return RG_SUCCESS;
}

extern "C" int rg_context_sync(rg_context *c)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipStreamSynchronize(c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_context_query(rg_context *c)
{
    RG_CHECK(c, "null context");
    hipError_t e = hipStreamQuery(c->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return rg_fail(RG_ERR_DEVICE, "hipStreamQuery: %s", hipGetErrorString(e));
}

extern "C" int rg_context_device_name(rg_context *c, char *buf, int cap)
{
    RG_CHECK(c && buf && cap > 0, "bad arguments");
    hipDeviceProp_t p;
    RG_HIP(hipGetDeviceProperties(&p, c->device));
    snprintf(buf, cap, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return RG_SUCCESS;
}

// ------------------------------------------------------------ memory
extern "C" int rg_malloc(rg_context *c, size_t bytes, void **p)
{
    RG_CHECK(c && p, "bad arguments");
    RG_HIP(hipSetDevice(c->device));
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return rg_fail(RG_ERR_OOM, "hipMalloc(%zu) failed", bytes);
    }
    return RG_SUCCESS;
}

extern "C" int rg_free(rg_context *c, void *p)
{
    RG_CHECK(c, "null context");
    if (p) RG_HIP(hipFree(p));
    return RG_SUCCESS;
}

extern "C" int rg_memcpy_h2d(rg_context *c, void *dst, const void *src, size_t bytes)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    RG_HIP(hipStreamSynchronize(c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_memcpy_h2d_async(rg_context *c, void *dst, const void *src, size_t bytes)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_memcpy_d2h(rg_context *c, void *dst, const void *src, size_t bytes)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    RG_HIP(hipStreamSynchronize(c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_memcpy_d2d(rg_context *c, void *dst, const void *src, size_t bytes)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_memset(rg_context *c, void *dst, int value, size_t bytes)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipMemsetAsync(dst, value, bytes, c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_host_alloc(rg_context *c, size_t bytes, void **p)
{
    RG_CHECK(c && p, "bad arguments");
    RG_HIP(hipSetDevice(c->device));
    if (hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return rg_fail(RG_ERR_OOM, "hipHostMalloc(%zu) failed", bytes);
    }
    return RG_SUCCESS;
}

extern "C" int rg_host_free(rg_context *c, void *p)
{
    RG_CHECK(c, "null context");
    if (p) RG_HIP(hipHostFree(p));
    return RG_SUCCESS;
}

extern "C" int rg_timer_start(rg_context *c)
{
    RG_CHECK(c, "null context");
    RG_HIP(hipEventRecord(c->ev_start, c->stream));
    return RG_SUCCESS;
}

extern "C" int rg_timer_stop(rg_context *c, float *ms)
{
    RG_CHECK(c && ms, "bad arguments");
    RG_HIP(hipEventRecord(c->ev_stop, c->stream));
    RG_HIP(hipEventSynchronize(c->ev_stop));
    RG_HIP(hipEventElapsedTime(ms, c->ev_start, c->ev_stop));
    return RG_SUCCESS;
}

// ------------------------------------------------------------ stages
extern "C" int rg_colour_fwd(rg_context *c, const uint8_t *src, int w, int h, int ncomp, int colour,
                             int out_float, void *planes)
{
    RG_CHECK(c && src && planes, "rg_colour_fwd: null argument");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "rg_colour_fwd: bad shape %dx%dx%d", w, h, ncomp);
    RG_CHECK(colour >= 0 && colour <= 2, "rg_colour_fwd: bad colour mode %d", colour);
    RG_CHECK(((uintptr_t)src & 15) == 0 && ((uintptr_t)planes & 15) == 0, "rg_colour_fwd: buffers must be 16-byte aligned");
    if (colour == RG_COLOUR_RCT) out_float = 0;
    if (colour == RG_COLOUR_ICT) out_float = 1;
    return rg_launch_colour_fwd(c->stream, src, w, h, ncomp, colour, out_float, planes);
}

extern "C" int rg_colour_inv(rg_context *c, const void *planes, int w, int h, int ncomp, int colour,
                             int in_float, uint8_t *dst)
{
    RG_CHECK(c && dst && planes, "rg_colour_inv: null argument");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "rg_colour_inv: bad shape %dx%dx%d", w, h, ncomp);
    RG_CHECK(colour >= 0 && colour <= 2, "rg_colour_inv: bad colour mode %d", colour);
    RG_CHECK(((uintptr_t)dst & 15) == 0 && ((uintptr_t)planes & 15) == 0, "rg_colour_inv: buffers must be 16-byte aligned");
    if (colour == RG_COLOUR_RCT) in_float = 0;
    if (colour == RG_COLOUR_ICT) in_float = 1;
    return rg_launch_colour_inv(c->stream, planes, w, h, ncomp, colour, in_float, dst);
}

extern "C" int rg_planar_split_i16(rg_context *c, const int16_t *src, int w, int h, int ncomp, int16_t *planes)
{
    RG_CHECK(c && src && planes, "rg_planar_split_i16: null argument");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "rg_planar_split_i16: bad shape");
    return rg_launch_planar_split_i16(c->stream, src, w, h, ncomp, planes);
}

static int check_levels(int w, int h, int levels)
{
    RG_CHECK(w > 0 && h > 0, "bad image size %dx%d", w, h);
    RG_CHECK(levels >= 0 && levels <= 32, "levels %d out of range 0..32", levels);
    return RG_SUCCESS;
}

extern "C" size_t rg_dwt_scratch_bytes(int w, int h, int nplanes, int filter)
{
    (void)filter;
    if (w <= 0 || h <= 0 || nplanes <= 0) return 512;
    return rg_ll_offset(w, h, nplanes, 33) + 512;   // the LL regions of up to 32 levels
}

// Level-0 (no transform) copy / quantisation: plane in -> q.
int rg_launch_copy_quant(hipStream_t s, int filter, const void *in, long n, int nplanes, float inv, int q_float,
                         void *q);
int rg_launch_copy_dequant(hipStream_t s, int filter, const void *q, long n, int nplanes, float step, int q_float,
                           void *out);

// Integer setting from the environment (`def` when unset).
static int rg_env_int(const char *name, int def)
{
    const char *v = getenv(name);
    return v && *v ? atoi(v) : def;
}

// Levels first..levels of the forward transform (first > 1: level first-1's
// LL is already in the scratch buffer, written by the fused first level).
static int dwt_fwd_impl(rg_context *c, int filter, const void *in, int w, int h, int nplanes, int levels,
                        const float *inv_step, int ll_fp16, void *dq, int q_float, void *scratch, int first = 1)
{
    int rc = check_levels(w, h, levels);
    if (rc) return rc;
    RG_CHECK(filter == RG_FILTER_53 || filter == RG_FILTER_97, "bad filter %d", filter);
    RG_CHECK((in || first > 1) && dq && nplanes >= 1, "null buffer");
    RG_CHECK(filter == RG_FILTER_53 || q_float || inv_step, "9/7 quantisation needs inv_step");
    if (levels == 0)
        return rg_launch_copy_quant(c->stream, filter, in, (long)w * h, nplanes,
                                    (filter == RG_FILTER_97 && !q_float) ? inv_step[0] : 1.0f, q_float, dq);
    RG_CHECK(scratch || levels == 1, "scratch buffer required");
    rg_level_dims d = rg_dims(w, h, levels);
    char *sc = reinterpret_cast<char *>(scratch);
    auto ll = [&](int l) { return sc + rg_ll_offset(w, h, nplanes, l); };
    const bool half = filter == RG_FILTER_97 && ll_fp16;
    for (int l = first; l <= levels; ++l) {
        rg_fwd_level_args a{};
        a.filter = filter;
        a.cw = d.w[l - 1]; a.ch = d.h[l - 1]; a.nplanes = nplanes;
        if (l == 1) {
            a.in = in; a.in_pitch = w; a.in_ps = (long)w * h; a.in_half = 0;
        } else {
            a.in = ll(l - 1);
            a.in_pitch = d.w[l - 1]; a.in_ps = (long)d.w[l - 1] * d.h[l - 1]; a.in_half = half;
        }
        if (l < levels) {
            a.ll_out = ll(l);
            a.ll_pitch = d.w[l]; a.ll_ps = (long)d.w[l] * d.h[l]; a.ll_half = half;
        } else {
            a.ll_out = nullptr;
        }
        a.q = dq; a.qp = w; a.q_ps = (long)w * h; a.q_float = q_float;
        if (filter == RG_FILTER_97 && !q_float) {
            const int b = 1 + 3 * (levels - l);  // HL of level l
            a.inv_hl = inv_step[b]; a.inv_lh = inv_step[b + 1]; a.inv_hh = inv_step[b + 2];
            a.inv_ll = inv_step[0];
        }
        rc = filter == RG_FILTER_53 ? rg_launch_fwd53_level(c->stream, a) : rg_launch_fwd97_level(c->stream, a);
        if (rc) return rc;
    }
    return RG_SUCCESS;
}

// Colour stage + forward transform from the interleaved uint8 image: the
// fused first level where rg_fwd1_colour_ok takes the shape (planes left
// unwritten, *fused = true), else the colour kernel into `planes` and the
// level kernels from there.  scratch: the LL pyramid (rg_dwt_scratch_bytes).
// The choice depends on the shape alone, as rg_dwt_fwd_u8_scratch_bytes's
// does: a caller's buffer sized by it always covers the path taken.
static int dwt_fwd_u8_impl(rg_context *c, const uint8_t *src, int w, int h, int ncomp, int colour, int filter,
                           int levels, const float *inv, void *dq, void *scratch, void *planes, bool *fused)
{
    int rc = check_levels(w, h, levels);
    if (rc) return rc;
    const bool f = levels >= 1 && rg_fwd1_colour_ok(w, h, ncomp, colour, filter);
    if (fused) *fused = f;
    if (!f) {
        RG_CHECK(planes, "colour planes buffer required for this shape");
        rc = rg_launch_colour_fwd(c->stream, src, w, h, ncomp, colour, filter == RG_FILTER_97, planes);
        if (rc) return rc;
        return dwt_fwd_impl(c, filter, planes, w, h, ncomp, levels, inv, 0, dq, 0, scratch);
    }
    RG_CHECK(filter == RG_FILTER_53 || inv, "9/7 quantisation needs inv_step");
    RG_CHECK(scratch || levels == 1, "scratch buffer required");
    const rg_level_dims d = rg_dims(w, h, 1);
    rg_fwd1c_args a{};
    a.src = src; a.w = w; a.h = h; a.ncomp = ncomp; a.colour = colour; a.filter = filter;
    a.ll_out = levels > 1 ? (char *)scratch + rg_ll_offset(w, h, ncomp, 1) : nullptr;
    a.ll_pitch = d.w[1]; a.ll_ps = (long)d.w[1] * d.h[1];
    a.q = (int32_t *)dq; a.qp = w; a.q_ps = (long)w * h;
    if (filter == RG_FILTER_97) {
        const int b = 1 + 3 * (levels - 1);   // HL of level 1
        a.inv_hl = inv[b]; a.inv_lh = inv[b + 1]; a.inv_hh = inv[b + 2]; a.inv_ll = inv[0];
    }
    rc = rg_launch_fwd1_colour(c->stream, a);
    if (rc || levels == 1) return rc;
    return dwt_fwd_impl(c, filter, nullptr, w, h, ncomp, levels, inv, 0, dq, 0, scratch, 2);
}

extern "C" size_t rg_dwt_fwd_u8_scratch_bytes(int w, int h, int ncomp, int colour, int filter)
{
    if (w <= 0 || h <= 0 || ncomp <= 0) return 512;
    size_t n = rg_dwt_scratch_bytes(w, h, ncomp, filter);
    if (!rg_fwd1_colour_ok(w, h, ncomp, colour, filter)) n += (size_t)w * h * ncomp * 4 + 256;
    return n;
}

extern "C" int rg_dwt_fwd_u8(rg_context *c, const uint8_t *src, int w, int h, int ncomp, int colour, int filter,
                             int levels, const float *inv_step, int32_t *dq, void *scratch)
{
    RG_CHECK(c && src && dq && scratch, "rg_dwt_fwd_u8: null argument");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "rg_dwt_fwd_u8: bad shape %dx%dx%d", w, h, ncomp);
    RG_CHECK(filter == RG_FILTER_53 || filter == RG_FILTER_97, "bad filter %d", filter);
    RG_CHECK(colour == RG_COLOUR_NONE || (ncomp == 3 && ((colour == RG_COLOUR_RCT && filter == RG_FILTER_53) ||
                                                         (colour == RG_COLOUR_ICT && filter == RG_FILTER_97))),
             "colour transform %d does not pair with filter %d and %d components", colour, filter, ncomp);
    RG_CHECK(((uintptr_t)src & 3) == 0, "rg_dwt_fwd_u8: source must be 4-byte aligned");
    RG_CHECK(levels >= 1, "rg_dwt_fwd_u8: levels %d, at least 1", levels);
    RG_HIP(hipSetDevice(c->device));
    // the planes of an unfused shape sit after the LL pyramid, 256-byte aligned
    const size_t ll_bytes = rg_dwt_scratch_bytes(w, h, ncomp, filter);
    void *planes = (char *)scratch + ((ll_bytes + 255) & ~(size_t)255);
    return dwt_fwd_u8_impl(c, src, w, h, ncomp, colour, filter, levels, inv_step, dq, scratch, planes, nullptr);
}

extern "C" int rg_dwt_fwd(rg_context *c, int filter, const void *in, int w, int h, int nplanes, int levels,
                          const float *inv_step, int ll_fp16, int32_t *dq, void *scratch)
{
    RG_CHECK(c, "null context");
    return dwt_fwd_impl(c, filter, in, w, h, nplanes, levels, inv_step, ll_fp16, dq, 0, scratch);
}

extern "C" int rg_dwt_fwd_noalloc(rg_context *c, int filter, const void *in, int w, int h, int nplanes, int levels,
                                  const float *inv_step, int ll_fp16, int32_t *dq)
{
    RG_CHECK(c, "null context");
    int rc = check_levels(w, h, levels);
    if (rc) return rc;
    RG_CHECK(nplanes >= 1 && nplanes <= 4, "nplanes %d out of range 1..4", nplanes);
    const size_t need = rg_dwt_scratch_bytes(w, h, nplanes, filter);
    RG_CHECK(need <= RG_NOALLOC_SCRATCH_BYTES,
             "rg_dwt_fwd_noalloc: %dx%d x %d planes needs %zu bytes of scratch, the static buffer holds %zu", w, h,
             nplanes, need, RG_NOALLOC_SCRATCH_BYTES);
    void *sc = c->device >= 0 && c->device < RG_MAX_DEVICES ? g_noalloc_scratch[c->device] : nullptr;
    if (!sc) return rg_fail(RG_ERR_NOT_READY, "static scratch of device %d was not set up", c->device);
    return dwt_fwd_impl(c, filter, in, w, h, nplanes, levels, inv_step, ll_fp16, dq, 0, sc);
}

extern "C" int rg_dwt97_fwd_float(rg_context *c, const float *in, int w, int h, int nplanes, int levels, float *out,
                                  void *scratch)
{
    RG_CHECK(c, "null context");
    return dwt_fwd_impl(c, RG_FILTER_97, in, w, h, nplanes, levels, nullptr, 0, out, 1, scratch);
}

// Levels `levels` down to `last` of the inverse transform (last > 1: the
// output is the LL band of level last - 1, left in the scratch buffer for
// the fused last level).
static int dwt_inv_impl(rg_context *c, int filter, const void *dq, int w, int h, int nplanes, int levels,
                        const float *step, void *out, void *scratch, int last = 1)
{
    int rc = check_levels(w, h, levels);
    if (rc) return rc;
    RG_CHECK(filter == RG_FILTER_53 || filter == RG_FILTER_97, "bad filter %d", filter);
    RG_CHECK(dq && (out || last > 1) && nplanes >= 1, "null buffer");
    const int q_float = filter == RG_FILTER_97 && step == nullptr;
    if (levels == 0)
        return rg_launch_copy_dequant(c->stream, filter, dq, (long)w * h, nplanes, step ? step[0] : 1.0f, q_float, out);
    RG_CHECK(scratch || levels == 1, "scratch buffer required");
    rg_level_dims d = rg_dims(w, h, levels);
    char *sc = reinterpret_cast<char *>(scratch);
    auto ll = [&](int l) { return sc + rg_ll_offset(w, h, nplanes, l); };
    for (int l = levels; l >= last; --l) {
        rg_inv_level_args a{};
        a.filter = filter;
        a.cw = d.w[l - 1]; a.ch = d.h[l - 1]; a.nplanes = nplanes;
        if (l == levels) {
            a.ll_is_q = 1;
        } else {
            a.ll_is_q = 0;
            a.ll_in = ll(l);   // LL after level l: the output of level l+1
            a.ll_pitch = d.w[l]; a.ll_ps = (long)d.w[l] * d.h[l];
        }
        a.q = dq; a.qp = w; a.q_ps = (long)w * h; a.q_float = q_float;
        if (filter == RG_FILTER_97 && step) {
            const int b = 1 + 3 * (levels - l);
            a.s_ll = step[0]; a.s_hl = step[b]; a.s_lh = step[b + 1]; a.s_hh = step[b + 2];
        }
        if (l > 1) {
            a.out = ll(l - 1);
            a.out_pitch = d.w[l - 1]; a.out_ps = (long)d.w[l - 1] * d.h[l - 1];
        } else {
            a.out = out; a.out_pitch = w; a.out_ps = (long)w * h;
        }
        rc = filter == RG_FILTER_53 ? rg_launch_inv53_level(c->stream, a) : rg_launch_inv97_level(c->stream, a);
        if (rc) return rc;
    }
    return RG_SUCCESS;
}

extern "C" int rg_dwt_inv(rg_context *c, int filter, const void *dq, int w, int h, int nplanes, int levels,
                          const float *step, void *out, void *scratch)
{
    RG_CHECK(c, "null context");
    return dwt_inv_impl(c, filter, dq, w, h, nplanes, levels, step, out, scratch);
}

// Inverse transform + inverse colour stage into the interleaved uint8 image:
// levels down to 2 as usual, then the fused last level where
// rg_inv1_colour_ok takes the shape (*fused = true), else level 1 into
// `planes` and the inverse colour kernel from there.  9/7 needs `step`
// (quantisation indices).  scratch: the LL pyramid.
static int dwt_inv_u8_impl(rg_context *c, const int32_t *dq, int w, int h, int ncomp, int colour, int filter,
                           int levels, const float *step, uint8_t *dst, void *scratch, void *planes, bool *fused)
{
    int rc = check_levels(w, h, levels);
    if (rc) return rc;
    RG_CHECK(filter == RG_FILTER_53 || step, "9/7 dequantisation needs the band steps");
    const bool f = levels >= 1 && rg_inv1_colour_ok(w, h, ncomp, colour, filter);
    if (fused) *fused = f;
    if (!f) {
        RG_CHECK(planes, "planes buffer required for this shape");
        if ((rc = dwt_inv_impl(c, filter, dq, w, h, ncomp, levels, step, planes, scratch))) return rc;
        return rg_launch_colour_inv(c->stream, planes, w, h, ncomp, colour, filter == RG_FILTER_97, dst);
    }
    RG_CHECK(scratch || levels == 1, "scratch buffer required");
    if (levels > 1 && (rc = dwt_inv_impl(c, filter, dq, w, h, ncomp, levels, step, nullptr, scratch, 2))) return rc;
    const rg_level_dims d = rg_dims(w, h, 1);
    rg_inv1c_args a{};
    a.ll_is_q = levels == 1;
    a.ll_in = levels > 1 ? (const void *)((char *)scratch + rg_ll_offset(w, h, ncomp, 1)) : nullptr;
    a.ll_pitch = d.w[1]; a.ll_ps = (long)d.w[1] * d.h[1];
    a.q = dq; a.qp = w; a.q_ps = (long)w * h;
    a.w = w; a.h = h; a.ncomp = ncomp; a.colour = colour; a.filter = filter;
    if (filter == RG_FILTER_97) {
        const int b = 1 + 3 * (levels - 1);   // HL of level 1
        a.s_ll = step[0]; a.s_hl = step[b]; a.s_lh = step[b + 1]; a.s_hh = step[b + 2];
    }
    a.dst = dst;
    return rg_launch_inv1_colour(c->stream, a);
}

extern "C" size_t rg_dwt_inv_u8_scratch_bytes(int w, int h, int ncomp, int colour, int filter)
{
    if (w <= 0 || h <= 0 || ncomp <= 0) return 512;
    size_t n = rg_dwt_scratch_bytes(w, h, ncomp, filter);
    if (!rg_inv1_colour_ok(w, h, ncomp, colour, filter)) n += (size_t)w * h * ncomp * 4 + 256;
    return n;
}

extern "C" int rg_dwt_inv_u8(rg_context *c, const int32_t *dq, int w, int h, int ncomp, int colour, int filter,
                             int levels, const float *step, uint8_t *dst, void *scratch)
{
    RG_CHECK(c && dq && dst && scratch, "rg_dwt_inv_u8: null argument");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "rg_dwt_inv_u8: bad shape %dx%dx%d", w, h, ncomp);
    RG_CHECK(filter == RG_FILTER_53 || filter == RG_FILTER_97, "bad filter %d", filter);
    RG_CHECK(colour == RG_COLOUR_NONE || (ncomp == 3 && ((colour == RG_COLOUR_RCT && filter == RG_FILTER_53) ||
                                                         (colour == RG_COLOUR_ICT && filter == RG_FILTER_97))),
             "colour transform %d does not pair with filter %d and %d components", colour, filter, ncomp);
    RG_CHECK(((uintptr_t)dst & 15) == 0, "rg_dwt_inv_u8: destination must be 16-byte aligned");
    RG_CHECK(levels >= 1, "rg_dwt_inv_u8: levels %d, at least 1", levels);
    RG_HIP(hipSetDevice(c->device));
    const size_t ll_bytes = rg_dwt_scratch_bytes(w, h, ncomp, filter);
    void *planes = (char *)scratch + ((ll_bytes + 255) & ~(size_t)255);
    return dwt_inv_u8_impl(c, dq, w, h, ncomp, colour, filter, levels, step, dst, scratch, planes, nullptr);
}

// ------------------------------------------------------------ code-blocks
extern "C" int rg_codeblocks(int w, int h, int levels, int cbw, int cbh, rg_cb_desc *out, int cap)
{
    if (check_levels(w, h, levels)) return RG_ERR_INVALID;
    if (cbw < 4 || cbh < 4 || cbw > 64 || cbh > 64 || cbw * cbh > 4096 || (cbw & (cbw - 1)) || (cbh & (cbh - 1)))
        return rg_fail(RG_ERR_INVALID, "code-block size %dx%d not supported", cbw, cbh);
    rg_level_dims d = rg_dims(w, h, levels);
    int n = 0;
    auto add_band = [&](int x0, int y0, int bw, int bh, int orient, int band, int res) {
        if (bw <= 0 || bh <= 0) return;
        for (int j = 0; j < (bh + cbh - 1) / cbh; ++j)
            for (int i = 0; i < (bw + cbw - 1) / cbw; ++i) {
                if (out && n < cap) {
                    rg_cb_desc &c = out[n];
                    c.x0 = x0 + i * cbw; c.y0 = y0 + j * cbh;
                    c.w = std::min(cbw, bw - i * cbw); c.h = std::min(cbh, bh - j * cbh);
                    c.comp = 0; c.orient = orient; c.band = band; c.res = res;
                }
                ++n;
            }
    };
    add_band(0, 0, d.w[levels], d.h[levels], 0, 0, 0);
    int band = 1;
    for (int l = levels; l >= 1; --l) {
        int wl = d.w[l], hl = d.h[l], wp = d.w[l - 1], hp = d.h[l - 1], res = levels - l + 1;
        add_band(wl, 0, wp - wl, hl, 1, band++, res);
        add_band(0, hl, wl, hp - hl, 2, band++, res);
        add_band(wl, hl, wp - wl, hp - hl, 3, band++, res);
    }
    return n;
}

extern "C" int rg_t1_encode(rg_context *c, const int32_t *dq, int pitch, size_t plane_stride, const rg_cb_desc *desc,
                            int ncb, uint8_t *out, int slot_bytes, rg_cb_result *res)
{
    RG_CHECK(c && dq && desc && out && res, "rg_t1_encode: null argument");
    RG_CHECK(ncb >= 0 && slot_bytes > 0 && pitch > 0, "rg_t1_encode: bad sizes");
    // the lane MQ coder writes 16-byte groups: it needs 16-byte aligned slots
    if (rg_t1_inline() || ncb == 0 || (slot_bytes & 15) != 0 || ((uintptr_t)out & 15) != 0)
        return rg_launch_t1_encode(c->stream, dq, pitch, plane_stride, desc, ncb, out, slot_bytes, res);
    // stream-ordered scratch for the symbol streams (cbw, cbh <= 64)
    const int symcap = rg_t1_symcap(64, 64);
    void *syms = nullptr, *cm = nullptr;
    RG_HIP(hipMallocAsync(&syms, (size_t)symcap * ncb, c->stream));
    RG_HIP(hipMallocAsync(&cm, sizeof(int4) * ncb, c->stream));
    int rc = rg_launch_t1_split(c->stream, dq, pitch, plane_stride, desc, ncb, (uint8_t *)syms, symcap, (int4 *)cm, out,
                                slot_bytes, res);
    hipFreeAsync(syms, c->stream);
    hipFreeAsync(cm, c->stream);
    return rc;
}

// ------------------------------------------------------------ buffers
namespace {
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap && p) return RG_SUCCESS;
        if (p) { hipFree(p); p = nullptr; cap = 0; }
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return rg_fail(RG_ERR_OOM, "hipMalloc(%zu) failed", bytes);
        }
        cap = bytes;
        return RG_SUCCESS;
    }
    void release() { if (p) hipFree(p); p = nullptr; cap = 0; }
};

struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap && p) return RG_SUCCESS;
        if (p) { hipHostFree(p); p = nullptr; cap = 0; }
        if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return rg_fail(RG_ERR_OOM, "hipHostMalloc(%zu) failed", bytes);
        }
        cap = bytes;
        return RG_SUCCESS;
    }
    void release() { if (p) hipHostFree(p); p = nullptr; cap = 0; }
};

// Per-band (3*levels+1) quantisation steps: Delta_b = 2^(R_b - eps_b)(1 + mu_b/2^11)
// with (eps_b, mu_b) the T.800 E.1 exponent/mantissa closest to `delta`.
void band_steps(int levels, int bitdepth, double delta, float *step, float *inv)
{
    const int nb = 3 * levels + 1;
    for (int b = 0; b < nb; ++b) {
        int orient = b == 0 ? 0 : 1 + (b - 1) % 3;
        int Rb = bitdepth + (orient == 0 ? 0 : (orient == 3 ? 2 : 1));
        int e;
        double m = frexp(delta, &e);
        m *= 2.0; e -= 1;
        int mu = (int)floor((m - 1.0) * 2048.0 + 0.5);
        if (mu >= 2048) { mu = 0; e += 1; }
        int eps = std::max(0, std::min(31, Rb - e));
        double s = ldexp(1.0 + mu / 2048.0, Rb - eps);
        step[b] = (float)s;
        inv[b] = (float)(1.0 / s);
    }
}
}  // namespace

// ------------------------------------------------------------ encoder
struct rg_encoder {
    rg_context *ctx = nullptr;
    int w = 0, h = 0, ncomp = 0, levels = 0, cbw = 0, cbh = 0, filter = 0;
    int lossy = 0, output_dwt = 0, ran = 0, t1_ran = 0, loaded = 0, i16_ran = 0, bitdepth = 8;
    int ncb_comp = 0, ncb = 0, slot_bytes = 0, colour = 0, style = 0;
    double delta = 1.0 / 64;
    float step[97], inv[97];
    DevBuf input[2], planes, q, fcoeff, scratch, desc, res, slots, chan16, syms, cm, poffs, arena, seglens, chains,
        chains_e[3], cms, sppm, rdws, rdw;
    int rd = 0;   // rate control enabled: the context modeller records significance-pass masks
    int rd_ran = 0;  // the last tier-1 run recorded those masks (sppm holds them for e->ncb blocks)
    int t1_phases = 0;   // modelling phases of the parallel-style tier-1 (0: the library default)
    int nlayers = 1;               // quality layers of the last rate control (1 after a new encode)
    std::vector<int32_t> layer_np; // per block, passes after each layer (nlayers > 1)
    DevBuf d_layer_np;
    HostBuf staging, map;
    // double-buffered ingest of host sources: input[k] is filled on the
    // context's upload stream (event up[k]) and released by the colour kernel
    // that reads it (event used[k]); consecutive encodes alternate k
    hipEvent_t ev_up[2] = {nullptr, nullptr}, ev_used[2] = {nullptr, nullptr};
    int in_slot = 0, used_rec[2] = {0, 0};
    // level-0 planes: rg_encoder_encode_u8 fuses the colour stage into the
    // first DWT level and leaves `planes` unwritten; they are computed from
    // the encode's source (src_dev, input slot src_slot or -1: the caller's
    // device buffer) only when asked for (ensure_planes)
    int planes_valid = 0, src_slot = -1;
    const uint8_t *src_dev = nullptr;
    // events of the overlapped phased tier-1 (created on first use; its side
    // streams belong to the context)
    hipEvent_t ev_in[3] = {nullptr, nullptr, nullptr}, ev_out[3] = {nullptr, nullptr, nullptr};
    // code-block descriptor tables, one per image shape seen (key: w, h,
    // levels, cbw, cbh, ncomp), kept in device memory so that a change of
    // shape queues no copy and waits for nothing; the least recently used
    // table is dropped beyond RG_DESC_TABLES
    struct DescTable {
        int key[6];
        int ncb_comp;
        std::vector<rg_cb_desc> h;
        void *d = nullptr;
        uint64_t used = 0;
    };
    std::vector<DescTable *> dtabs;
    DescTable *cur = nullptr;
    uint64_t dclock = 0;
};

extern "C" int rg_encoder_create(rg_context *c, rg_encoder **out)
{
    RG_CHECK(c && out, "rg_encoder_create: null argument");
    rg_encoder *e = new rg_encoder();
    e->ctx = c;
    *out = e;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_destroy(rg_encoder *e)
{
    if (!e) return RG_SUCCESS;
    hipSetDevice(e->ctx->device);
    hipStreamSynchronize(e->ctx->stream);
    for (hipStream_t st : e->ctx->side)
        if (st) hipStreamSynchronize(st);
    for (DevBuf *b : {&e->input[0], &e->input[1], &e->planes, &e->q, &e->fcoeff, &e->scratch, &e->res, &e->slots, &e->chan16,
                      &e->syms, &e->cm, &e->poffs, &e->arena, &e->seglens, &e->chains, &e->chains_e[0], &e->chains_e[1], &e->chains_e[2],
                      &e->cms, &e->sppm, &e->rdws, &e->rdw, &e->d_layer_np})
        b->release();
    for (int k = 0; k < 3; ++k) {
        if (e->ev_in[k]) hipEventDestroy(e->ev_in[k]);
        if (e->ev_out[k]) hipEventDestroy(e->ev_out[k]);
    }
    for (int k = 0; k < 2; ++k) {
        if (e->ev_up[k]) hipEventDestroy(e->ev_up[k]);
        if (e->ev_used[k]) hipEventDestroy(e->ev_used[k]);
    }
    for (rg_encoder::DescTable *t : e->dtabs) {
        if (t->d) hipFree(t->d);
        delete t;
    }
    e->staging.release();
    e->map.release();
    delete e;
    return RG_SUCCESS;
}

constexpr size_t RG_DESC_TABLES = 512;

static int prepare_codeblocks(rg_encoder *e)
{
    const int key[6] = {e->w, e->h, e->levels, e->cbw, e->cbh, e->ncomp};
    rg_encoder::DescTable *t = nullptr;
    if (e->cur && memcmp(key, e->cur->key, sizeof(key)) == 0) {
        t = e->cur;
    } else {
        for (rg_encoder::DescTable *x : e->dtabs)
            if (memcmp(key, x->key, sizeof(key)) == 0) { t = x; break; }
    }
    if (!t) {
        const int n = rg_codeblocks(e->w, e->h, e->levels, e->cbw, e->cbh, nullptr, 0);
        if (n < 0) return n;
        if (e->dtabs.size() >= RG_DESC_TABLES) {
            // drop the least recently used table once queued work is done with it
            auto lru = std::min_element(e->dtabs.begin(), e->dtabs.end(),
                                        [](const rg_encoder::DescTable *a, const rg_encoder::DescTable *b) {
                                            return a->used < b->used;
                                        });
            RG_HIP(hipStreamSynchronize(e->ctx->stream));
            if ((*lru)->d) RG_HIP(hipFree((*lru)->d));
            if (e->cur == *lru) e->cur = nullptr;
            delete *lru;
            e->dtabs.erase(lru);
        }
        t = new rg_encoder::DescTable();
        memcpy(t->key, key, sizeof(key));
        t->ncb_comp = n;
        t->h.resize((size_t)n * e->ncomp);
        rg_codeblocks(e->w, e->h, e->levels, e->cbw, e->cbh, t->h.data(), n);
        for (int c = 1; c < e->ncomp; ++c)
            for (int i = 0; i < n; ++i) { t->h[(size_t)c * n + i] = t->h[i]; t->h[(size_t)c * n + i].comp = c; }
        const size_t bytes = sizeof(rg_cb_desc) * t->h.size();
        if (hipMalloc(&t->d, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            delete t;
            return rg_fail(RG_ERR_OOM, "hipMalloc(%zu) failed", bytes);
        }
        // t->h stays unchanged while the table lives, so the copy needs no wait
        if (hipMemcpyAsync(t->d, t->h.data(), bytes, hipMemcpyHostToDevice, e->ctx->stream) != hipSuccess) {
            hipFree(t->d);
            delete t;
            return rg_fail(RG_ERR_DEVICE, "hipMemcpyAsync of the code-block table failed");
        }
        e->dtabs.push_back(t);
    }
    t->used = ++e->dclock;
    e->cur = t;
    e->ncb_comp = t->ncb_comp;
    e->ncb = t->ncb_comp * e->ncomp;
    return RG_SUCCESS;
}

static int run_t1(rg_encoder *e)
{
    e->nlayers = 1;
    e->layer_np.clear();
    e->rd_ran = 0;
    int rc = prepare_codeblocks(e);
    if (rc) return rc;
    e->slot_bytes = std::max(256, e->cbw * e->cbh * 3);
    if ((rc = e->res.ensure(sizeof(rg_cb_result) * e->ncb))) return rc;
    if ((rc = e->slots.ensure((size_t)e->slot_bytes * e->ncb))) return rc;
    if (e->style & RG_STYLE_TERMALL) {
        const int symcap = rg_t1_symcap(e->cbw, e->cbh), maxp = rg_t1_maxp();
        if ((rc = e->syms.ensure((size_t)symcap * e->ncb)) || (rc = e->cm.ensure(sizeof(int4) * e->ncb)) ||
            (rc = e->poffs.ensure(sizeof(int32_t) * maxp * e->ncb)) ||
            (rc = e->seglens.ensure(sizeof(int32_t) * maxp * e->ncb)) ||
            (rc = e->arena.ensure(rg_t1_seg_bytes(symcap) * e->ncb)) ||
            (rc = e->chains.ensure(rg_t1_chain_ws_bytes(e->ncb))))
            return rc;
        rg_t1_overlap ov{};
        const int phases = e->t1_phases ? e->t1_phases : rg_t1_phases(e->style);
        const bool two = phases > 1;
        if (two) {
            ov.nearly = phases - 1;
            if ((rc = e->cms.ensure((size_t)e->ncb * 64 * 16))) return rc;
            for (int k = 0; k < ov.nearly; ++k) {
                if ((rc = e->chains_e[k].ensure(rg_t1_chain_ws_bytes(e->ncb)))) return rc;
                if (!e->ctx->side[k]) RG_HIP(hipStreamCreateWithFlags(&e->ctx->side[k], hipStreamNonBlocking));
                if (!e->ev_in[k]) RG_HIP(hipEventCreateWithFlags(&e->ev_in[k], hipEventDisableTiming));
                if (!e->ev_out[k]) RG_HIP(hipEventCreateWithFlags(&e->ev_out[k], hipEventDisableTiming));
                ov.se[k] = e->ctx->side[k]; ov.ev_in[k] = e->ev_in[k]; ov.ev_out[k] = e->ev_out[k];
                ov.ws[k] = (int32_t *)e->chains_e[k].p;
            }
            ov.cms = e->cms.p;
        }
        if (e->rd && (rc = e->sppm.ensure(rg_rd_sppm_bytes(e->ncb)))) return rc;
        rc = rg_launch_t1_parallel(e->ctx->stream, (const int32_t *)e->q.p, e->w, (size_t)e->w * e->h,
                                   (const rg_cb_desc *)e->cur->d, e->ncb, (uint8_t *)e->syms.p, symcap,
                                   (int4 *)e->cm.p, (int32_t *)e->poffs.p, (uint8_t *)e->arena.p,
                                   (int32_t *)e->seglens.p, (uint8_t *)e->slots.p, e->slot_bytes,
                                   (rg_cb_result *)e->res.p, e->style, (int32_t *)e->chains.p, two ? &ov : nullptr,
                                   e->rd ? (uint64_t *)e->sppm.p : nullptr);
        if (rc == RG_SUCCESS) {
            e->t1_ran = 1;
            e->rd_ran = e->rd;
        }
        return rc;
    }
    if (!rg_t1_inline()) {
        const int symcap = rg_t1_symcap(e->cbw, e->cbh);
        if ((rc = e->syms.ensure((size_t)symcap * e->ncb)) || (rc = e->cm.ensure(sizeof(int4) * e->ncb))) return rc;
        rc = rg_launch_t1_split(e->ctx->stream, (const int32_t *)e->q.p, e->w, (size_t)e->w * e->h,
                                (const rg_cb_desc *)e->cur->d, e->ncb, (uint8_t *)e->syms.p, symcap, (int4 *)e->cm.p,
                                (uint8_t *)e->slots.p, e->slot_bytes, (rg_cb_result *)e->res.p);
        if (rc == RG_SUCCESS) e->t1_ran = 1;
        return rc;
    }
    rc = rg_launch_t1_encode(e->ctx->stream, (const int32_t *)e->q.p, e->w, (size_t)e->w * e->h,
                             (const rg_cb_desc *)e->cur->d, e->ncb, (uint8_t *)e->slots.p, e->slot_bytes,
                             (rg_cb_result *)e->res.p);
    if (rc == RG_SUCCESS) e->t1_ran = 1;
    return rc;
}

static int check_params(const rg_encode_params *p, int w, int h, int ncomp)
{
    RG_CHECK(p, "null params");
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "bad image shape %dx%dx%d", w, h, ncomp);
    RG_CHECK(p->levels >= 0 && p->levels <= 15, "levels %d out of range 0..15", p->levels);
    RG_CHECK(p->filter == RG_FILTER_53 || p->filter == RG_FILTER_97, "bad filter %d", p->filter);
    RG_CHECK(p->colour == RG_COLOUR_NONE || ncomp == 3, "colour transform needs 3 components");
    RG_CHECK(p->colour != RG_COLOUR_RCT || p->filter == RG_FILTER_53, "RCT pairs with the 5/3 filter");
    RG_CHECK(p->colour != RG_COLOUR_ICT || p->filter == RG_FILTER_97, "ICT pairs with the 9/7 filter");
    RG_CHECK(p->filter == RG_FILTER_53 || (p->delta > 0 && std::isfinite(p->delta)), "bad quantisation step");
    RG_CHECK(p->bitdepth >= 1 && p->bitdepth <= 16, "bad bitdepth %d", p->bitdepth);
    RG_CHECK(p->style == 0 || p->style == RG_STYLE_PARALLEL || p->style == RG_STYLE_PARALLEL_BYPASS,
             "code-block style %d: only 0, RG_STYLE_PARALLEL and RG_STYLE_PARALLEL_BYPASS", p->style);
    if (rg_codeblocks(w, h, p->levels, p->cbw, p->cbh, nullptr, 0) < 0) return RG_ERR_INVALID;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_encode_u8(rg_encoder *e, const uint8_t *src, int on_device, int w, int h, int ncomp,
                                    const rg_encode_params *p)
{
    RG_CHECK(e && src, "rg_encoder_encode_u8: null argument");
    int rc = check_params(p, w, h, ncomp);
    if (rc) return rc;
    RG_HIP(hipSetDevice(e->ctx->device));
    const size_t npix = (size_t)w * h;
    e->w = w; e->h = h; e->ncomp = ncomp; e->levels = p->levels; e->cbw = p->cbw; e->cbh = p->cbh;
    e->filter = p->filter; e->lossy = p->filter == RG_FILTER_97; e->output_dwt = 0; e->bitdepth = p->bitdepth;
    e->colour = p->colour; e->delta = p->delta; e->style = p->style;
    e->ran = 0; e->t1_ran = 0; e->rd_ran = 0; e->loaded = 1; e->i16_ran = 0;
    band_steps(p->levels, p->bitdepth, p->delta, e->step, e->inv);
    const uint8_t *d_src = src;
    int slot = -1;
    if (!on_device) {
        // host source: upload on the context's upload stream into the input
        // buffer the encode before last used, so this copy overlaps the
        // previous encode's kernels; the main stream waits for it
        slot = e->in_slot;
        e->in_slot ^= 1;
        if ((rc = e->input[slot].ensure(npix * ncomp))) return rc;
        rg_context *c = e->ctx;
        if (!c->upload) RG_HIP(hipStreamCreateWithFlags(&c->upload, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k) {
            if (!e->ev_up[k]) RG_HIP(hipEventCreateWithFlags(&e->ev_up[k], hipEventDisableTiming));
            if (!e->ev_used[k]) RG_HIP(hipEventCreateWithFlags(&e->ev_used[k], hipEventDisableTiming));
        }
        if (e->used_rec[slot]) RG_HIP(hipStreamWaitEvent(c->upload, e->ev_used[slot], 0));
        RG_HIP(hipMemcpyAsync(e->input[slot].p, src, npix * ncomp, hipMemcpyHostToDevice, c->upload));
        RG_HIP(hipEventRecord(e->ev_up[slot], c->upload));
        RG_HIP(hipStreamWaitEvent(c->stream, e->ev_up[slot], 0));
        d_src = (const uint8_t *)e->input[slot].p;
    }
    if ((rc = e->q.ensure(npix * ncomp * 4))) return rc;
    if ((rc = e->scratch.ensure(rg_dwt_scratch_bytes(w, h, ncomp, p->filter)))) return rc;
    e->src_dev = d_src; e->src_slot = slot;
    if (p->ll_fp16 || p->levels == 0 || !rg_fwd1_colour_ok(w, h, ncomp, p->colour, p->filter) ||
        !rg_env_int("RG_DWT_FUSED", 1)) {
        // FP16 inter-level LL, no transform, or a shape the fused first level
        // does not take: colour planes first
        if ((rc = e->planes.ensure(npix * ncomp * 4))) return rc;
        rc = rg_launch_colour_fwd(e->ctx->stream, d_src, w, h, ncomp, p->colour, p->filter == RG_FILTER_97,
                                  e->planes.p);
        if (rc) return rc;
        if (slot >= 0) {
            RG_HIP(hipEventRecord(e->ev_used[slot], e->ctx->stream));
            e->used_rec[slot] = 1;
        }
        e->planes_valid = 1;
        rc = dwt_fwd_impl(e->ctx, p->filter, e->planes.p, w, h, ncomp, p->levels, e->inv, p->ll_fp16, e->q.p, 0,
                          e->scratch.p);
    } else {
        // the colour stage inside the first DWT level: the image is read once
        // and the planes are computed only if asked for (ensure_planes)
        rc = dwt_fwd_u8_impl(e->ctx, d_src, w, h, ncomp, p->colour, p->filter, p->levels, e->inv, e->q.p,
                             e->scratch.p, nullptr, nullptr);
        if (slot >= 0) {   // the transform has read the input: a later upload into it may go ahead
            RG_HIP(hipEventRecord(e->ev_used[slot], e->ctx->stream));
            e->used_rec[slot] = 1;
        }
        e->planes_valid = 0;
    }
    if (rc) return rc;
    e->ran = 1;
    return run_t1(e);
}

// The level-0 planes of an rg_encoder_encode_u8 encode, computed on first
// request from the image that encode read (its own input buffer for a host
// source, which stays untouched until the encode after next; the caller's
// device buffer otherwise, which must still hold the image).
static int ensure_planes(rg_encoder *e)
{
    if (e->planes_valid) return RG_SUCCESS;
    if (!e->src_dev) return rg_fail(RG_ERR_NOT_READY, "no level-0 planes and no source image");
    const size_t npix = (size_t)e->w * e->h;
    int rc;
    if ((rc = e->planes.ensure(npix * e->ncomp * 4))) return rc;
    rc = rg_launch_colour_fwd(e->ctx->stream, e->src_dev, e->w, e->h, e->ncomp, e->colour,
                              e->filter == RG_FILTER_97, e->planes.p);
    if (rc) return rc;
    if (e->src_slot >= 0) {   // a later upload into this input waits for this read too
        RG_HIP(hipEventRecord(e->ev_used[e->src_slot], e->ctx->stream));
        e->used_rec[e->src_slot] = 1;
    }
    e->planes_valid = 1;
    return RG_SUCCESS;
}

// Host conversion of T components into the kernels' sample type (float for
// 9/7, int32 for 5/3) in pinned staging memory.
static int stage_components(HostBuf &staging, hipStream_t s, const void *const *comps, int ncomp, int type,
                            size_t npix, int lossy)
{
    int rc;
    if ((rc = staging.ensure(npix * ncomp * 4))) return rc;
    RG_HIP(hipStreamSynchronize(s));  // staging may still feed a previous copy
    for (int c = 0; c < ncomp; ++c) {
        if (lossy) {
            float *dst = (float *)staging.p + c * npix;
            if (type == RG_F32) memcpy(dst, comps[c], npix * 4);
            else if (type == RG_I16) for (size_t i = 0; i < npix; ++i) dst[i] = ((const int16_t *)comps[c])[i];
            else for (size_t i = 0; i < npix; ++i) dst[i] = (float)((const int32_t *)comps[c])[i];
        } else {
            int32_t *dst = (int32_t *)staging.p + c * npix;
            if (type == RG_F32) for (size_t i = 0; i < npix; ++i) dst[i] = (int32_t)lrintf(((const float *)comps[c])[i]);
            else if (type == RG_I16) for (size_t i = 0; i < npix; ++i) dst[i] = ((const int16_t *)comps[c])[i];
            else memcpy(dst, comps[c], npix * 4);
        }
    }
    return RG_SUCCESS;
}

extern "C" int rg_encoder_load_planar(rg_encoder *e, const void *const *comps, int ncomp, int type, int w, int h,
                                      int precision, int lossy)
{
    RG_CHECK(e && comps, "rg_encoder_load_planar: null argument");
    RG_CHECK(type == RG_F32 || type == RG_I16 || type == RG_I32, "bad sample type %d", type);
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "bad image shape %dx%dx%d", w, h, ncomp);
    RG_CHECK(precision >= 0 && precision <= 16, "bad precision %d", precision);
    for (int c = 0; c < ncomp; ++c) RG_CHECK(comps[c], "component %d is null", c);
    RG_HIP(hipSetDevice(e->ctx->device));
    const size_t npix = (size_t)w * h;
    int rc = stage_components(e->staging, e->ctx->stream, comps, ncomp, type, npix, lossy);
    if (rc) return rc;
    if ((rc = e->planes.ensure(npix * ncomp * 4))) return rc;
    RG_HIP(hipMemcpyAsync(e->planes.p, e->staging.p, npix * ncomp * 4, hipMemcpyHostToDevice, e->ctx->stream));
    e->planes_valid = 1; e->src_dev = nullptr; e->src_slot = -1;
    e->w = w; e->h = h; e->ncomp = ncomp; e->lossy = lossy; e->bitdepth = precision > 0 ? precision : 8;
    e->filter = lossy ? RG_FILTER_97 : RG_FILTER_53;
    e->colour = RG_COLOUR_NONE;
    e->style = 0;   // the stage path (OCLEncoder::run) codes one codeword per block
    e->loaded = 1; e->ran = 0; e->t1_ran = 0; e->rd_ran = 0; e->i16_ran = 0;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_dwt(rg_encoder *e, int levels, int output_dwt, double delta)
{
    RG_CHECK(e, "null encoder");
    if (!e->loaded) return rg_fail(RG_ERR_NOT_READY, "no image loaded");
    RG_CHECK(levels >= 0 && levels <= 15, "levels %d out of range 0..15", levels);
    RG_CHECK(!e->lossy || (delta > 0 && std::isfinite(delta)), "bad quantisation step");
    const size_t npix = (size_t)e->w * e->h;
    int rc;
    if ((rc = ensure_planes(e))) return rc;
    if ((rc = e->q.ensure(npix * e->ncomp * 4))) return rc;
    if ((rc = e->scratch.ensure(rg_dwt_scratch_bytes(e->w, e->h, e->ncomp, e->filter)))) return rc;
    e->levels = levels;
    e->output_dwt = e->lossy && output_dwt;
    e->t1_ran = 0; e->rd_ran = 0; e->i16_ran = 0;
    band_steps(levels, e->bitdepth, delta, e->step, e->inv);
    e->delta = delta;
    if (e->output_dwt) {
        if ((rc = e->fcoeff.ensure(npix * e->ncomp * 4))) return rc;
        rc = dwt_fwd_impl(e->ctx, RG_FILTER_97, e->planes.p, e->w, e->h, e->ncomp, levels, nullptr, 0, e->fcoeff.p, 1,
                          e->scratch.p);
    } else {
        rc = dwt_fwd_impl(e->ctx, e->filter, e->planes.p, e->w, e->h, e->ncomp, levels, e->inv, 0, e->q.p, 0,
                          e->scratch.p);
    }
    if (rc) return rc;
    e->ran = 1;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_bpc(rg_encoder *e, int cbw, int cbh)
{
    RG_CHECK(e, "null encoder");
    if (!e->ran || e->output_dwt) return rg_fail(RG_ERR_NOT_READY, "no quantised coefficients to code");
    if (rg_codeblocks(e->w, e->h, e->levels, cbw, cbh, nullptr, 0) < 0) return RG_ERR_INVALID;
    e->cbw = cbw; e->cbh = cbh;
    return run_t1(e);
}

// ---- rate control ----

// Energy of the 1-D synthesis basis function of a coefficient of the lowpass
// (hi = 0) or highpass band at decomposition level `level`: the band's
// synthesis filter, then level-1 lowpass synthesis steps (upsample by two,
// filter).  Filters: T.800's 5/3 (g0 = [1 2 1]/2, g1 = [-1 -2 6 -2 -1]/8) and
// CDF 9/7 synthesis pairs with unit DC gain of the lowpass.
static double synth_energy(int filter, int level, int hi)
{
    static const double g0_53[] = {0.5, 1.0, 0.5};
    static const double g1_53[] = {-0.125, -0.25, 0.75, -0.25, -0.125};
    static const double g0_97[] = {-0.091271763114250, -0.057543526228500, 0.591271763114250, 1.115087052457000,
                                   0.591271763114250, -0.057543526228500, -0.091271763114250};
    static const double g1_97[] = {0.026748757410810, 0.016864118442875, -0.078223266528990, -0.266864118442875,
                                   0.602949018236360, -0.266864118442875, -0.078223266528990, 0.016864118442875,
                                   0.026748757410810};
    const double *g0 = filter == RG_FILTER_97 ? g0_97 : g0_53, *g1 = filter == RG_FILTER_97 ? g1_97 : g1_53;
    const int n0 = filter == RG_FILTER_97 ? 7 : 3, n1 = filter == RG_FILTER_97 ? 9 : 5;
    std::vector<double> v(hi ? g1 : g0, (hi ? g1 : g0) + (hi ? n1 : n0));
    for (int l = 1; l < level; ++l) {
        std::vector<double> u(2 * v.size() - 1, 0.0), o(2 * v.size() - 1 + n0 - 1, 0.0);
        for (size_t i = 0; i < v.size(); ++i) u[2 * i] = v[i];
        for (size_t i = 0; i < u.size(); ++i)
            for (int k = 0; k < n0; ++k) o[i + k] += u[i] * g0[k];
        v.swap(o);
    }
    double e = 0;
    for (double x : v) e += x * x;
    return e;
}

// Default weight of each code-block: squared-error contribution in the image
// domain of one squared doubled index unit, step^2 * (horizontal synthesis
// energy) * (vertical synthesis energy) / 4 (components weighted equally).
static void default_rd_weights(const rg_encoder *e, std::vector<double> &w)
{
    w.resize(e->ncb);
    // one weight per band (3 * levels + 1 of them), then per block
    std::vector<double> band_w(3 * e->levels + 1, 0.0);
    for (int band = 0; band <= 3 * e->levels; ++band) {
        const int orient = band == 0 ? 0 : 1 + (band - 1) % 3;
        const int level = band == 0 ? e->levels : e->levels - (band - 1) / 3;
        const int hx = orient == 1 || orient == 3, hy = orient == 2 || orient == 3;
        double g = 1.0;
        if (e->levels > 0) g = synth_energy(e->filter, level, hx) * synth_energy(e->filter, level, hy);
        const double st = e->lossy ? (double)e->step[band] : 1.0;
        band_w[band] = st * st * g * 0.25;
    }
    for (int i = 0; i < e->ncb; ++i) w[i] = band_w[e->cur->h[i].band];
}

extern "C" int rg_encoder_set_t1_phases(rg_encoder *e, int phases)
{
    RG_CHECK(e, "null encoder");
    RG_CHECK(phases >= 0 && phases <= RG_T1_MAX_EARLY + 1, "phases %d out of range 0..%d", phases,
             RG_T1_MAX_EARLY + 1);
    e->t1_phases = phases;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_enable_rate_control(rg_encoder *e, int on)
{
    RG_CHECK(e, "null encoder");
    e->rd = on ? 1 : 0;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_rate_control(rg_encoder *e, long target_bytes, const double *weights, long *kept)
{
    return rg_encoder_rate_control_layers(e, 1, &target_bytes, weights, kept);
}

extern "C" int rg_encoder_rate_control_layers(rg_encoder *e, int nlayers, const long *targets, const double *weights,
                                              long *kept)
{
    RG_CHECK(e, "null encoder");
    if (!e->t1_ran) return rg_fail(RG_ERR_NOT_READY, "tier-1 has not run");
    RG_CHECK(e->style & RG_STYLE_TERMALL, "rate control needs a parallel code-block style (TERMALL)");
    // the masks must come from the tier-1 run whose codewords are cut: enabling
    // rate control after an encode leaves none (or a stale, smaller set)
    if (!e->rd_ran) return rg_fail(RG_ERR_NOT_READY, "rate control was not enabled before tier-1 ran");
    RG_CHECK(nlayers >= 1 && nlayers <= 65535 && targets, "bad layer count %d", nlayers);
    for (int l = 0; l < nlayers; ++l)
        RG_CHECK(targets[l] >= 0 && (l == 0 || targets[l] >= targets[l - 1]),
                 "rate targets must be non-negative and non-decreasing");
    std::vector<double> w;
    if (weights) {
        w.assign(weights, weights + e->ncb);
        for (int i = 0; i < e->ncb; ++i)
            RG_CHECK(std::isfinite(w[i]) && w[i] > 0, "block weight %d is not a positive finite number", i);
    } else {
        default_rd_weights(e, w);
    }
    int rc;
    if ((rc = e->rdws.ensure(rg_rd_ws_bytes(e->ncb))) || (rc = e->rdw.ensure(sizeof(double) * e->ncb)) ||
        (nlayers > 1 && (rc = e->d_layer_np.ensure(sizeof(int32_t) * (size_t)e->ncb * nlayers))))
        return rc;
    RG_HIP(hipSetDevice(e->ctx->device));
    RG_HIP(hipMemcpyAsync(e->rdw.p, w.data(), sizeof(double) * e->ncb, hipMemcpyHostToDevice, e->ctx->stream));
    long k = 0;
    int err = 0;
    rc = rg_launch_rate_control(e->ctx->stream, (const int32_t *)e->q.p, e->w, (size_t)e->w * e->h,
                                (const rg_cb_desc *)e->cur->d, e->ncb, (const int4 *)e->cm.p,
                                (const uint64_t *)e->sppm.p, (const int32_t *)e->seglens.p, (const double *)e->rdw.p,
                                nlayers, targets, e->rdws.p, nlayers > 1 ? (int32_t *)e->d_layer_np.p : nullptr,
                                (rg_cb_result *)e->res.p, &k, &err);
    if (rc) return rc;
    if (err) return rg_fail(RG_ERR_INVALID, "rate control needs coefficient magnitudes below 2^24");
    e->nlayers = nlayers;
    e->layer_np.clear();
    if (nlayers > 1) {
        e->layer_np.resize((size_t)e->ncb * nlayers);
        RG_HIP(hipMemcpy(e->layer_np.data(), e->d_layer_np.p, sizeof(int32_t) * e->layer_np.size(),
                         hipMemcpyDeviceToHost));
    }
    if (kept) *kept = k;
    return RG_SUCCESS;
}

int rg_launch_sat_i16(hipStream_t s, const int32_t *src, long n, int16_t *dst);

extern "C" int rg_encoder_planar_i16(rg_encoder *e)
{
    RG_CHECK(e, "null encoder");
    if (!e->ran || e->output_dwt) return rg_fail(RG_ERR_NOT_READY, "no quantised coefficients");
    const size_t n = (size_t)e->w * e->h * e->ncomp;
    int rc;
    if ((rc = e->chan16.ensure(n * 2))) return rc;
    rc = rg_launch_sat_i16(e->ctx->stream, (const int32_t *)e->q.p, (long)n, (int16_t *)e->chan16.p);
    if (rc) return rc;
    e->i16_ran = 1;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_channel_i16(rg_encoder *e, int c, int16_t **d_plane)
{
    RG_CHECK(e && d_plane, "null argument");
    if (!e->i16_ran) return rg_fail(RG_ERR_NOT_READY, "planar int16 channels not produced");
    RG_CHECK(c >= 0 && c < e->ncomp, "channel %d out of range", c);
    *d_plane = (int16_t *)e->chan16.p + (size_t)c * e->w * e->h;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_run_planar(rg_encoder *e, const void *const *comps, int ncomp, int type, int w, int h,
                                     int levels, int precision, int lossy, int output_dwt, int cbw, int cbh,
                                     double delta)
{
    int rc = rg_encoder_load_planar(e, comps, ncomp, type, w, h, precision, lossy);
    if (rc) return rc;
    if ((rc = rg_encoder_dwt(e, levels, output_dwt, delta))) return rc;
    if (e->output_dwt) return RG_SUCCESS;  // outputDwt: no bit-plane coder (src/OCLEncoder.cpp:46-50)
    if ((rc = rg_encoder_planar_i16(e))) return rc;
    return rg_encoder_bpc(e, cbw, cbh);
}

extern "C" int rg_encoder_coefficients(rg_encoder *e, int32_t **dq, int *ncomp, int *w, int *h)
{
    RG_CHECK(e, "null encoder");
    if (!e->ran) return rg_fail(RG_ERR_NOT_READY, "encoder has not run");
    if (dq) *dq = (int32_t *)e->q.p;
    if (ncomp) *ncomp = e->ncomp;
    if (w) *w = e->w;
    if (h) *h = e->h;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_float_coefficients(rg_encoder *e, float **d)
{
    RG_CHECK(e && d, "null argument");
    if (!e->ran || !e->output_dwt) return rg_fail(RG_ERR_NOT_READY, "no float coefficients (output_dwt not set)");
    *d = (float *)e->fcoeff.p;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_dwt_in(rg_encoder *e, int level, void **d_ptr, int *w, int *h)
{
    RG_CHECK(e && d_ptr, "null argument");
    *d_ptr = nullptr;
    if (!e->loaded) return rg_fail(RG_ERR_NOT_READY, "no image loaded");
    RG_CHECK(level >= 0, "level %d is negative", level);
    if (level > 0) {
        if (!e->ran) return rg_fail(RG_ERR_NOT_READY, "the transform has not run");
        if (level >= e->levels)
            return rg_fail(RG_ERR_INVALID, "level %d: the %d-level transform has inputs 0..%d", level, e->levels,
                           e->levels - 1);
    }
    if (level == 0) {
        const int rc = ensure_planes(e);
        if (rc) return rc;
    }
    const rg_level_dims d = rg_dims(e->w, e->h, level);
    *d_ptr = level == 0 ? e->planes.p : (char *)e->scratch.p + rg_ll_offset(e->w, e->h, e->ncomp, level);
    if (w) *w = d.w[level];
    if (h) *h = d.h[level];
    return RG_SUCCESS;
}

extern "C" int rg_encoder_num_codeblocks(rg_encoder *e)
{
    RG_CHECK(e, "null encoder");
    return e->t1_ran ? e->ncb : 0;
}

extern "C" int rg_encoder_get_codeblocks(rg_encoder *e, rg_cb_desc *desc, rg_cb_result *res, long *offsets,
                                         uint8_t *data, long cap, long *total)
{
    RG_CHECK(e, "null encoder");
    if (!e->t1_ran) return rg_fail(RG_ERR_NOT_READY, "tier-1 has not run");
    std::vector<rg_cb_result> r(e->ncb);
    RG_HIP(hipMemcpyAsync(r.data(), e->res.p, sizeof(rg_cb_result) * e->ncb, hipMemcpyDeviceToHost, e->ctx->stream));
    RG_HIP(hipStreamSynchronize(e->ctx->stream));
    long sum = 0;
    for (int i = 0; i < e->ncb; ++i) {
        if (r[i].flags & 1)
            return rg_fail(RG_ERR_OVERFLOW, "code-block %d needs more than %d bytes", i, e->slot_bytes);
        sum += r[i].len;
    }
    if (total) *total = sum;
    if (desc) memcpy(desc, e->cur->h.data(), sizeof(rg_cb_desc) * e->ncb);
    if (res) memcpy(res, r.data(), sizeof(rg_cb_result) * e->ncb);
    if (data) {
        RG_CHECK(cap >= sum, "data buffer too small (%ld < %ld)", cap, sum);
        std::vector<uint8_t> all((size_t)e->slot_bytes * e->ncb);
        RG_HIP(hipMemcpyAsync(all.data(), e->slots.p, all.size(), hipMemcpyDeviceToHost, e->ctx->stream));
        RG_HIP(hipStreamSynchronize(e->ctx->stream));
        long o = 0;
        for (int i = 0; i < e->ncb; ++i) {
            if (offsets) offsets[i] = o;
            memcpy(data + o, all.data() + (size_t)i * e->slot_bytes, r[i].len);
            o += r[i].len;
        }
    } else if (offsets) {
        long o = 0;
        for (int i = 0; i < e->ncb; ++i) { offsets[i] = o; o += r[i].len; }
    }
    return RG_SUCCESS;
}

extern "C" int rg_encoder_get_pass_lengths(rg_encoder *e, int32_t *seglen)
{
    RG_CHECK(e && seglen, "null argument");
    if (!e->t1_ran) return rg_fail(RG_ERR_NOT_READY, "tier-1 has not run");
    RG_CHECK(e->style & RG_STYLE_TERMALL, "per-pass segments exist only with the TERMALL style");
    static_assert(RG_MAX_PASSES == 96, "pass-array stride");
    RG_CHECK(rg_t1_maxp() == RG_MAX_PASSES, "pass-array stride mismatch");
    RG_HIP(hipMemcpyAsync(seglen, e->seglens.p, sizeof(int32_t) * RG_MAX_PASSES * e->ncb, hipMemcpyDeviceToHost,
                          e->ctx->stream));
    RG_HIP(hipStreamSynchronize(e->ctx->stream));
    return RG_SUCCESS;
}

extern "C" int rg_encoder_write_j2k(rg_encoder *e, uint8_t *out, long cap, long *size)
{
    RG_CHECK(e, "null encoder");
    if (!e->t1_ran) return rg_fail(RG_ERR_NOT_READY, "tier-1 has not run");
    long total = 0;
    int rc = rg_encoder_get_codeblocks(e, nullptr, nullptr, nullptr, nullptr, 0, &total);
    if (rc) return rc;
    std::vector<rg_cb_result> res(e->ncb);
    std::vector<long> offs(e->ncb);
    std::vector<uint8_t> data(total > 0 ? total : 1);
    if ((rc = rg_encoder_get_codeblocks(e, nullptr, res.data(), offs.data(), data.data(), (long)data.size(), nullptr)))
        return rc;
    std::vector<const uint8_t *> ptr(e->ncb);
    for (int i = 0; i < e->ncb; ++i) ptr[i] = data.data() + offs[i];
    std::vector<int32_t> seglen;
    if (e->style & RG_STYLE_TERMALL) {
        seglen.resize((size_t)RG_MAX_PASSES * e->ncb);
        if ((rc = rg_encoder_get_pass_lengths(e, seglen.data()))) return rc;
    }
    rg_t2_image im{e->w, e->h, e->ncomp, e->levels, e->cbw, e->cbh, e->filter == RG_FILTER_53,
                   e->ncomp == 3 && e->colour != RG_COLOUR_NONE, 2, e->bitdepth, e->delta, e->style};
    std::vector<uint8_t> cs;
    if ((rc = rg_j2k_assemble(im, e->cur->h.data(), res.data(), ptr.data(), seglen.empty() ? nullptr : seglen.data(),
                              RG_MAX_PASSES, e->ncb_comp, cs, e->nlayers,
                              e->nlayers > 1 ? e->layer_np.data() : nullptr)))
        return rc;
    if (size) *size = (long)cs.size();
    if (out) {
        RG_CHECK(cap >= (long)cs.size(), "output buffer too small (%ld < %zu)", cap, cs.size());
        memcpy(out, cs.data(), cs.size());
    }
    return RG_SUCCESS;
}

extern "C" int rg_encoder_map_dwt_out(rg_encoder *e, void **host)
{
    RG_CHECK(e && host, "null argument");
    if (!e->ran) return rg_fail(RG_ERR_NOT_READY, "encoder has not run");
    const size_t npix = (size_t)e->w * e->h, nc = e->ncomp;
    std::vector<uint8_t> tmp(npix * nc * 4);
    const void *src = e->output_dwt ? e->fcoeff.p : e->q.p;
    RG_HIP(hipMemcpyAsync(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost, e->ctx->stream));
    RG_HIP(hipStreamSynchronize(e->ctx->stream));
    int rc;
    if (e->output_dwt) {
        if ((rc = e->map.ensure(npix * nc * 4))) return rc;
        const float *f = (const float *)tmp.data();
        float *o = (float *)e->map.p;
        for (size_t i = 0; i < npix; ++i)
            for (size_t c = 0; c < nc; ++c) o[i * nc + c] = f[c * npix + i];
    } else {
        if ((rc = e->map.ensure(npix * nc * 2))) return rc;
        const int32_t *q = (const int32_t *)tmp.data();
        int16_t *o = (int16_t *)e->map.p;
        for (size_t i = 0; i < npix; ++i)
            for (size_t c = 0; c < nc; ++c) {
                int32_t v = q[c * npix + i];
                o[i * nc + c] = (int16_t)std::max(-32768, std::min(32767, v));
            }
    }
    *host = e->map.p;
    return RG_SUCCESS;
}

extern "C" int rg_encoder_unmap_dwt_out(rg_encoder *e, void *host)
{
    RG_CHECK(e, "null encoder");
    RG_CHECK(host == e->map.p, "pointer was not mapped by this encoder");
    return RG_SUCCESS;
}

// ------------------------------------------------------------ decoder
struct rg_decoder {
    rg_context *ctx = nullptr;
    int w = 0, h = 0, ncomp = 0, lossy = 0, ran = 0, loaded = 0, u8_ran = 0;
    DevBuf coeff, out, scratch, cbdata, cboffs, cbres, cbdesc, u8, cbseg, t1ws;
    HostBuf staging, map;
    // a code-block decode ends in the fused last inverse level, which writes
    // the uint8 image and no FP32 planes (`out`); those are computed from the
    // coefficients (still in `coeff`) if rg_decoder_map_out asks for them
    int out_valid = 0, levels = 0, filter = 0;
    float step[97];
};

int rg_launch_t1_decode(hipStream_t s, const uint8_t *data, const long *offs, const rg_cb_result *res,
                        const rg_cb_desc *desc, int ncb, int32_t *q, int pitch, size_t plane_stride, int style,
                        const int32_t *seglens, void *ws);
size_t rg_t1_decode_ws_bytes(int ncb);

extern "C" int rg_t1_decode(rg_context *c, const uint8_t *d_data, const long *d_offsets, const rg_cb_result *d_res,
                            const rg_cb_desc *d_desc, int ncb, int32_t *d_q, int pitch, size_t plane_stride)
{
    RG_CHECK(c && d_data && d_offsets && d_res && d_desc && d_q, "rg_t1_decode: null argument");
    RG_CHECK(ncb >= 0 && pitch > 0, "rg_t1_decode: bad sizes");
    return rg_launch_t1_decode(c->stream, d_data, d_offsets, d_res, d_desc, ncb, d_q, pitch, plane_stride, 0, nullptr,
                               nullptr);
}

extern "C" int rg_t1_decode_style(rg_context *c, const uint8_t *d_data, const long *d_offsets,
                                  const rg_cb_result *d_res, const rg_cb_desc *d_desc, int ncb, int32_t *d_q, int pitch,
                                  size_t plane_stride, int style, const int32_t *d_seglen)
{
    RG_CHECK(c && d_data && d_offsets && d_res && d_desc && d_q, "rg_t1_decode_style: null argument");
    RG_CHECK(ncb >= 0 && pitch > 0, "rg_t1_decode_style: bad sizes");
    RG_CHECK(style == 0 || style == RG_STYLE_PARALLEL || style == RG_STYLE_PARALLEL_BYPASS,
             "code-block style %d: only 0, RG_STYLE_PARALLEL and RG_STYLE_PARALLEL_BYPASS", style);
    RG_CHECK(!(style & RG_STYLE_TERMALL) || d_seglen, "TERMALL style needs the per-pass segment lengths");
    return rg_launch_t1_decode(c->stream, d_data, d_offsets, d_res, d_desc, ncb, d_q, pitch, plane_stride, style,
                               (style & RG_STYLE_TERMALL) ? d_seglen : nullptr, nullptr);
}

static int decode_codeblocks(rg_decoder *d, const rg_cb_desc *desc, const rg_cb_result *res, const long *offsets,
                             const uint8_t *data, long total, int ncb, int w, int h, int ncomp,
                             const rg_encode_params *p, const int32_t *seglen, const float *steps);

extern "C" int rg_decoder_decode_codeblocks(rg_decoder *d, const rg_cb_desc *desc, const rg_cb_result *res,
                                            const long *offsets, const uint8_t *data, long total, int ncb, int w,
                                            int h, int ncomp, const rg_encode_params *p, const int32_t *seglen)
{
    return decode_codeblocks(d, desc, res, offsets, data, total, ncb, w, h, ncomp, p, seglen, nullptr);
}

extern "C" int rg_decoder_decode_j2k(rg_decoder *d, const uint8_t *cs, long size, rg_j2k_info *info_out)
{
    RG_CHECK(d && cs, "rg_decoder_decode_j2k: null argument");
    rg_j2k_info info;
    std::vector<rg_cb_desc> desc;
    std::vector<rg_cb_result> res;
    std::vector<long> offs;
    std::vector<uint8_t> data;
    std::vector<int32_t> seg;
    if (int rc = rg_j2k_read(cs, size, info, desc, res, offs, data, seg)) return rc;
    rg_encode_params p{};
    p.levels = info.levels; p.filter = info.filter; p.colour = info.colour; p.cbw = info.cbw; p.cbh = info.cbh;
    p.delta = 1.0 / 64; p.ll_fp16 = 0; p.bitdepth = info.bitdepth; p.style = info.style;
    if (info_out) *info_out = info;
    return decode_codeblocks(d, desc.data(), res.data(), offs.data(), data.empty() ? nullptr : data.data(),
                             (long)data.size(), info.ncb, info.w, info.h, info.ncomp, &p,
                             seg.empty() ? nullptr : seg.data(), info.step);
}

static int decode_codeblocks(rg_decoder *d, const rg_cb_desc *desc, const rg_cb_result *res, const long *offsets,
                             const uint8_t *data, long total, int ncb, int w, int h, int ncomp,
                             const rg_encode_params *p, const int32_t *seglen, const float *steps)
{
    RG_CHECK(d && desc && res && offsets && (data || total == 0), "rg_decoder_decode_codeblocks: null argument");
    int rc = check_params(p, w, h, ncomp);
    if (rc) return rc;
    const bool termall = (p->style & RG_STYLE_TERMALL) != 0;
    RG_CHECK(!termall || seglen, "TERMALL style needs the per-pass segment lengths");
    const int per = rg_codeblocks(w, h, p->levels, p->cbw, p->cbh, nullptr, 0);
    RG_CHECK(per * ncomp == ncb, "expected %d code-blocks, got %d", per * ncomp, ncb);
    for (int i = 0; i < ncb; ++i) {
        RG_CHECK(desc[i].comp >= 0 && desc[i].comp < ncomp && desc[i].w <= p->cbw && desc[i].h <= p->cbh &&
                     desc[i].x0 >= 0 && desc[i].y0 >= 0 && desc[i].x0 + desc[i].w <= w && desc[i].y0 + desc[i].h <= h,
                 "code-block %d lies outside the image", i);
        RG_CHECK(res[i].len >= 0 && res[i].len <= 12288 && offsets[i] >= 0 && offsets[i] + res[i].len <= total,
                 "code-block %d: bad codeword extent", i);
        RG_CHECK(res[i].nbp >= 0 && res[i].nbp <= 31 && res[i].npasses >= 0 && res[i].npasses <= 3 * res[i].nbp,
                 "code-block %d: bad pass counts", i);
        if (termall) {
            long sum = 0;
            for (int k = 0; k < res[i].npasses; ++k) {
                RG_CHECK(seglen[(size_t)i * RG_MAX_PASSES + k] >= 0, "code-block %d: negative segment length", i);
                sum += seglen[(size_t)i * RG_MAX_PASSES + k];
            }
            RG_CHECK(sum == res[i].len, "code-block %d: segment lengths do not add up to the codeword", i);
        }
    }
    RG_HIP(hipSetDevice(d->ctx->device));
    const size_t npix = (size_t)w * h;
    if ((rc = d->cbdata.ensure(total + 16)) || (rc = d->cboffs.ensure(sizeof(long) * ncb)) ||
        (rc = d->cbres.ensure(sizeof(rg_cb_result) * ncb)) || (rc = d->cbdesc.ensure(sizeof(rg_cb_desc) * ncb)) ||
        (rc = d->coeff.ensure(npix * ncomp * 4)) || (rc = d->out.ensure(npix * ncomp * 4)) ||
        (rc = d->scratch.ensure(rg_dwt_scratch_bytes(w, h, ncomp, p->filter))) || (rc = d->u8.ensure(npix * ncomp)) ||
        (termall && (rc = d->cbseg.ensure(sizeof(int32_t) * RG_MAX_PASSES * ncb))) ||
        (rc = d->t1ws.ensure(rg_t1_decode_ws_bytes(ncb))))
        return rc;
    hipStream_t s = d->ctx->stream;
    // stage the host arrays in pinned memory so the uploads are truly
    // asynchronous; the previous call's copies must have drained first
    RG_HIP(hipStreamSynchronize(s));
    const size_t o1 = (total + 15) / 16 * 16, o2 = o1 + sizeof(long) * ncb, o3 = o2 + sizeof(rg_cb_result) * ncb;
    const size_t o4 = o3 + sizeof(rg_cb_desc) * ncb, segb = termall ? sizeof(int32_t) * RG_MAX_PASSES * ncb : 0;
    if ((rc = d->staging.ensure(o4 + segb))) return rc;
    char *st = (char *)d->staging.p;
    if (total) memcpy(st, data, total);
    memcpy(st + o1, offsets, sizeof(long) * ncb);
    memcpy(st + o2, res, sizeof(rg_cb_result) * ncb);
    memcpy(st + o3, desc, sizeof(rg_cb_desc) * ncb);
    if (termall) {
        memcpy(st + o4, seglen, segb);
        RG_HIP(hipMemcpyAsync(d->cbseg.p, st + o4, segb, hipMemcpyHostToDevice, s));
    }
    RG_HIP(hipMemcpyAsync(d->cbdata.p, st, total, hipMemcpyHostToDevice, s));
    RG_HIP(hipMemcpyAsync(d->cboffs.p, st + o1, sizeof(long) * ncb, hipMemcpyHostToDevice, s));
    RG_HIP(hipMemcpyAsync(d->cbres.p, st + o2, sizeof(rg_cb_result) * ncb, hipMemcpyHostToDevice, s));
    RG_HIP(hipMemcpyAsync(d->cbdesc.p, st + o3, sizeof(rg_cb_desc) * ncb, hipMemcpyHostToDevice, s));
    if ((rc = rg_launch_t1_decode(s, (const uint8_t *)d->cbdata.p, (const long *)d->cboffs.p,
                                  (const rg_cb_result *)d->cbres.p, (const rg_cb_desc *)d->cbdesc.p, ncb,
                                  (int32_t *)d->coeff.p, w, npix, p->style, termall ? (const int32_t *)d->cbseg.p : nullptr,
                                  d->t1ws.p)))
        return rc;
    float step[97], inv[97];
    band_steps(p->levels, p->bitdepth, p->delta, step, inv);
    if (steps) memcpy(step, steps, sizeof(float) * (3 * p->levels + 1));
    // inverse transform and inverse colour stage; the last level writes the
    // uint8 image itself where the shape allows (RG_DWT_FUSED=0: never)
    bool fused = false;
    if (p->levels >= 1 && rg_env_int("RG_DWT_FUSED", 1))
        rc = dwt_inv_u8_impl(d->ctx, (const int32_t *)d->coeff.p, w, h, ncomp, p->colour, p->filter, p->levels,
                             step, (uint8_t *)d->u8.p, d->scratch.p, d->out.p, &fused);
    else if (!(rc = dwt_inv_impl(d->ctx, p->filter, d->coeff.p, w, h, ncomp, p->levels,
                                 p->filter == RG_FILTER_97 ? step : nullptr, d->out.p, d->scratch.p)))
        rc = rg_launch_colour_inv(s, d->out.p, w, h, ncomp, p->colour, p->filter == RG_FILTER_97, (uint8_t *)d->u8.p);
    if (rc) return rc;
    d->w = w; d->h = h; d->ncomp = ncomp; d->lossy = p->filter == RG_FILTER_97; d->ran = 1; d->u8_ran = 1;
    d->out_valid = !fused; d->levels = p->levels; d->filter = p->filter;
    memcpy(d->step, step, sizeof(float) * (3 * p->levels + 1));
    return RG_SUCCESS;
}

extern "C" int rg_decoder_image_u8(rg_decoder *d, uint8_t **d_img)
{
    RG_CHECK(d && d_img, "null argument");
    if (!d->u8_ran) return rg_fail(RG_ERR_NOT_READY, "no decoded image");
    *d_img = (uint8_t *)d->u8.p;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_coefficients(rg_decoder *d, int32_t **d_q)
{
    RG_CHECK(d && d_q, "null argument");
    if (!d->u8_ran) return rg_fail(RG_ERR_NOT_READY, "no decoded code-blocks");
    *d_q = (int32_t *)d->coeff.p;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_create(rg_context *c, rg_decoder **out)
{
    RG_CHECK(c && out, "null argument");
    rg_decoder *d = new rg_decoder();
    d->ctx = c;
    *out = d;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_destroy(rg_decoder *d)
{
    if (!d) return RG_SUCCESS;
    hipSetDevice(d->ctx->device);
    hipStreamSynchronize(d->ctx->stream);
    for (DevBuf *b : {&d->coeff, &d->out, &d->scratch, &d->cbdata, &d->cboffs, &d->cbres, &d->cbdesc, &d->u8, &d->cbseg,
                      &d->t1ws})
        b->release();
    d->staging.release(); d->map.release();
    delete d;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_load_planar(rg_decoder *d, const void *const *comps, int ncomp, int type, int w, int h,
                                      int precision, int lossy)
{
    (void)precision;
    RG_CHECK(d && comps, "null argument");
    RG_CHECK(type == RG_F32 || type == RG_I16 || type == RG_I32, "bad sample type %d", type);
    RG_CHECK(w > 0 && h > 0 && ncomp >= 1 && ncomp <= 4, "bad image shape");
    for (int c = 0; c < ncomp; ++c) RG_CHECK(comps[c], "component %d is null", c);
    RG_HIP(hipSetDevice(d->ctx->device));
    const size_t npix = (size_t)w * h;
    int rc = stage_components(d->staging, d->ctx->stream, comps, ncomp, type, npix, lossy);
    if (rc) return rc;
    if ((rc = d->coeff.ensure(npix * ncomp * 4))) return rc;
    RG_HIP(hipMemcpyAsync(d->coeff.p, d->staging.p, npix * ncomp * 4, hipMemcpyHostToDevice, d->ctx->stream));
    d->w = w; d->h = h; d->ncomp = ncomp; d->lossy = lossy; d->loaded = 1; d->ran = 0;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_idwt(rg_decoder *d, int levels)
{
    RG_CHECK(d, "null decoder");
    if (!d->loaded) return rg_fail(RG_ERR_NOT_READY, "no coefficients loaded");
    RG_CHECK(levels >= 0 && levels <= 15, "levels out of range");
    const size_t npix = (size_t)d->w * d->h;
    int rc;
    if ((rc = d->out.ensure(npix * d->ncomp * 4))) return rc;
    if ((rc = d->scratch.ensure(rg_dwt_scratch_bytes(d->w, d->h, d->ncomp, d->lossy)))) return rc;
    rc = rg_dwt_inv(d->ctx, d->lossy ? RG_FILTER_97 : RG_FILTER_53, d->coeff.p, d->w, d->h, d->ncomp, levels, nullptr,
                    d->out.p, d->scratch.p);
    if (rc) return rc;
    d->ran = 1;
    d->out_valid = 1;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_run_planar(rg_decoder *d, const void *const *comps, int ncomp, int type, int w, int h,
                                     int levels, int precision, int lossy)
{
    int rc = rg_decoder_load_planar(d, comps, ncomp, type, w, h, precision, lossy);
    if (rc) return rc;
    return rg_decoder_idwt(d, levels);
}

extern "C" int rg_decoder_map_out(rg_decoder *d, void **host)
{
    RG_CHECK(d && host, "null argument");
    if (!d->ran) return rg_fail(RG_ERR_NOT_READY, "decoder has not run");
    const size_t npix = (size_t)d->w * d->h, nc = d->ncomp;
    if (!d->out_valid) {   // after a fused code-block decode: the planes from the coefficients
        const int rc = dwt_inv_impl(d->ctx, d->filter, d->coeff.p, d->w, d->h, d->ncomp, d->levels,
                                    d->filter == RG_FILTER_97 ? d->step : nullptr, d->out.p, d->scratch.p);
        if (rc) return rc;
        d->out_valid = 1;
    }
    std::vector<uint8_t> tmp(npix * nc * 4);
    RG_HIP(hipMemcpyAsync(tmp.data(), d->out.p, tmp.size(), hipMemcpyDeviceToHost, d->ctx->stream));
    RG_HIP(hipStreamSynchronize(d->ctx->stream));
    int rc;
    if (d->lossy) {
        if ((rc = d->map.ensure(npix * nc * 4))) return rc;
        const float *f = (const float *)tmp.data();
        float *o = (float *)d->map.p;
        for (size_t i = 0; i < npix; ++i)
            for (size_t c = 0; c < nc; ++c) o[i * nc + c] = f[c * npix + i];
    } else {
        if ((rc = d->map.ensure(npix * nc * 2))) return rc;
        const int32_t *q = (const int32_t *)tmp.data();
        int16_t *o = (int16_t *)d->map.p;
        for (size_t i = 0; i < npix; ++i)
            for (size_t c = 0; c < nc; ++c)
                o[i * nc + c] = (int16_t)std::max(-32768, std::min(32767, q[c * npix + i]));
    }
    *host = d->map.p;
    return RG_SUCCESS;
}

extern "C" int rg_decoder_unmap_out(rg_decoder *d, void *host)
{
    RG_CHECK(d, "null decoder");
    RG_CHECK(host == d->map.p, "pointer was not mapped by this decoder");
    return RG_SUCCESS;
}
