// Batched encode: many images of uneven sizes, each stage one launch for the
// whole batch, the pipeline captured once in a HIP graph and replayed.
//
// The reference encodes one image per call (src/OCLEncoder.cpp:43-52) with
// one launch per level and per channel; here the plan step lays out every
// image's buffers in a few large HBM allocations and builds job tables:
//   colour : one workgroup per {image, 4096-pixel chunk};
//   DWT    : per level, one workgroup per {image, plane, 128x16 tile}, both
//            filters in the same launch;
//   tier-1 : code-blocks sorted largest first, either one workgroup per
//            block or a persistent grid (waves_per_cu per CU) pulling block
//            indices from a device counter; the split styles add the
//            lane-parallel MQ launch (and, for the parallel style, packing).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>
#include "rg_internal.h"
#include "rg_t2.h"

namespace {
struct Buf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (p && bytes <= cap) return RG_SUCCESS;
        if (p) { hipFree(p); p = nullptr; cap = 0; }
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            return rg_fail(RG_ERR_OOM, "hipMalloc(%zu) failed", bytes);
        }
        cap = bytes;
        return RG_SUCCESS;
    }
    void release() { if (p) hipFree(p); p = nullptr; cap = 0; }
};

void steps_for(int levels, int bitdepth, double delta, float *step, float *inv)
{
    for (int b = 0; b < 3 * levels + 1; ++b) {
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

struct rg_batch {
    rg_context *ctx = nullptr;
    std::vector<rg_batch_image> img;
    int cbw = 64, cbh = 64, maxlev = 0, slot = 0, planned = 0, use_graph = 0, ran = 0, style = 0, nwaves = 0;
    std::vector<size_t> plane_off, scratch_off;     // element / byte offsets per image
    std::vector<int> cb_first, cb_count;            // per image, into the block list
    std::vector<rg_cb_desc> desc;                    // all blocks, image-major
    std::vector<int> desc_img;
    Buf planes, q, scratch, slots, res, t1jobs, cparams, cjobs, syms, cm, poffs, arena, seglens, counter, chains;
    Buf fargs[16], fjobs[16];
    int nfjobs[16] = {0}, ncjobs = 0, ncb = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<rg_cb_result> h_res;
};

extern "C" int rg_batch_create(rg_context *ctx, rg_batch **out)
{
    RG_CHECK(ctx && out, "rg_batch_create: null argument");
    rg_batch *b = new rg_batch();
    b->ctx = ctx;
    *out = b;
    return RG_SUCCESS;
}

static void drop_graph(rg_batch *b)
{
    if (b->exec) { hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
}

extern "C" int rg_batch_destroy(rg_batch *b)
{
    if (!b) return RG_SUCCESS;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (Buf *x : {&b->planes, &b->q, &b->scratch, &b->slots, &b->res, &b->t1jobs, &b->cparams, &b->cjobs,
                   &b->syms, &b->cm, &b->poffs, &b->arena, &b->seglens, &b->counter, &b->chains})
        x->release();
    for (int l = 0; l < 16; ++l) { b->fargs[l].release(); b->fjobs[l].release(); }
    delete b;
    return RG_SUCCESS;
}

// Issues the whole pipeline on the context stream (also used under capture).
static int issue(rg_batch *b)
{
    hipStream_t s = b->ctx->stream;
    int rc = rg_launch_colour_fwd_batch(s, (const rg_colour_job_params *)b->cparams.p, (const int2 *)b->cjobs.p,
                                        b->ncjobs);
    if (rc) return rc;
    for (int l = 1; l <= b->maxlev; ++l)
        if ((rc = rg_launch_fwd_level_batch(s, (const rg_fwd_level_args *)b->fargs[l].p, (const int4 *)b->fjobs[l].p,
                                            b->nfjobs[l])))
            return rc;
    if (b->nwaves > 0)
        return rg_launch_t1_queue(s, b->t1jobs.p, b->ncb, b->style, (int *)b->counter.p, b->nwaves,
                                  (int32_t *)b->chains.p);
    if (b->style & RG_STYLE_TERMALL)
        return rg_launch_t1_jobs_parallel(s, b->t1jobs.p, b->ncb, b->style, (int32_t *)b->chains.p);
    if (rg_t1_inline()) return rg_launch_t1_jobs(s, b->t1jobs.p, b->ncb);
    return rg_launch_t1_jobs_split(s, b->t1jobs.p, b->ncb);
}

extern "C" int rg_batch_plan(rg_batch *b, const rg_batch_image *imgs, int n, int cbw, int cbh, int use_graph,
                             int style, int waves_per_cu)
{
    RG_CHECK(b && imgs && n > 0, "rg_batch_plan: bad arguments");
    RG_CHECK(waves_per_cu >= 0 && waves_per_cu <= 32, "waves_per_cu %d out of range 0..32", waves_per_cu);
    RG_CHECK(style == 0 || style == RG_STYLE_PARALLEL || style == RG_STYLE_PARALLEL_BYPASS,
             "code-block style %d: only 0, RG_STYLE_PARALLEL and RG_STYLE_PARALLEL_BYPASS", style);
    RG_HIP(hipSetDevice(b->ctx->device));
    RG_HIP(hipStreamSynchronize(b->ctx->stream));
    drop_graph(b);
    b->planned = 0; b->ran = 0;
    if (rg_codeblocks(64, 64, 1, cbw, cbh, nullptr, 0) < 0) return RG_ERR_INVALID;
    b->img.assign(imgs, imgs + n);
    b->cbw = cbw; b->cbh = cbh; b->maxlev = 0;
    for (int i = 0; i < n; ++i) {
        const rg_batch_image &m = imgs[i];
        RG_CHECK(m.d_src && m.w > 0 && m.h > 0 && (m.ncomp == 1 || m.ncomp == 3), "image %d: bad shape", i);
        // every image needs a DWT level: the batch has no level-0 copy/quantise job,
        // so a 0-level image's coefficient plane would never be written
        RG_CHECK(m.levels >= 1 && m.levels <= 15, "image %d: levels %d out of range 1..15", i, m.levels);
        RG_CHECK(m.filter == RG_FILTER_53 || m.filter == RG_FILTER_97, "image %d: bad filter", i);
        RG_CHECK(m.colour == RG_COLOUR_NONE || m.ncomp == 3, "image %d: colour transform needs 3 components", i);
        RG_CHECK(m.filter == RG_FILTER_53 || m.delta > 0, "image %d: bad quantisation step", i);
        RG_CHECK(((uintptr_t)m.d_src & 15) == 0, "image %d: source must be 16-byte aligned", i);
        b->maxlev = std::max(b->maxlev, m.levels);
    }
    // ---- buffer layout (element offsets, 64-element aligned) ----
    b->plane_off.resize(n + 1);
    b->scratch_off.resize(n + 1);
    size_t pe = 0, sb = 0;
    for (int i = 0; i < n; ++i) {
        b->plane_off[i] = pe;
        pe += ((size_t)imgs[i].w * imgs[i].h * imgs[i].ncomp + 63) / 64 * 64;
        b->scratch_off[i] = sb;
        sb += (rg_dwt_scratch_bytes(imgs[i].w, imgs[i].h, imgs[i].ncomp, imgs[i].filter) + 255) / 256 * 256;
    }
    b->plane_off[n] = pe; b->scratch_off[n] = sb;
    int rc;
    if ((rc = b->planes.ensure(pe * 4)) || (rc = b->q.ensure(pe * 4)) || (rc = b->scratch.ensure(sb))) return rc;
    // ---- colour jobs ----
    std::vector<rg_colour_job_params> cp(n);
    std::vector<int2> cj;
    for (int i = 0; i < n; ++i) {
        const rg_batch_image &m = imgs[i];
        cp[i].src = m.d_src; cp[i].npix = (long)m.w * m.h; cp[i].ncomp = m.ncomp;
        cp[i].colour = m.colour; cp[i].out_float = m.filter == RG_FILTER_97;
        cp[i].planes = (char *)b->planes.p + b->plane_off[i] * 4;
        long chunks = (cp[i].npix + 15) / 16;
        for (long c = 0; c < chunks; c += 256) cj.push_back(make_int2(i, (int)c));
    }
    b->ncjobs = (int)cj.size();
    if ((rc = b->cparams.ensure(sizeof(rg_colour_job_params) * n)) || (rc = b->cjobs.ensure(sizeof(int2) * cj.size())))
        return rc;
    RG_HIP(hipMemcpy(b->cparams.p, cp.data(), sizeof(rg_colour_job_params) * n, hipMemcpyHostToDevice));
    RG_HIP(hipMemcpy(b->cjobs.p, cj.data(), sizeof(int2) * cj.size(), hipMemcpyHostToDevice));
    // ---- DWT level jobs ----
    for (int l = 1; l <= b->maxlev; ++l) {
        std::vector<rg_fwd_level_args> args;
        std::vector<int4> jobs;
        for (int i = 0; i < n; ++i) {
            const rg_batch_image &m = imgs[i];
            if (l > m.levels) continue;
            rg_level_dims d = rg_dims(m.w, m.h, m.levels);
            float step[97], inv[97];
            steps_for(m.levels, 8, m.filter == RG_FILTER_97 ? m.delta : 1.0, step, inv);
            const size_t np = (size_t)m.w * m.h;
            char *sc = (char *)b->scratch.p + b->scratch_off[i];
            rg_fwd_level_args a{};
            a.filter = m.filter;
            a.cw = d.w[l - 1]; a.ch = d.h[l - 1]; a.nplanes = m.ncomp;
            if (l == 1) { a.in = (char *)b->planes.p + b->plane_off[i] * 4; a.in_pitch = m.w; a.in_ps = (long)np; }
            else {
                a.in = sc + rg_ll_offset(m.w, m.h, m.ncomp, l - 1);
                a.in_pitch = d.w[l - 1]; a.in_ps = (long)d.w[l - 1] * d.h[l - 1];
            }
            if (l < m.levels) {
                a.ll_out = sc + rg_ll_offset(m.w, m.h, m.ncomp, l);
                a.ll_pitch = d.w[l]; a.ll_ps = (long)d.w[l] * d.h[l];
            }
            a.q = (char *)b->q.p + b->plane_off[i] * 4; a.qp = m.w; a.q_ps = (long)np;
            if (m.filter == RG_FILTER_97) {
                const int bi = 1 + 3 * (m.levels - l);
                a.inv_ll = inv[0]; a.inv_hl = inv[bi]; a.inv_lh = inv[bi + 1]; a.inv_hh = inv[bi + 2];
            }
            const int ai = (int)args.size();
            args.push_back(a);
            const int tx = (a.cw + RG_DWT_FWD_TW - 1) / RG_DWT_FWD_TW, ty = (a.ch + RG_DWT_FWD_TH - 1) / RG_DWT_FWD_TH;
            for (int p = 0; p < m.ncomp; ++p)
                for (int y = 0; y < ty; ++y)
                    for (int x = 0; x < tx; ++x) jobs.push_back(make_int4(ai, p, x, y));
        }
        b->nfjobs[l] = (int)jobs.size();
        if ((rc = b->fargs[l].ensure(sizeof(rg_fwd_level_args) * std::max<size_t>(1, args.size()))) ||
            (rc = b->fjobs[l].ensure(sizeof(int4) * std::max<size_t>(1, jobs.size()))))
            return rc;
        if (!args.empty())
            RG_HIP(hipMemcpy(b->fargs[l].p, args.data(), sizeof(rg_fwd_level_args) * args.size(), hipMemcpyHostToDevice));
        if (!jobs.empty())
            RG_HIP(hipMemcpy(b->fjobs[l].p, jobs.data(), sizeof(int4) * jobs.size(), hipMemcpyHostToDevice));
    }
    // ---- code-blocks and tier-1 jobs (largest first) ----
    b->desc.clear(); b->desc_img.clear();
    b->cb_first.assign(n, 0); b->cb_count.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        const rg_batch_image &m = imgs[i];
        int k = rg_codeblocks(m.w, m.h, m.levels, cbw, cbh, nullptr, 0);
        std::vector<rg_cb_desc> one(k);
        rg_codeblocks(m.w, m.h, m.levels, cbw, cbh, one.data(), k);
        b->cb_first[i] = (int)b->desc.size();
        b->cb_count[i] = k * m.ncomp;
        for (int c = 0; c < m.ncomp; ++c)
            for (int j = 0; j < k; ++j) {
                rg_cb_desc dd = one[j];
                dd.comp = c;
                b->desc.push_back(dd);
                b->desc_img.push_back(i);
            }
    }
    b->ncb = (int)b->desc.size();
    b->slot = std::max(256, cbw * cbh * 3);
    const int symcap = rg_t1_symcap(cbw, cbh);
    if ((rc = b->slots.ensure((size_t)b->slot * b->ncb)) || (rc = b->res.ensure(sizeof(rg_cb_result) * b->ncb)) ||
        (rc = b->cm.ensure(sizeof(int4) * b->ncb)) || (rc = b->counter.ensure(256)) ||
        ((style || !rg_t1_inline()) && (rc = b->syms.ensure((size_t)symcap * b->ncb))))
        return rc;
    b->style = style;
    const int maxp = rg_t1_maxp();
    const size_t astride = rg_t1_seg_bytes(symcap);
    if (style && ((rc = b->poffs.ensure(sizeof(int32_t) * maxp * b->ncb)) ||
                  (rc = b->seglens.ensure(sizeof(int32_t) * maxp * b->ncb)) ||
                  (rc = b->arena.ensure(astride * b->ncb)) ||
                  (rc = b->chains.ensure(rg_t1_chain_ws_bytes(b->ncb)))))
        return rc;
    std::vector<int> order(b->ncb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->desc[x].w * b->desc[x].h > b->desc[y].w * b->desc[y].h; });
    const int jb = rg_t1_job_bytes();
    std::vector<char> jobs((size_t)jb * b->ncb);
    for (int k = 0; k < b->ncb; ++k) {
        const int idx = order[k];
        const rg_cb_desc &d = b->desc[idx];
        const rg_batch_image &m = imgs[b->desc_img[idx]];
        const int32_t *base = (const int32_t *)b->q.p + b->plane_off[b->desc_img[idx]] +
                              (size_t)d.comp * m.w * m.h + (size_t)d.y0 * m.w + d.x0;
        rg_t1_job_fill(jobs.data() + (size_t)k * jb, base, m.w, d.w, d.h, d.orient,
                       (uint8_t *)b->slots.p + (size_t)idx * b->slot, b->slot, (rg_cb_result *)b->res.p + idx,
                       b->syms.p ? (uint8_t *)b->syms.p + (size_t)k * symcap : nullptr, symcap, (int4 *)b->cm.p + k,
                       style ? (int32_t *)b->poffs.p + (size_t)k * maxp : nullptr,
                       style ? (uint8_t *)b->arena.p + (size_t)k * astride : nullptr,
                       style ? (int32_t *)b->seglens.p + (size_t)idx * maxp : nullptr, style);
    }
    if ((rc = b->t1jobs.ensure(jobs.size()))) return rc;
    RG_HIP(hipMemcpy(b->t1jobs.p, jobs.data(), jobs.size(), hipMemcpyHostToDevice));
    b->nwaves = 0;
    if (waves_per_cu > 0) {
        hipDeviceProp_t prop;
        RG_HIP(hipGetDeviceProperties(&prop, b->ctx->device));
        b->nwaves = std::max(1, std::min(b->ncb, prop.multiProcessorCount * waves_per_cu));
    }
    b->use_graph = use_graph;
    if (use_graph) {
        RG_HIP(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
        rc = issue(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
        if (rc) { if (g) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return rg_fail(RG_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
        b->graph = g;
        RG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    b->planned = 1;
    return RG_SUCCESS;
}

extern "C" int rg_batch_run(rg_batch *b)
{
    RG_CHECK(b, "null batch");
    if (!b->planned) return rg_fail(RG_ERR_NOT_READY, "batch not planned");
    RG_HIP(hipSetDevice(b->ctx->device));
    int rc = RG_SUCCESS;
    if (b->use_graph) RG_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    else rc = issue(b);
    if (rc == RG_SUCCESS) { b->ran = 1; b->h_res.clear(); }
    return rc;
}

extern "C" int rg_batch_num_codeblocks(rg_batch *b, int image)
{
    RG_CHECK(b && b->planned, "batch not planned");
    RG_CHECK(image >= -1 && image < (int)b->img.size(), "image index out of range");
    return image < 0 ? b->ncb : b->cb_count[image];
}

extern "C" int rg_batch_coefficients(rg_batch *b, int image, int32_t **d_q)
{
    RG_CHECK(b && b->planned && d_q, "bad arguments");
    RG_CHECK(image >= 0 && image < (int)b->img.size(), "image index out of range");
    *d_q = (int32_t *)b->q.p + b->plane_off[image];
    return RG_SUCCESS;
}

extern "C" int rg_batch_get_codeblocks(rg_batch *b, int image, rg_cb_desc *desc, rg_cb_result *res, long *offsets,
                                       uint8_t *data, long cap, long *total)
{
    RG_CHECK(b && b->planned, "batch not planned");
    RG_CHECK(image >= 0 && image < (int)b->img.size(), "image index out of range");
    if (!b->ran) return rg_fail(RG_ERR_NOT_READY, "batch has not run");
    if (b->h_res.empty()) {
        b->h_res.resize(b->ncb);
        RG_HIP(hipMemcpyAsync(b->h_res.data(), b->res.p, sizeof(rg_cb_result) * b->ncb, hipMemcpyDeviceToHost,
                              b->ctx->stream));
        RG_HIP(hipStreamSynchronize(b->ctx->stream));
    }
    const int first = b->cb_first[image], cnt = b->cb_count[image];
    long sum = 0;
    for (int i = 0; i < cnt; ++i) {
        if (b->h_res[first + i].flags & 1)
            return rg_fail(RG_ERR_OVERFLOW, "image %d code-block %d overflowed its slot", image, i);
        sum += b->h_res[first + i].len;
    }
    if (total) *total = sum;
    if (desc) memcpy(desc, b->desc.data() + first, sizeof(rg_cb_desc) * cnt);
    if (res) memcpy(res, b->h_res.data() + first, sizeof(rg_cb_result) * cnt);
    std::vector<uint8_t> all;
    if (data) {
        RG_CHECK(sum <= cap, "data buffer too small (%ld < %ld)", cap, sum);
        all.resize((size_t)b->slot * cnt);
        RG_HIP(hipMemcpyAsync(all.data(), (uint8_t *)b->slots.p + (size_t)first * b->slot, all.size(),
                              hipMemcpyDeviceToHost, b->ctx->stream));
        RG_HIP(hipStreamSynchronize(b->ctx->stream));
    }
    long o = 0;
    for (int i = 0; i < cnt; ++i) {
        if (offsets) offsets[i] = o;
        if (data) memcpy(data + o, all.data() + (size_t)i * b->slot, b->h_res[first + i].len);
        o += b->h_res[first + i].len;
    }
    return RG_SUCCESS;
}

extern "C" int rg_batch_get_pass_lengths(rg_batch *b, int image, int32_t *seglen)
{
    RG_CHECK(b && b->planned && seglen, "bad arguments");
    RG_CHECK(image >= 0 && image < (int)b->img.size(), "image index out of range");
    if (!b->ran) return rg_fail(RG_ERR_NOT_READY, "batch has not run");
    RG_CHECK(b->style & RG_STYLE_TERMALL, "per-pass segments exist only with the TERMALL style");
    RG_CHECK(rg_t1_maxp() == RG_MAX_PASSES, "pass-array stride mismatch");
    const size_t first = (size_t)b->cb_first[image] * RG_MAX_PASSES;
    RG_HIP(hipMemcpyAsync(seglen, (int32_t *)b->seglens.p + first, sizeof(int32_t) * RG_MAX_PASSES * b->cb_count[image],
                          hipMemcpyDeviceToHost, b->ctx->stream));
    RG_HIP(hipStreamSynchronize(b->ctx->stream));
    return RG_SUCCESS;
}

extern "C" int rg_batch_write_j2k(rg_batch *b, int image, uint8_t *out, long cap, long *size)
{
    RG_CHECK(b && b->planned, "batch not planned");
    RG_CHECK(image >= 0 && image < (int)b->img.size(), "image index out of range");
    long total = 0;
    int rc = rg_batch_get_codeblocks(b, image, nullptr, nullptr, nullptr, nullptr, 0, &total);
    if (rc) return rc;
    const int cnt = b->cb_count[image];
    std::vector<rg_cb_result> res(cnt);
    std::vector<long> offs(cnt);
    std::vector<uint8_t> data(total > 0 ? total : 1);
    if ((rc = rg_batch_get_codeblocks(b, image, nullptr, res.data(), offs.data(), data.data(), (long)data.size(),
                                      nullptr)))
        return rc;
    std::vector<const uint8_t *> ptr(cnt);
    for (int i = 0; i < cnt; ++i) ptr[i] = data.data() + offs[i];
    std::vector<int32_t> seglen;
    if (b->style & RG_STYLE_TERMALL) {
        seglen.resize((size_t)RG_MAX_PASSES * cnt);
        if ((rc = rg_batch_get_pass_lengths(b, image, seglen.data()))) return rc;
    }
    const rg_batch_image &m = b->img[image];
    rg_t2_image im{m.w, m.h, m.ncomp, m.levels, b->cbw, b->cbh, m.filter == RG_FILTER_53,
                   m.ncomp == 3 && m.colour != RG_COLOUR_NONE, 2, 8, m.delta, b->style};
    std::vector<uint8_t> cs;
    if ((rc = rg_j2k_assemble(im, b->desc.data() + b->cb_first[image], res.data(), ptr.data(),
                              seglen.empty() ? nullptr : seglen.data(), RG_MAX_PASSES, cnt / m.ncomp, cs)))
        return rc;
    if (size) *size = (long)cs.size();
    if (out) {
        RG_CHECK(cap >= (long)cs.size(), "output buffer too small (%ld < %zu)", cap, cs.size());
        memcpy(out, cs.data(), cs.size());
    }
    return RG_SUCCESS;
}
