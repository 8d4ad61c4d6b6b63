// Runs the tier-1 encoder cores (roger_amd/csrc/t1_core.h) on the CPU under
// the lock-step wavefront emulator of wave_emu.h, so the kernels' own code
// (not a restatement) can be checked against the oracle without a GPU.
//
//   t1_emu MODE DIR PITCH PLANE_STRIDE SLOT [STYLE]
//     MODE  inline | split | queue_split | parallel | queue_parallel | sorted | two_phase | three_phase |
//           cm_only
//     STYLE code-block style bits of the parallel modes (default 6 =
//           RESET|TERMALL; 7 adds BYPASS)
//     DIR   holds q.bin (int32 coefficient planes) and desc.bin (rg_cb_desc
//           records); res.bin (rg_cb_result per block) and out.bin (SLOT
//           bytes per block) are written there.
// The three modes drive the cores the way t1_encode.hip's kernels do:
//   inline       t1_block<false>, one wave per block
//   split        t1_block<true> per block, then t1_mq_lanes<false> per 64 blocks
//   queue_split  the same over a job list (t1_mq_lanes<true>)
//   parallel     RESET|TERMALL style: t1_block<true> recording pass offsets,
//                t1_mq_pass_lane per (64 blocks, pass), t1_pack_block per block
//   queue_parallel  the same reading every pointer from the job list
//   sorted       parallel styles with the chains in the longest-first lists
//                (built here with the kernels' chain_key) coded 64 per
//                wavefront by t1_chain_lane, as chain_code_kernel does
//   two_phase    sorted, with context modelling split at pass 10: passes
//                0-9 modelled and coded, then the rest from the saved state
//   three_phase  the same with phases [0, 7), [7, 10) and [10, ...)
//   cm_only      context modelling alone; writes cm.bin and poff.bin
// The parallel modes also write seglen.bin (MAXP int32 per block).
//
//   t1_emu decode DIR PITCH PLANE_STRIDE NPLANES [STYLE]
// runs the decoder core (t1_dec_core.h) per block, as t1_decode_kernel does:
// DIR holds desc.bin, res.bin, cw.bin (codewords), offs.bin (int64 offset of
// each block's codeword) and, for TERMALL styles, seglen.bin; the decoded
// coefficients are written to dq.bin (NPLANES planes of int32).
// Exit status 3: an emulator check failed (message on stderr).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "wave_emu.h"
#include "../../roger_amd/csrc/t1_core.h"
#include "../../roger_amd/csrc/t1_dec_core.h"
#include "../../roger_amd/csrc/t1_dec_lane.h"

using namespace rg_t1;

// ---- fiber switch (x86-64 System V: callee-saved registers on the stack) ----
extern "C" void emu_swap(void **save_sp, void *to_sp);
asm(R"(
    .text
    .globl emu_swap
    .type emu_swap, @function
emu_swap:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_swap, .-emu_swap
)");

namespace emu {
namespace {
constexpr size_t STACK = 1 << 18;
void *sched_sp;
void *lane_sp[64];
char *stacks[64];
int cur = 0;
bool done[64];
uint64_t buf[2][64];
int phase[64];
void *site[64];
int want_uniform[64];
void (*body)(void *) = nullptr;
void *body_arg = nullptr;

[[noreturn]] void fail(const char *msg)
{
    fprintf(stderr, "wave emulator: %s\n", msg);
    exit(3);
}

void entry()
{
    body(body_arg);
    done[cur] = true;
    emu_swap(&lane_sp[cur], sched_sp);
    fail("finished lane resumed");
}
}  // namespace

int lane() { return cur; }

__attribute__((noinline)) const uint64_t *exchange(uint64_t v, int check_uniform)
{
    const int L = cur, ph = phase[L];
    buf[ph][L] = v;
    phase[L] ^= 1;
    site[L] = __builtin_return_address(0);
    want_uniform[L] = check_uniform;
    emu_swap(&lane_sp[L], sched_sp);
    return buf[ph];
}

// Runs fn(arg) as one 64-lane wavefront to completion.
void run(void (*fn)(void *), void *arg)
{
    body = fn;
    body_arg = arg;
    for (int l = 0; l < 64; ++l) {
        if (!stacks[l]) stacks[l] = static_cast<char *>(aligned_alloc(16, STACK));
        const uintptr_t top = reinterpret_cast<uintptr_t>(stacks[l] + STACK) & ~uintptr_t(15);
        void **sp = reinterpret_cast<void **>(top - 64);
        for (int k = 0; k < 8; ++k) sp[k] = nullptr;
        sp[6] = reinterpret_cast<void *>(&entry);   // popped by `ret` after the six registers
        lane_sp[l] = sp;
        done[l] = false;
        phase[l] = 0;
    }
    for (;;) {
        for (int l = 0; l < 64; ++l) {
            if (done[l]) continue;
            cur = l;
            emu_swap(&sched_sp, lane_sp[l]);
        }
        int ndone = 0;
        for (int l = 0; l < 64; ++l) ndone += done[l];
        if (ndone == 64) return;
        if (ndone) fail("a lane left the kernel while the others wait at a cross-lane call");
        for (int l = 1; l < 64; ++l)
            if (site[l] != site[0]) fail("lanes reached different cross-lane call sites (non-uniform control flow)");
        if (want_uniform[0]) {
            const int ph = phase[0] ^ 1;
            for (int l = 1; l < 64; ++l)
                if (buf[ph][l] != buf[ph][0]) fail("a value declared wave-uniform differs between lanes");
        }
    }
}
}  // namespace emu

namespace {
uint32_t g_tab[MQ_TT_WORDS];
alignas(16) uint8_t g_lut[CM_LDS_BYTES];   // the kernels' workgroup memory: context tables, symbol selectors
alignas(16) uint8_t g_ring[MQ_LDS_BYTES]; // and the lane MQ coders' output rings and context words

void inline_body(void *a) { t1_block<false>(*static_cast<T1Job *>(a), g_lut); }
void cm_body(void *a) { t1_block<true>(*static_cast<T1Job *>(a), g_lut); }
CMPhase g_phase;
void cm_phase_body(void *a) { t1_block<true>(*static_cast<T1Job *>(a), g_lut, g_phase); }

struct MQFlat {
    int i0, ncb, symcap, slot;
    const uint8_t *syms;
    const i32x4 *cm;
    uint8_t *slots;
    rg_cb_result *res;
};
void mq_flat_body(void *a)
{
    const MQFlat &m = *static_cast<MQFlat *>(a);
    t1_mq_lanes<false>(g_tab, g_ring, nullptr, m.i0 + wv_lane(), m.ncb, m.syms, m.symcap, m.cm, m.slots, m.slot,
                       m.res);
}

struct MQQueue {
    const T1Job *jobs;
    int i0, njobs;
};

// parallel style, driven the way t1_mqp_kernel / t1_mqp_jobs_kernel drive it
struct MQPass {
    const T1Job *jobs;
    int i0, njobs, p, style;
};
void mq_pass_body(void *a)
{
    const MQPass &m = *static_cast<MQPass *>(a);
    const int b = m.i0 + wv_lane();
    bool active = b < m.njobs;
    int p = 0, s0 = 0, s1 = 0, at = 0, cap = 0;
    const T1Job &j = m.jobs[active ? b : 0];
    if (active) active = pass_range(*j.cm, j.poff, m.p, p, s0, s1, at, cap);
    if (pass_is_raw(m.style, m.p))
        t1_raw_pass_lane(g_ring, active, j.sym, s0, s1, j.seg + at, cap, j.seglen + p);
    else
        t1_mq_pass_lane(g_tab, g_ring, active, j.sym, s0, s1, j.seg + at, cap, j.seglen + p);
}
struct JobAt {
    const T1Job *jobs;
    const T1Job &operator()(int b) const { return jobs[b]; }
};
struct ChainRun {
    const T1Job *jobs;
    const int32_t *list;
    int base, end;
    bool raw;
};
void chain_body(void *a)
{
    const ChainRun &c = *static_cast<ChainRun *>(a);
    t1_chain_lane(g_tab, g_ring, c.raw, c.list, c.base, c.end, JobAt{c.jobs});
}
void pack_body(void *a)
{
    const T1Job &j = *static_cast<T1Job *>(a);
    static int32_t sh[PACK_WORDS];
    t1_pack_block(*j.cm, j.poff, j.seg, j.seglen, j.out, j.cap, j.res, sh);
}
uint8_t g_dlut[LUT_BYTES];
uint32_t g_rawbuf[DEC_RAW_WORDS];
void dec_body(void *a) { t1_dec_block(*static_cast<DecJob *>(a), g_dlut, g_rawbuf); }
void mq_queue_body(void *a)
{
    const MQQueue &m = *static_cast<MQQueue *>(a);
    t1_mq_lanes<true>(g_tab, g_ring, m.jobs, m.i0 + wv_lane(), m.njobs, nullptr, 0, nullptr, nullptr, 0, nullptr);
}

std::vector<char> slurp(const std::string &path)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path.c_str()); exit(2); }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<char> v((size_t)n);
    if (n && fread(v.data(), 1, (size_t)n, f) != (size_t)n) { fprintf(stderr, "short read %s\n", path.c_str()); exit(2); }
    fclose(f);
    return v;
}

void spill(const std::string &path, const void *p, size_t n)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f || fwrite(p, 1, n, f) != n) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    fclose(f);
}
}  // namespace

int main(int argc, char **argv)
{
    if (argc != 6 && argc != 7) {
        fprintf(stderr, "usage: t1_emu MODE DIR PITCH PLANE_STRIDE SLOT [STYLE]\n");
        return 2;
    }
    const std::string mode = argv[1], dir = argv[2];
    const int pitch = atoi(argv[3]);
    const long plane_stride = atol(argv[4]);
    if (mode == "decode_lane") {
        // the lane decoder (t1_dec_lane.h) per block with plain host state,
        // then the magnitude fold of the epilogue
        const int nplanes = atoi(argv[5]);
        const int dstyle = argc == 7 ? atoi(argv[6]) : 0;
        std::vector<char> draw = slurp(dir + "/desc.bin"), rraw = slurp(dir + "/res.bin");
        std::vector<char> cw = slurp(dir + "/cw.bin"), oraw = slurp(dir + "/offs.bin");
        const int ncb = (int)(draw.size() / sizeof(rg_cb_desc));
        const rg_cb_desc *desc = reinterpret_cast<const rg_cb_desc *>(draw.data());
        const rg_cb_result *res = reinterpret_cast<const rg_cb_result *>(rraw.data());
        const int64_t *offs = reinterpret_cast<const int64_t *>(oraw.data());
        std::vector<char> sraw;
        if (dstyle & STYLE_TERMALL) sraw = slurp(dir + "/seglen.bin");
        std::vector<int32_t> dq((size_t)nplanes * plane_stride, 0);
        struct HostMem {
            uint64_t s[64], n[64], v[64], r[64], b[64];
            int cx[19];
            int orient;
            uint64_t &S(int c) { return s[c]; }
            uint64_t &N(int c) { return n[c]; }
            uint64_t &V(int c) { return v[c]; }
            uint64_t &R(int c) { return r[c]; }
            uint64_t &B(int c) { return b[c]; }
            int ctx(int k) const { return cx[k]; }
            void set_ctx(int k, int x) { cx[k] = x; }
            uint32_t mqt(int k) const { return MQT[k]; }
            uint32_t lut(int x) const { return lut_entry(orient, x); }
        };
        struct HostRd {
            const uint8_t *p;
            int n;
            uint32_t get(int i) const
            {
                if (i < 0 || i >= n) { fprintf(stderr, "lane decoder read byte %d of %d\n", i, n); exit(3); }
                return p[i];
            }
        };
        for (int k = 0; k < ncb; ++k) {
            HostMem mem;
            mem.orient = desc[k].orient;
            std::vector<uint64_t> planes(64 * 32, 0), negs(64, 0);
            LaneJob j{};
            j.len = res[k].len; j.npasses = res[k].npasses; j.nbp = res[k].nbp;
            j.w = desc[k].w; j.h = desc[k].h; j.orient = desc[k].orient;
            j.style = dstyle;
            j.seglen = (dstyle & STYLE_TERMALL) ? reinterpret_cast<const int32_t *>(sraw.data()) + (size_t)k * MAXP
                                                : nullptr;
            j.planes = planes.data();
            j.negs = negs.data();
            HostRd rd{reinterpret_cast<const uint8_t *>(cw.data()) + offs[k], res[k].len};
            t1_dec_lane(j, mem, rd);
            int32_t *dst = dq.data() + (size_t)desc[k].comp * plane_stride + (size_t)desc[k].y0 * pitch + desc[k].x0;
            for (int c = 0; c < desc[k].w; ++c)
                for (int r = 0; r < desc[k].h; ++r) {
                    uint32_t mag = 0;
                    for (int p = j.nbp - 1; p >= 0; --p) mag = (mag << 1) | (uint32_t)((planes[p * 64 + c] >> r) & 1);
                    const int v = (int)mag;
                    dst[(size_t)r * pitch + c] = ((negs[c] >> r) & 1) ? -v : v;
                }
        }
        spill(dir + "/dq.bin", dq.data(), sizeof(int32_t) * dq.size());
        return 0;
    }
    if (mode == "decode") {
        const int nplanes = atoi(argv[5]);
        const int dstyle = argc == 7 ? atoi(argv[6]) : 0;
        std::vector<char> draw = slurp(dir + "/desc.bin"), rraw = slurp(dir + "/res.bin");
        std::vector<char> cw = slurp(dir + "/cw.bin"), oraw = slurp(dir + "/offs.bin");
        const int ncb = (int)(draw.size() / sizeof(rg_cb_desc));
        const rg_cb_desc *desc = reinterpret_cast<const rg_cb_desc *>(draw.data());
        const rg_cb_result *res = reinterpret_cast<const rg_cb_result *>(rraw.data());
        const int64_t *offs = reinterpret_cast<const int64_t *>(oraw.data());
        std::vector<char> sraw;
        if (dstyle & STYLE_TERMALL) sraw = slurp(dir + "/seglen.bin");
        cw.resize(cw.size() + 16);
        std::vector<int32_t> dq((size_t)nplanes * plane_stride, 0);
        for (int k = 0; k < ncb; ++k) {
            DecJob j{};
            j.data = reinterpret_cast<const uint8_t *>(cw.data()) + offs[k];
            j.len = res[k].len; j.npasses = res[k].npasses; j.nbp = res[k].nbp;
            j.w = desc[k].w; j.h = desc[k].h; j.orient = desc[k].orient; j.pitch = pitch;
            j.dst = dq.data() + (size_t)desc[k].comp * plane_stride + (size_t)desc[k].y0 * pitch + desc[k].x0;
            j.style = dstyle;
            j.seglen = (dstyle & STYLE_TERMALL) ? reinterpret_cast<const int32_t *>(sraw.data()) + (size_t)k * MAXP
                                                : nullptr;
            emu::run(dec_body, &j);
        }
        spill(dir + "/dq.bin", dq.data(), sizeof(int32_t) * dq.size());
        return 0;
    }
    const int style = argc == 7 ? atoi(argv[6]) : STYLE_RESET | STYLE_TERMALL;
    const int slot = atoi(argv[5]);
    std::vector<char> qraw = slurp(dir + "/q.bin"), draw = slurp(dir + "/desc.bin");
    const int32_t *q = reinterpret_cast<const int32_t *>(qraw.data());
    const int ncb = (int)(draw.size() / sizeof(rg_cb_desc));
    const rg_cb_desc *desc = reinterpret_cast<const rg_cb_desc *>(draw.data());

    const int symcap = ((64 * 64 * 24) + 15) / 16 * 16;
    std::vector<uint8_t> slots((size_t)slot * ncb, 0);
    std::vector<rg_cb_result> res(ncb);
    uint8_t *syms = static_cast<uint8_t *>(aligned_alloc(16, (size_t)symcap * (ncb ? ncb : 1)));
    std::vector<i32x4> cm(ncb);
    std::vector<u32x4> cms((size_t)64 * (ncb ? ncb : 1));
    std::vector<T1Job> jobs(ncb);
    const size_t astride = (size_t)seg_bytes(symcap);
    uint8_t *arena = static_cast<uint8_t *>(aligned_alloc(16, astride * (ncb ? ncb : 1)));
    std::vector<int32_t> poffs((size_t)MAXP * (ncb ? ncb : 1), 0), seglens((size_t)MAXP * (ncb ? ncb : 1), 0);
    for (int k = 0; k < ncb; ++k) {
        const rg_cb_desc &d = desc[k];
        T1Job &j = jobs[k];
        j.src = q + (size_t)d.comp * plane_stride + (size_t)d.y0 * pitch + d.x0;
        j.pitch = pitch; j.w = d.w; j.h = d.h; j.orient = d.orient; j.cap = slot;
        j.out = slots.data() + (size_t)k * slot;
        j.res = &res[k];
        j.sym = syms + (size_t)k * symcap;
        j.symcap = symcap;
        j.cm = &cm[k];
        const bool par = mode == "parallel" || mode == "queue_parallel" || mode == "sorted" || mode == "two_phase" ||
                         mode == "three_phase";
        j.style = par ? style : 0;
        j.poff = par ? &poffs[(size_t)k * MAXP] : nullptr;
        j.seg = arena + (size_t)k * astride;
        j.seglen = &seglens[(size_t)k * MAXP];
        j.cms = &cms[(size_t)k * 64];
    }
    if (mode == "inline") {
        for (int k = 0; k < ncb; ++k) emu::run(inline_body, &jobs[k]);
    } else if (mode == "split" || mode == "queue_split") {
        for (int k = 0; k < ncb; ++k) emu::run(cm_body, &jobs[k]);
        for (int i0 = 0; i0 < ncb; i0 += 64) {
            if (mode == "split") {
                MQFlat m{i0, ncb, symcap, slot, syms, cm.data(), slots.data(), res.data()};
                emu::run(mq_flat_body, &m);
            } else {
                MQQueue m{jobs.data(), i0, ncb};
                emu::run(mq_queue_body, &m);
            }
        }
    } else if (mode == "cm_only") {
        // context modelling alone: writes cm.bin (4 int32 per block) and
        // poff.bin (MAXP pass offsets per block) for workload statistics
        for (int k = 0; k < ncb; ++k) jobs[k].poff = &poffs[(size_t)k * MAXP];
        for (int k = 0; k < ncb; ++k) emu::run(cm_body, &jobs[k]);
        spill(dir + "/cm.bin", cm.data(), sizeof(i32x4) * cm.size());
        spill(dir + "/poff.bin", poffs.data(), sizeof(int32_t) * poffs.size());
    } else if (mode == "sorted" || mode == "two_phase" || mode == "three_phase") {
        // counting sort by chain_key over passes [qlo, qhi): MQ buckets
        // (longest first), then raw ones; 64 list entries per wavefront
        auto code_chains = [&](int qlo, int qhi) {
            std::vector<std::vector<int32_t>> bucket(CH_HIST);
            for (int k = 0; k < ncb; ++k)
                for (int q = 0; q < MAXP; ++q) {
                    const int key = chain_key(cm[k], jobs[k].poff, q, style, qlo, qhi);
                    if (key >= 0) bucket[key].push_back(k * MAXP + q);
                }
            std::vector<int32_t> list;
            int nmq = 0;
            for (int key = 0; key < CH_HIST; ++key) {
                list.insert(list.end(), bucket[key].begin(), bucket[key].end());
                if (key == CH_BUCKETS - 1) nmq = (int)list.size();
            }
            const int ntot = (int)list.size();
            for (int base = 0; base < nmq; base += 64) {
                ChainRun c{jobs.data(), list.data(), base, nmq, false};
                emu::run(chain_body, &c);
            }
            for (int base = nmq; base < ntot; base += 64) {
                ChainRun c{jobs.data(), list.data(), base, ntot, true};
                emu::run(chain_body, &c);
            }
        };
        if (mode == "sorted") {
            for (int k = 0; k < ncb; ++k) emu::run(cm_body, &jobs[k]);
            code_chains(0, MAXP);
        } else {
            // phase boundaries: passes [0, 10), [10, ...) or [0, 7), [7, 10), [10, ...)
            const std::vector<int> bounds =
                mode == "two_phase" ? std::vector<int>{0, 10, 1 << 30} : std::vector<int>{0, 7, 10, 1 << 30};
            for (size_t k = 0; k + 1 < bounds.size(); ++k) {
                g_phase.qbeg = bounds[k]; g_phase.qend = bounds[k + 1];
                for (int b = 0; b < ncb; ++b) emu::run(cm_phase_body, &jobs[b]);
                code_chains(bounds[k], bounds[k + 1] < MAXP ? bounds[k + 1] : MAXP);
            }
        }
        for (int k = 0; k < ncb; ++k) emu::run(pack_body, &jobs[k]);
        spill(dir + "/seglen.bin", seglens.data(), sizeof(int32_t) * seglens.size());
    } else if (mode == "parallel" || mode == "queue_parallel") {
        for (int k = 0; k < ncb; ++k) emu::run(cm_body, &jobs[k]);
        for (int p = 0; p < MAXP - 2; ++p)
            for (int i0 = 0; i0 < ncb; i0 += 64) {
                MQPass m{jobs.data(), i0, ncb, p, style};
                emu::run(mq_pass_body, &m);
            }
        for (int k = 0; k < ncb; ++k) emu::run(pack_body, &jobs[k]);
        spill(dir + "/seglen.bin", seglens.data(), sizeof(int32_t) * seglens.size());
    } else {
        fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    spill(dir + "/res.bin", res.data(), sizeof(rg_cb_result) * ncb);
    spill(dir + "/out.bin", slots.data(), slots.size());
    free(syms);
    free(arena);
    return 0;
}
