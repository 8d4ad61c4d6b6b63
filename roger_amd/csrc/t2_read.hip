// Tier-2 reader (host code): a JPEG 2000 codestream (.j2k) into the
// code-blocks the tier-1 decoder takes.  The reference has no tier-2 at all
// (it stops after its bit-plane coder, src/OCLEncoder.cpp:43-52).
//
// Supported (T.800 Annex A / B): one tile in one tile-part, LRCP progression,
// any number of quality layers, maximal precincts, no component
// subsampling, unsigned 8-bit samples, 1-4 components, code-block styles 0,
// RESET|TERMALL and RESET|TERMALL|BYPASS, SOP / EPH markers, quantisation
// styles none / scalar derived / scalar expounded; COM, TLM, PLM, PLT and CRG
// segments are skipped.  Anything else is refused with RG_ERR_INVALID.
//
// Packet headers (B.10): bits MSB first, a byte after 0xFF carries 7; the
// inclusion tag tree is decoded against layer + 1 until a block is
// included, then one bit per layer; the missing-MSB tree at first
// inclusion; pass counts; Lblock increments; each codeword segment's length
// in Lblock (+ floor(log2 passes) for a multi-pass segment) bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include <new>
#include <stdexcept>
#include "rg_internal.h"
#include "rg_t2.h"

namespace {

struct Bytes {
    const uint8_t *p;
    long n, pos = 0;
    bool ok = true;
    int u8() { if (pos >= n) { ok = false; return 0; } return p[pos++]; }
    int u16() { int a = u8(); return (a << 8) | u8(); }
    uint32_t u32() { uint32_t a = (uint32_t)u16(); return (a << 16) | (uint32_t)u16(); }
};

// Packet-header bit reader with the 0xFF bit-stuffing rule.
struct BitIn {
    Bytes &b;
    int cur = 0, left = 0;
    bool ff = false;
    explicit BitIn(Bytes &bb) : b(bb) {}
    int bit()
    {
        if (left == 0) {
            cur = b.u8();
            left = ff ? 7 : 8;
            ff = cur == 0xFF;
        }
        --left;
        return (cur >> left) & 1;
    }
    uint32_t bits(int k)
    {
        uint32_t v = 0;
        for (int i = 0; i < k; ++i) v = (v << 1) | (uint32_t)bit();
        return v;
    }
    // end of a header: the rest of the byte is padding; after a 0xFF the
    // writer added one more byte
    void align()
    {
        left = 0;
        if (ff) b.u8();
        ff = false;
    }
};

// Tag-tree decoder (B.10.2): node values start unknown (1 << 30).
class TagIn {
public:
    TagIn(int w, int h)
    {
        for (;;) {
            lw_.push_back(w); lh_.push_back(h); base_.push_back(total_);
            total_ += w * h;
            if (w == 1 && h == 1) break;
            w = (w + 1) / 2; h = (h + 1) / 2;
        }
        val_.assign(total_, 1 << 30);
        low_.assign(total_, 0);
    }
    // true when the leaf's value is below threshold
    bool decode(BitIn &in, int x, int y, int threshold)
    {
        int low = 0, n = 0;
        for (int l = (int)lw_.size() - 1; l >= 0; --l) {
            n = base_[l] + (y >> l) * lw_[l] + (x >> l);
            if (low > low_[n]) low_[n] = low;
            else low = low_[n];
            while (low < threshold && low < val_[n]) {
                if (in.bit()) val_[n] = low;
                else ++low;
            }
            low_[n] = low;
        }
        return val_[n] < threshold;
    }

private:
    std::vector<int> lw_, lh_, base_, val_, low_;
    int total_ = 0;
};

int pass_count(BitIn &in)
{
    if (!in.bit()) return 1;
    if (!in.bit()) return 2;
    int v = (int)in.bits(2);
    if (v != 3) return 3 + v;
    v = (int)in.bits(5);
    if (v != 31) return 6 + v;
    return 37 + (int)in.bits(7);
}

int ilog2(int v) { int l = -1; while (v) { v >>= 1; ++l; } return l; }

struct Block {
    std::vector<uint8_t> data;
    std::vector<int32_t> seg;
    int npasses = 0, nbp = 0, lblock = 3;
    bool incl = false;
};

}  // namespace

// Caps on what a header may ask the reader to hold: samples of all components
// (the decode needs about 10 bytes of HBM per sample) and code-blocks (about
// 100 bytes of host state each).  A codestream beyond them is refused before
// anything is allocated.
constexpr long RG_J2K_MAX_SAMPLES = 1L << 30;
constexpr long RG_J2K_MAX_BLOCKS = 1L << 22;

static int j2k_read_impl(const uint8_t *cs, long size, rg_j2k_info &info, std::vector<rg_cb_desc> &desc,
                         std::vector<rg_cb_result> &res, std::vector<long> &offsets, std::vector<uint8_t> &data,
                         std::vector<int32_t> &seglen)
{
    RG_CHECK(cs && size > 0, "empty codestream");
    Bytes b{cs, size};
    RG_CHECK(b.u16() == 0xFF4F, "not a JPEG 2000 codestream (no SOC)");
    memset(&info, 0, sizeof info);
    bool siz = false, cod = false, qcd = false;
    int scod = 0, prog = 0, mct = 0, xform = 0, qstyle = 0, nq = 0;
    int qv[97] = {0};
    // ---- main header ----
    for (;;) {
        const int m = b.u16();
        RG_CHECK(b.ok, "codestream ends in the main header");
        if (m == 0xFF90) break;   // SOT
        const int len = b.u16();
        RG_CHECK(len >= 2 && b.pos + len - 2 <= size, "marker 0x%04X: bad length", m);
        const long end = b.pos + len - 2;
        if (m == 0xFF51) {         // SIZ
            b.u16();
            const uint32_t xs = b.u32(), ys = b.u32(), xo = b.u32(), yo = b.u32();
            const uint32_t txs = b.u32(), tys = b.u32(), txo = b.u32(), tyo = b.u32();
            const int nc = b.u16();
            RG_CHECK(xo == 0 && yo == 0 && txo == 0 && tyo == 0, "image or tile offsets are not supported");
            RG_CHECK(txs >= xs && tys >= ys, "more than one tile is not supported");
            RG_CHECK(xs >= 1 && ys >= 1 && xs <= 65536 && ys <= 65536, "image size %ux%u out of range", xs, ys);
            RG_CHECK(nc >= 1 && nc <= 4, "%d components (1-4 supported)", nc);
            for (int c = 0; c < nc; ++c) {
                const int ss = b.u8(), dx = b.u8(), dy = b.u8();
                RG_CHECK(ss == 7, "component %d: only unsigned 8-bit samples are supported", c);
                RG_CHECK(dx == 1 && dy == 1, "component subsampling is not supported");
            }
            info.w = (int)xs; info.h = (int)ys; info.ncomp = nc; info.bitdepth = 8;
            siz = true;
        } else if (m == 0xFF52) {  // COD
            scod = b.u8();
            prog = b.u8();
            info.nlayers = b.u16();
            mct = b.u8();
            info.levels = b.u8();
            const int xcb = b.u8(), ycb = b.u8();
            RG_CHECK(xcb <= 4 && ycb <= 4, "code-block size exponents %d, %d (at most 64x64)", xcb + 2, ycb + 2);
            info.cbw = 1 << (xcb + 2);
            info.cbh = 1 << (ycb + 2);
            info.style = b.u8();
            xform = b.u8();
            cod = true;
        } else if (m == 0xFF5C) {  // QCD
            const int s = b.u8();
            info.guard_bits = s >> 5;
            qstyle = s & 0x1F;
            const bool wide = qstyle != 0;
            nq = 0;
            while (b.pos + (wide ? 2 : 1) <= end && nq < 97) qv[nq++] = wide ? b.u16() : b.u8();
            qcd = true;
        } else if (m == 0xFF64 || m == 0xFF55 || m == 0xFF57 || m == 0xFF63) {
            // COM, TLM, PLM, CRG: nothing the decode needs
        } else {
            return rg_fail(RG_ERR_INVALID, "marker 0x%04X is not supported", m);
        }
        b.pos = end;
    }
    RG_CHECK(siz && cod && qcd, "main header lacks SIZ, COD or QCD");
    RG_CHECK(!(scod & 1), "user-defined precincts are not supported");
    RG_CHECK(prog == 0, "progression order %d (only LRCP is supported)", prog);
    RG_CHECK(info.nlayers >= 1, "no quality layers");
    RG_CHECK(info.levels <= 15 && info.cbw <= 64 && info.cbh <= 64 && info.cbw * info.cbh <= 4096,
             "code-block or level parameters out of range");
    RG_CHECK(info.style == 0 || info.style == RG_STYLE_PARALLEL || info.style == RG_STYLE_PARALLEL_BYPASS,
             "code-block style 0x%02X is not supported", info.style);
    RG_CHECK(!mct || info.ncomp == 3, "component transform needs three components");
    info.filter = xform == 1 ? RG_FILTER_53 : RG_FILTER_97;
    info.colour = mct ? (xform == 1 ? RG_COLOUR_RCT : RG_COLOUR_ICT) : RG_COLOUR_NONE;
    const int nb = 3 * info.levels + 1;
    // ---- per-band exponents and steps (E.1) ----
    int eps[97], mu[97];
    for (int k = 0; k < nb; ++k) {
        const int orient = k == 0 ? 0 : 1 + (k - 1) % 3;
        const int Rb = info.bitdepth + (orient == 0 ? 0 : (orient == 3 ? 2 : 1));
        if (qstyle == 0) {
            RG_CHECK(nq >= nb, "QCD: %d exponents for %d bands", nq, nb);
            eps[k] = qv[k] >> 3; mu[k] = 0;
        } else if (qstyle == 1) {
            RG_CHECK(nq >= 1, "QCD: no step");
            const int lev = k == 0 ? info.levels : info.levels - (k - 1) / 3;   // decomposition level of the band
            eps[k] = (qv[0] >> 11) - info.levels + lev; mu[k] = qv[0] & 0x7FF;
        } else if (qstyle == 2) {
            RG_CHECK(nq >= nb, "QCD: %d steps for %d bands", nq, nb);
            eps[k] = qv[k] >> 11; mu[k] = qv[k] & 0x7FF;
        } else {
            return rg_fail(RG_ERR_INVALID, "quantisation style %d is not supported", qstyle);
        }
        info.step[k] = xform == 1 ? 1.0f : (float)std::ldexp(1.0 + mu[k] / 2048.0, Rb - eps[k]);
    }
    RG_CHECK((xform == 1) == (qstyle == 0), "transform and quantisation style disagree");
    // ---- tile-part header ----
    {
        const int lsot = b.u16(), isot = b.u16();
        const uint32_t psot = b.u32();
        const int tpsot = b.u8(), tnsot = b.u8();
        RG_CHECK(b.ok && lsot == 10 && isot == 0 && tpsot == 0 && tnsot <= 1, "only one tile in one tile-part is supported");
        const long sot_start = b.pos - 12;
        for (;;) {
            const int m = b.u16();
            RG_CHECK(b.ok, "codestream ends in the tile header");
            if (m == 0xFF93) break;   // SOD
            const int len = b.u16();
            RG_CHECK(len >= 2 && b.pos + len - 2 <= size, "marker 0x%04X: bad length", m);
            RG_CHECK(m == 0xFF64 || m == 0xFF58, "tile-header marker 0x%04X is not supported", m);   // COM, PLT
            b.pos += len - 2;
        }
        if (psot) {
            RG_CHECK(sot_start + (long)psot <= size, "tile-part runs past the codestream");
            b.n = sot_start + (long)psot;
        }
    }
    // ---- code-block layout (one component's list in band order, raster in a band) ----
    RG_CHECK((long)info.w * info.h * info.ncomp <= RG_J2K_MAX_SAMPLES, "image of %dx%dx%d samples is beyond the "
             "reader's limit of %ld", info.w, info.h, info.ncomp, RG_J2K_MAX_SAMPLES);
    const int per = rg_codeblocks(info.w, info.h, info.levels, info.cbw, info.cbh, nullptr, 0);
    RG_CHECK(per > 0, "bad image geometry");
    RG_CHECK((long)per * info.ncomp <= RG_J2K_MAX_BLOCKS, "%ld code-blocks are beyond the reader's limit of %ld",
             (long)per * info.ncomp, RG_J2K_MAX_BLOCKS);
    std::vector<rg_cb_desc> d1(per);
    rg_codeblocks(info.w, info.h, info.levels, info.cbw, info.cbh, d1.data(), per);
    std::vector<int> bfirst(nb + 1, 0), bnx(nb, 0), bny(nb, 0);
    {
        int k = 0;
        for (int band = 0; band < nb; ++band) {
            bfirst[band] = k;
            int maxx = 0, maxy = 0, cnt = 0;
            while (k < per && d1[k].band == band) { maxx = d1[k].x0; maxy = d1[k].y0; ++cnt; ++k; }
            if (cnt) {
                bnx[band] = (maxx - d1[bfirst[band]].x0) / info.cbw + 1;
                bny[band] = (maxy - d1[bfirst[band]].y0) / info.cbh + 1;
                RG_CHECK(bnx[band] * bny[band] == cnt, "band %d: block grid mismatch", band);
            }
        }
        bfirst[nb] = k;
    }
    const int ncb = per * info.ncomp;
    std::vector<Block> blk(ncb);
    std::vector<TagIn> incl, msb;
    incl.reserve((size_t)info.ncomp * nb);
    msb.reserve((size_t)info.ncomp * nb);
    for (int c = 0; c < info.ncomp; ++c)
        for (int band = 0; band < nb; ++band) {
            incl.emplace_back(bnx[band] ? bnx[band] : 1, bny[band] ? bny[band] : 1);
            msb.emplace_back(bnx[band] ? bnx[band] : 1, bny[band] ? bny[band] : 1);
        }
    const bool termall = (info.style & RG_STYLE_TERMALL) != 0;
    std::vector<int> news;
    std::vector<int32_t> lens;
    // ---- packets, LRCP ----
    for (int l = 0; l < info.nlayers; ++l)
        for (int r = 0; r <= info.levels; ++r) {
            const int b0 = r == 0 ? 0 : 1 + 3 * (r - 1), b1 = r == 0 ? 1 : b0 + 3;
            for (int c = 0; c < info.ncomp; ++c) {
                if (b.pos >= b.n) goto done;   // truncated after whole packets: the rest is empty
                if ((scod & 2) && b.pos + 6 <= b.n && cs[b.pos] == 0xFF && cs[b.pos + 1] == 0x91) b.pos += 6;   // SOP
                BitIn in(b);
                const int cbase = c * per;
                news.assign(bfirst[b1] - bfirst[b0], 0);
                lens.clear();
                if (in.bit()) {
                    for (int band = b0; band < b1; ++band) {
                        if (!bnx[band]) continue;
                        TagIn &ti = incl[(size_t)c * nb + band], &tm = msb[(size_t)c * nb + band];
                        const int Mb = info.guard_bits + eps[band] - 1;
                        for (int i = 0; i < bnx[band] * bny[band]; ++i) {
                            Block &q = blk[cbase + bfirst[band] + i];
                            const int x = i % bnx[band], y = i / bnx[band];
                            bool in_now;
                            if (!q.incl) {
                                in_now = ti.decode(in, x, y, l + 1);
                                if (in_now) {
                                    int z = 0;
                                    while (!tm.decode(in, x, y, z + 1)) {
                                        ++z;
                                        RG_CHECK(z <= 64 && b.ok, "bad missing-MSB tag tree");
                                    }
                                    q.nbp = Mb - z;
                                    RG_CHECK(q.nbp >= 0 && q.nbp <= 31, "code-block with %d bit-planes", q.nbp);
                                    q.incl = true;
                                }
                            } else {
                                in_now = in.bit() != 0;
                            }
                            if (!in_now) continue;
                            const int nnew = pass_count(in);
                            RG_CHECK(q.npasses + nnew <= std::min(RG_MAX_PASSES, 3 * q.nbp - 2),
                                     "code-block with too many coding passes");
                            while (in.bit()) {
                                ++q.lblock;
                                RG_CHECK(q.lblock < 32 && b.ok, "bad Lblock increment");
                            }
                            news[bfirst[band] + i - bfirst[b0]] = nnew;
                            if (termall) {
                                for (int g = 0; g < nnew; ++g) lens.push_back((int32_t)in.bits(q.lblock));
                            } else {
                                lens.push_back((int32_t)in.bits(q.lblock + ilog2(nnew)));
                            }
                        }
                    }
                }
                in.align();
                RG_CHECK(b.ok, "codestream ends in a packet header");
                if ((scod & 4)) {
                    RG_CHECK(b.pos + 2 <= b.n && cs[b.pos] == 0xFF && cs[b.pos + 1] == 0x92, "missing EPH marker");
                    b.pos += 2;
                }
                // body: the contributions in the same block order
                size_t li = 0;
                for (int i = bfirst[b0]; i < bfirst[b1]; ++i) {
                    const int nnew = news[i - bfirst[b0]];
                    if (!nnew) continue;
                    Block &q = blk[cbase + i];
                    const int nseg = termall ? nnew : 1;
                    for (int g = 0; g < nseg; ++g) {
                        const int32_t len = lens[li++];
                        RG_CHECK(len >= 0 && b.pos + len <= b.n, "packet body runs past the tile-part");
                        q.data.insert(q.data.end(), cs + b.pos, cs + b.pos + len);
                        if (termall) q.seg.push_back(len);
                        b.pos += len;
                    }
                    q.npasses += nnew;
                }
            }
        }
done:
    desc.resize(ncb);
    res.resize(ncb);
    offsets.resize(ncb);
    data.clear();
    seglen.assign(termall ? (size_t)ncb * RG_MAX_PASSES : 0, 0);
    for (int c = 0; c < info.ncomp; ++c)
        for (int i = 0; i < per; ++i) {
            const int k = c * per + i;
            desc[k] = d1[i];
            desc[k].comp = c;
            const Block &q = blk[k];
            RG_CHECK(q.data.size() <= 12288, "code-block %d: %zu codeword bytes (at most 12288)", k, q.data.size());
            res[k] = rg_cb_result{(int32_t)q.data.size(), q.npasses, q.npasses ? q.nbp : 0, 0};
            offsets[k] = (long)data.size();
            data.insert(data.end(), q.data.begin(), q.data.end());
            if (termall)
                for (size_t g = 0; g < q.seg.size(); ++g) seglen[(size_t)k * RG_MAX_PASSES + g] = q.seg[g];
        }
    info.ncb = ncb;
    info.data_bytes = (int64_t)data.size();
    return RG_SUCCESS;
}

int rg_j2k_read(const uint8_t *cs, long size, rg_j2k_info &info, std::vector<rg_cb_desc> &desc,
                std::vector<rg_cb_result> &res, std::vector<long> &offsets, std::vector<uint8_t> &data,
                std::vector<int32_t> &seglen)
{
    // no C++ exception may cross the C ABI: allocation failures become error codes
    try {
        return j2k_read_impl(cs, size, info, desc, res, offsets, data, seglen);
    } catch (const std::bad_alloc &) {
        return rg_fail(RG_ERR_OOM, "tier-2 reader: out of host memory");
    } catch (const std::exception &ex) {
        return rg_fail(RG_ERR_INVALID, "tier-2 reader: %s", ex.what());
    }
}

extern "C" int rg_j2k_parse(const uint8_t *cs, long size, rg_j2k_info *info, rg_cb_desc *desc, rg_cb_result *res,
                            long *offsets, uint8_t *data, int32_t *seglen)
{
    RG_CHECK(info, "rg_j2k_parse: null info");
    std::vector<rg_cb_desc> d;
    std::vector<rg_cb_result> r;
    std::vector<long> o;
    std::vector<uint8_t> bytes;
    std::vector<int32_t> s;
    if (int rc = rg_j2k_read(cs, size, *info, d, r, o, bytes, s)) return rc;
    if (desc) memcpy(desc, d.data(), sizeof(rg_cb_desc) * d.size());
    if (res) memcpy(res, r.data(), sizeof(rg_cb_result) * r.size());
    if (offsets) memcpy(offsets, o.data(), sizeof(long) * o.size());
    if (data && !bytes.empty()) memcpy(data, bytes.data(), bytes.size());
    if (seglen && !s.empty()) memcpy(seglen, s.data(), sizeof(int32_t) * s.size());
    return RG_SUCCESS;
}
