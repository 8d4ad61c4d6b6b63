// Tier-2: JPEG 2000 codestream assembly on the host (T.800 Annex A markers,
// Annex B packets) from the tier-1 results of an encode.  One tile, one
// quality layer, LRCP progression, maximal precincts, default code-block
// style.  The reference stops before tier-2 (src/OCLEncoder.cpp:43-52); this
// is what turns the device output into a .j2k file any decoder reads.
//
// Packet headers: tag trees for inclusion and missing MSBs (B.10.2), coding
// pass counts (Table B.4), Lblock length signalling (B.10.7.1) and bit
// stuffing after 0xFF (B.10.1).
#include <cmath>
#include <cstring>
#include <vector>
#include "rg_internal.h"
#include "rg_t2.h"

namespace {

class ByteSink {
public:
    explicit ByteSink(std::vector<uint8_t> &v) : v_(v) {}
    void u8(int x) { v_.push_back((uint8_t)x); }
    void u16(int x) { u8(x >> 8); u8(x); }
    void u32(uint32_t x) { u16((int)(x >> 16)); u16((int)(x & 0xFFFF)); }
    size_t size() const { return v_.size(); }
    std::vector<uint8_t> &buf() { return v_; }

private:
    std::vector<uint8_t> &v_;
};

// Packet-header bit writer: after a 0xFF byte the next byte carries 7 bits.
class HeaderBits {
public:
    explicit HeaderBits(ByteSink &s) : s_(s) {}
    void bit(int b)
    {
        cur_ = (cur_ << 1) | (b & 1);
        if (++n_ == room_) emit();
    }
    void bits(uint32_t v, int n)
    {
        while (n-- > 0) bit((int)(v >> n) & 1);
    }
    void finish()
    {
        if (n_ > 0) {
            cur_ <<= (room_ - n_);
            n_ = room_;
            emit();
        }
        if (last_ == 0xFF) s_.u8(0);  // a header never ends on 0xFF
    }

private:
    void emit()
    {
        s_.u8(cur_);
        last_ = cur_;
        room_ = cur_ == 0xFF ? 7 : 8;
        cur_ = 0;
        n_ = 0;
    }
    ByteSink &s_;
    int cur_ = 0, n_ = 0, room_ = 8, last_ = 0;
};

// Quad tree of minima over a grid of leaves (B.10.2).
class TagTree {
public:
    TagTree(int w, int h)
    {
        for (;;) {
            lw_.push_back(w); lh_.push_back(h); base_.push_back(total_);
            total_ += w * h;
            if (w == 1 && h == 1) break;
            w = (w + 1) / 2; h = (h + 1) / 2;
        }
        val_.assign(total_, 1 << 30);
        low_.assign(total_, 0);
        known_.assign(total_, 0);
    }
    void set(int x, int y, int v)
    {
        for (size_t l = 0; l < lw_.size(); ++l, x >>= 1, y >>= 1) {
            int &node = val_[base_[l] + y * lw_[l] + x];
            if (v < node) node = v;
        }
    }
    // Signals, root first, whether leaf (x, y) is below `threshold`.
    void code(HeaderBits &hb, int x, int y, int threshold)
    {
        int low = 0;
        for (int l = (int)lw_.size() - 1; l >= 0; --l) {
            const int n = base_[l] + (y >> l) * lw_[l] + (x >> l);
            if (low > low_[n]) low_[n] = low; else low = low_[n];
            while (low < threshold) {
                if (low >= val_[n]) {
                    if (!known_[n]) { hb.bit(1); known_[n] = 1; }
                    break;
                }
                hb.bit(0);
                ++low;
            }
            low_[n] = low;
        }
    }

private:
    std::vector<int> lw_, lh_, base_, val_, low_, known_;
    int total_ = 0;
};

void pass_count(HeaderBits &hb, int n)
{
    if (n == 1) hb.bit(0);
    else if (n == 2) hb.bits(2, 2);
    else if (n <= 5) { hb.bits(3, 2); hb.bits((uint32_t)(n - 3), 2); }
    else if (n <= 36) { hb.bits(15, 4); hb.bits((uint32_t)(n - 6), 5); }
    else { hb.bits(0x1FF, 9); hb.bits((uint32_t)(n - 37), 7); }
}

int ilog2(int v) { int l = -1; while (v) { v >>= 1; ++l; } return l; }

}  // namespace

// Exponent / mantissa of a band step (T.800 E.1) closest to delta.
void rg_band_expmant(int Rb, double delta, int *eps, int *mu)
{
    int e;
    double m = frexp(delta, &e);
    m *= 2.0; e -= 1;
    int u = (int)floor((m - 1.0) * 2048.0 + 0.5);
    if (u >= 2048) { u = 0; e += 1; }
    int ep = Rb - e;
    *eps = ep < 0 ? 0 : (ep > 31 ? 31 : ep);
    *mu = u;
}

int rg_j2k_assemble(const rg_t2_image &im, const rg_cb_desc *desc, const rg_cb_result *res, const uint8_t *const *data,
                    const int32_t *seglen, int maxp, int ncb_per_comp, std::vector<uint8_t> &out, int nlayers,
                    const int32_t *layer_np)
{
    const bool termall = (im.style & RG_STYLE_TERMALL) != 0;
    if (termall && !seglen) return rg_fail(RG_ERR_INVALID, "TERMALL style needs the per-pass segment lengths");
    if (!layer_np) nlayers = 1;
    if (nlayers < 1 || nlayers > 65535) return rg_fail(RG_ERR_INVALID, "bad layer count %d", nlayers);
    if (nlayers > 1 && !termall) return rg_fail(RG_ERR_INVALID, "quality layers need a TERMALL code-block style");
    const int nb = 3 * im.levels + 1;
    int eps[97], mu[97];
    for (int b = 0; b < nb; ++b) {
        const int orient = b == 0 ? 0 : 1 + (b - 1) % 3;
        const int Rb = im.bitdepth + (orient == 0 ? 0 : (orient == 3 ? 2 : 1));
        if (im.reversible) { eps[b] = Rb + 1; mu[b] = 0; }   // no quantisation; +1 bit for RCT chroma
        else rg_band_expmant(Rb, im.delta, &eps[b], &mu[b]);
    }
    out.clear();
    ByteSink s(out);
    s.u16(0xFF4F);                                            // SOC
    s.u16(0xFF51); s.u16(38 + 3 * im.ncomp); s.u16(0);        // SIZ
    s.u32(im.w); s.u32(im.h); s.u32(0); s.u32(0);
    s.u32(im.w); s.u32(im.h); s.u32(0); s.u32(0);
    s.u16(im.ncomp);
    for (int c = 0; c < im.ncomp; ++c) { s.u8(im.bitdepth - 1); s.u8(1); s.u8(1); }
    s.u16(0xFF52); s.u16(12); s.u8(0);                        // COD
    s.u8(0); s.u16(nlayers); s.u8(im.mct);
    s.u8(im.levels); s.u8(ilog2(im.cbw) - 2); s.u8(ilog2(im.cbh) - 2); s.u8(im.style); s.u8(im.reversible ? 1 : 0);
    s.u16(0xFF5C);                                            // QCD
    if (im.reversible) {
        s.u16(3 + nb); s.u8((im.guard_bits << 5) | 0);
        for (int b = 0; b < nb; ++b) s.u8(eps[b] << 3);
    } else {
        s.u16(3 + 2 * nb); s.u8((im.guard_bits << 5) | 2);
        for (int b = 0; b < nb; ++b) s.u16((eps[b] << 11) | mu[b]);
    }
    const size_t sot = s.size();
    s.u16(0xFF90); s.u16(10); s.u16(0); s.u32(0); s.u8(0); s.u8(1);   // SOT (Psot patched)
    s.u16(0xFF93);                                                    // SOD

    // block index ranges per band (blocks are per component in band order)
    std::vector<int> bfirst(nb + 1, 0), bnx(nb, 0), bny(nb, 0);
    {
        int k = 0;
        for (int b = 0; b < nb; ++b) {
            bfirst[b] = k;
            int maxx = 0, maxy = 0, cnt = 0;
            while (k < ncb_per_comp && desc[k].band == b) {
                maxx = desc[k].x0; maxy = desc[k].y0; ++cnt; ++k;
            }
            if (cnt) {
                const int x0 = desc[bfirst[b]].x0, y0 = desc[bfirst[b]].y0;
                bnx[b] = (maxx - x0) / im.cbw + 1;
                bny[b] = (maxy - y0) / im.cbh + 1;
                if (bnx[b] * bny[b] != cnt) return rg_fail(RG_ERR_INVALID, "band %d: block grid mismatch", b);
            }
        }
        bfirst[nb] = k;
        if (k != ncb_per_comp) return rg_fail(RG_ERR_INVALID, "code-block list out of band order");
    }
    // Packets in LRCP order.  Per block across the layers: passes sent so
    // far and Lblock; per (component, band): the inclusion tag tree (value:
    // the first layer with a pass, nlayers if none) and the missing-MSB tree.
    const int ncb_all = ncb_per_comp * im.ncomp;
    auto np_at = [&](int blk, int l) { return layer_np ? layer_np[(size_t)blk * nlayers + l] : res[blk].npasses; };
    std::vector<int> sent(ncb_all, 0), lblock(ncb_all, 3);
    std::vector<TagTree> incl, msb;
    incl.reserve((size_t)im.ncomp * nb);
    msb.reserve((size_t)im.ncomp * nb);
    for (int c = 0; c < im.ncomp; ++c)
        for (int b = 0; b < nb; ++b) {
            incl.emplace_back(bnx[b] ? bnx[b] : 1, bny[b] ? bny[b] : 1);
            msb.emplace_back(bnx[b] ? bnx[b] : 1, bny[b] ? bny[b] : 1);
            const int Mb = im.guard_bits + eps[b] - 1;
            for (int i = 0; i < bnx[b] * bny[b]; ++i) {
                const int blk = c * ncb_per_comp + bfirst[b] + i;
                const rg_cb_result &q = res[blk];
                if (q.nbp > Mb) return rg_fail(RG_ERR_INVALID, "block needs %d bit-planes, band allows %d", q.nbp, Mb);
                int first = nlayers;
                for (int l = nlayers - 1; l >= 0; --l)
                    if (np_at(blk, l) > 0) first = l;
                for (int l = 1; l < nlayers; ++l)
                    if (np_at(blk, l) < np_at(blk, l - 1)) return rg_fail(RG_ERR_INVALID, "layer passes decrease");
                incl.back().set(i % bnx[b], i / bnx[b], first);
                msb.back().set(i % bnx[b], i / bnx[b], Mb - q.nbp);
            }
        }
    for (int l = 0; l < nlayers; ++l)
        for (int r = 0; r <= im.levels; ++r) {
            const int b0 = r == 0 ? 0 : 1 + 3 * (r - 1), b1 = r == 0 ? 1 : b0 + 3;
            for (int c = 0; c < im.ncomp; ++c) {
                const int cbase = c * ncb_per_comp;
                bool any = false;
                for (int i = bfirst[b0]; i < bfirst[b1]; ++i) any |= np_at(cbase + i, l) > sent[cbase + i];
                HeaderBits hb(s);
                if (!any) { hb.bit(0); hb.finish(); continue; }
                hb.bit(1);
                for (int b = b0; b < b1; ++b) {
                    if (!bnx[b]) continue;
                    TagTree &ti = incl[(size_t)c * nb + b], &tm = msb[(size_t)c * nb + b];
                    for (int i = 0; i < bnx[b] * bny[b]; ++i) {
                        const int blk = cbase + bfirst[b] + i;
                        const rg_cb_result &q = res[blk];
                        const int x = i % bnx[b], y = i / bnx[b];
                        const int prev = sent[blk], nnew = np_at(blk, l) - prev;
                        if (prev == 0) {
                            ti.code(hb, x, y, l + 1);
                            if (nnew <= 0) continue;
                            tm.code(hb, x, y, 1 << 20);
                        } else {
                            hb.bit(nnew > 0 ? 1 : 0);
                            if (nnew <= 0) continue;
                        }
                        pass_count(hb, nnew);
                        // B.10.7: one Lblock increment, then each codeword segment's
                        // length in Lblock + log2(passes in the segment) bits (one
                        // segment per new pass with TERMALL, else one for all passes)
                        const int32_t *sl = termall ? seglen + (size_t)blk * maxp + prev : nullptr;
                        const int nseg = termall ? nnew : 1, lp = termall ? 0 : ilog2(nnew);
                        int inc = 0;
                        for (int g = 0; g < nseg; ++g) {
                            const int need = ilog2(termall ? sl[g] : q.len) + 1 - (lblock[blk] + lp);
                            inc = need > inc ? need : inc;
                        }
                        for (int t = 0; t < inc; ++t) hb.bit(1);
                        hb.bit(0);
                        lblock[blk] += inc;
                        for (int g = 0; g < nseg; ++g) hb.bits((uint32_t)(termall ? sl[g] : q.len), lblock[blk] + lp);
                    }
                }
                hb.finish();
                // body: each contributing block's new passes, in block order
                for (int i = bfirst[b0]; i < bfirst[b1]; ++i) {
                    const int blk = cbase + i;
                    const int prev = sent[blk], now = np_at(blk, l);
                    if (now <= prev) continue;
                    long from = 0, to = res[blk].len;
                    if (termall) {
                        const int32_t *sl = seglen + (size_t)blk * maxp;
                        to = 0;
                        for (int t = 0; t < now; ++t) {
                            if (t == prev) from = to;
                            to += sl[t];
                        }
                    }
                    out.insert(out.end(), data[blk] + from, data[blk] + to);
                    sent[blk] = now;
                }
            }
        }
    const uint32_t psot = (uint32_t)(out.size() - sot);
    out[sot + 6] = (uint8_t)(psot >> 24); out[sot + 7] = (uint8_t)(psot >> 16);
    out[sot + 8] = (uint8_t)(psot >> 8); out[sot + 9] = (uint8_t)psot;
    s.u16(0xFFD9);                                                    // EOC
    return RG_SUCCESS;
}
