// Host-side tier-2 assembly shared by the encoder and the batch.
#pragma once
#include <cstdint>
#include <vector>
#include "../../include/roger_amd.h"

struct rg_t2_image {
    int w, h, ncomp, levels, cbw, cbh;
    int reversible, mct, guard_bits, bitdepth;
    double delta;
    int style;   // code-block style byte of COD (RG_STYLE_*)
};

// desc: one component's blocks in band order (ncb_per_comp); res / data: all
// components, component-major; seglen (TERMALL style): per block, the length
// of each pass's codeword segment (maxp entries per block).  Appends the
// codestream to out.
// layer_np (optional, TERMALL styles): per block, the passes it has after
// each of nlayers quality layers (non-decreasing); null: one layer.
int rg_j2k_assemble(const rg_t2_image &im, const rg_cb_desc *desc, const rg_cb_result *res, const uint8_t *const *data,
                    const int32_t *seglen, int maxp, int ncb_per_comp, std::vector<uint8_t> &out, int nlayers = 1,
                    const int32_t *layer_np = nullptr);
void rg_band_expmant(int Rb, double delta, int *eps, int *mu);

// Tier-2 reader: the codestream's parameters (info) and its code-blocks in the
// encoder's order (desc / res / offsets into data; seglen for TERMALL styles,
// RG_MAX_PASSES per block).
int rg_j2k_read(const uint8_t *cs, long size, rg_j2k_info &info, std::vector<rg_cb_desc> &desc,
                std::vector<rg_cb_result> &res, std::vector<long> &offsets, std::vector<uint8_t> &data,
                std::vector<int32_t> &seglen);
