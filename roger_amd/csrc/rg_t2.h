// Host-side tier-2 assembly shared by the encoder and the batch.
#pragma once
#include <cstdint>
#include <vector>
#include "../../include/roger_amd.h"

struct rg_t2_image {
    int w, h, ncomp, levels, cbw, cbh;
    int reversible, mct, guard_bits, bitdepth;
    double delta;
};

// desc: one component's blocks in band order (ncb_per_comp); res / data: all
// components, component-major.  Appends the codestream to out.
int rg_j2k_assemble(const rg_t2_image &im, const rg_cb_desc *desc, const rg_cb_result *res, const uint8_t *const *data,
                    int ncb_per_comp, std::vector<uint8_t> &out);
void rg_band_expmant(int Rb, double delta, int *eps, int *mu);
