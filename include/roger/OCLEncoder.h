// Encoder facade (reference: src/OCLEncoder.h:27-37, src/OCLEncoder.cpp:26-52):
// upload -> DWT over all levels -> planar split (more than one component) ->
// bit-plane coder on 32x32 code-blocks; with outputDwt (lossy only) the float
// DWT coefficients are the output and no bit-plane coding runs.
#pragma once
#include <vector>
#include "OCLBPC.h"
#include "OCLDWTForward.h"
#include "OCLEncodeDecode.h"
#include "OCLRGBtoPlanar.h"

template <typename T> class OCLEncoder : public OCLEncodeDecode<T> {
public:
    OCLEncoder(ocl_args_d_t *ocl, bool isLossy, bool outputDwt)
        : OCLEncodeDecode<T>(ocl, isLossy, outputDwt),
          dwt(new OCLDWTForward<T>(KernelInitInfoBase(ocl->commandQueue, "-D WIN_SIZE_X=8 -D WIN_SIZE_Y=128"),
                                   this->memoryManager)),
          bpc(new OCLBPC<T>(KernelInitInfoBase(ocl->commandQueue, "-D CODEBLOCKX=32 -D CODEBLOCKY=32"),
                            this->memoryManager)),
          rgbToPlanar(new OCLRGBtoPlanar<T>(KernelInitInfoBase(ocl->commandQueue, "-D WIN_SIZE_X=16 -D WIN_SIZE_Y=16"),
                                            this->memoryManager)) {}
    ~OCLEncoder()
    {
        delete dwt;
        delete bpc;
        delete rgbToPlanar;
    }
    void run(std::vector<T *> components, size_t w, size_t h, size_t levels, size_t precision)
    {
        OCLEncodeDecode<T>::run(components, w, h, levels, precision);
        dwt->run(this->lossy, w, h, precision, 128, 0, levels);
        if (!this->memoryManager->isOnlyDwtOut()) {
            if (components.size() > 1) rgbToPlanar->run();
            bpc->run(codeblockX, codeblockY);
        }
    }
    // Beyond the reference class: code-block size and the tier-1 output.
    void setCodeblockSize(size_t cbx, size_t cby) { codeblockX = cbx; codeblockY = cby; }
    tDeviceRC getCodeblocks(std::vector<rg_cb_desc> &desc, std::vector<rg_cb_result> &res,
                            std::vector<long> &offsets, std::vector<uint8_t> &data)
    {
        rg_encoder *e = this->memoryManager->encoder();
        int n = rg_encoder_num_codeblocks(e);
        if (n <= 0) return n < 0 ? n : RG_ERR_NOT_READY;
        long total = 0;
        desc.resize(n); res.resize(n); offsets.resize(n);
        int rc = rg_encoder_get_codeblocks(e, desc.data(), res.data(), nullptr, nullptr, 0, &total);
        if (rc != RG_SUCCESS) return rc;
        data.resize(total > 0 ? total : 1);
        return rg_encoder_get_codeblocks(e, nullptr, nullptr, offsets.data(), data.data(), (long)data.size(), nullptr);
    }

private:
    OCLDWTForward<T> *dwt;
    OCLBPC<T> *bpc;
    OCLRGBtoPlanar<T> *rgbToPlanar;
    size_t codeblockX = 32, codeblockY = 32;
};
