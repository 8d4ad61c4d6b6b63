// Inverse DWT stage (reference: src/OCLDWTRev.h:25-37, src/OCLDWTRev.cpp:39-59,
// whose kernels were never reachable).  Reconstructs the samples from the
// Mallat-packed coefficients loaded by the memory manager, over the number of
// levels given to OCLDecoder::run.
#pragma once
#include "OCLMemoryManager.h"

template <typename T> class OCLDWTRev {
public:
    OCLDWTRev(KernelInitInfoBase initInfo, OCLMemoryManager<T> *memMgr) : info(initInfo), memoryManager(memMgr) {}
    ~OCLDWTRev() {}
    void run(bool lossy, size_t w, size_t h, size_t windowX, size_t windowY)
    {
        (void)windowX; (void)windowY;
        if (lossy != memoryManager->isLossy() || w != memoryManager->getWidth() || h != memoryManager->getHeight()) {
            LogError("OCLDWTRev::run: arguments do not match the loaded coefficients\n");
            return;
        }
        if (rg_decoder_idwt(memoryManager->decoder(), (int)memoryManager->getNumLevels()) != RG_SUCCESS)
            LogError("OCLDWTRev::run: %s\n", rg_last_error());
    }

private:
    KernelInitInfoBase info;
    OCLMemoryManager<T> *memoryManager;
};
