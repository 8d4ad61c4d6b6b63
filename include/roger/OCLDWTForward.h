// Forward DWT stage (reference: src/OCLDWTForward.h:25-39,
// src/OCLDWTForward.cpp:49-137).  run(level = 0) transforms all `levels`
// levels (5/3 when !lossy; 9/7 + quantisation when lossy, float output when
// the memory manager is in outputDwt mode).  windowX/windowY were the
// reference's OpenCL work-group shape; the HIP kernels use fixed 64x64 tiles.
#pragma once
#include "OCLMemoryManager.h"

template <typename T> class OCLDWTForward {
public:
    OCLDWTForward(KernelInitInfoBase initInfo, OCLMemoryManager<T> *memMgr) : info(initInfo), memoryManager(memMgr) {}
    ~OCLDWTForward() {}
    void run(bool lossy, size_t w, size_t h, size_t windowX, size_t windowY, size_t level, size_t levels)
    {
        (void)windowX; (void)windowY;
        if (lossy != memoryManager->isLossy() || w != memoryManager->getWidth() || h != memoryManager->getHeight()) {
            LogError("OCLDWTForward::run: arguments do not match the loaded image\n");
            return;
        }
        if (level != 0) {
            LogError("OCLDWTForward::run: only a run starting at level 0 is supported\n");
            return;
        }
        if (rg_encoder_dwt(memoryManager->encoder(), (int)levels, memoryManager->isOnlyDwtOut(),
                           memoryManager->getQuantStep()) != RG_SUCCESS)
            LogError("OCLDWTForward::run: %s\n", rg_last_error());
    }

private:
    KernelInitInfoBase info;
    OCLMemoryManager<T> *memoryManager;
};
