// Planar stage (reference: src/OCLRGBtoPlanar.h:23-34,
// src/OCLRGBtoPlanar.cpp:36-80): leaves one int16 plane per channel
// (OCLMemoryManager::getDWTOutByChannel), saturated from the int32
// coefficient planes.  The colour transform itself (interleaved RGB -> planar
// + RCT/ICT) is rg_colour_fwd on the native path.
#pragma once
#include "OCLMemoryManager.h"

template <typename T> class OCLRGBtoPlanar {
public:
    OCLRGBtoPlanar(KernelInitInfoBase initInfo, OCLMemoryManager<T> *memMgr) : initInfo(initInfo), memoryManager(memMgr) {}
    ~OCLRGBtoPlanar() {}
    void run()
    {
        if (rg_encoder_planar_i16(memoryManager->encoder()) != RG_SUCCESS)
            LogError("OCLRGBtoPlanar::run: %s\n", rg_last_error());
    }

private:
    KernelInitInfoBase initInfo;
    OCLMemoryManager<T> *memoryManager;
};
