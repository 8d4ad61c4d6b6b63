// Shared base of the encoder and decoder (reference: src/OCLEncodeDecode.h:23-38,
// src/OCLEncodeDecode.cpp:22-49).
#pragma once
#include <vector>
#include "OCLMemoryManager.h"

template <typename T> class OCLEncodeDecode {
public:
    OCLEncodeDecode(ocl_args_d_t *ocl, bool isLossy, bool outputDwt)
        : _ocl(ocl), lossy(isLossy), memoryManager(new OCLMemoryManager<T>(ocl, isLossy, outputDwt)) {}
    virtual ~OCLEncodeDecode() { delete memoryManager; }
    tDeviceRC mapDWTOut(void **mappedPtr) { return memoryManager->mapImage(nullptr, mappedPtr); }
    tDeviceRC unmapDWTOut(void *mappedPtr) { return memoryManager->unmapMemory(nullptr, mappedPtr); }
    void finish()
    {
        if (rg_context_sync(_ocl->commandQueue) != RG_SUCCESS) LogError("finish: %s\n", rg_last_error());
    }
    OCLMemoryManager<T> *getMemoryManager() { return memoryManager; }

protected:
    void run(std::vector<T *> components, size_t w, size_t h, size_t levels, size_t precision)
    {
        memoryManager->init(components, w, h, levels, precision);
    }
    ocl_args_d_t *_ocl;
    bool lossy;
    OCLMemoryManager<T> *memoryManager;
};
