// Bit-plane coder stage: EBCOT tier-1 with the MQ coder on every code-block of
// every component (reference: src/OCLBPC.h:23-34, src/OCLBPC.cpp:38-51).
#pragma once
#include "OCLMemoryManager.h"

template <typename T> class OCLBPC {
public:
    OCLBPC(KernelInitInfoBase initInfo, OCLMemoryManager<T> *memMgr) : initInfo(initInfo), memoryManager(memMgr) {}
    ~OCLBPC() {}
    void run(size_t codeblockX, size_t codeblockY)
    {
        if (rg_encoder_bpc(memoryManager->encoder(), (int)codeblockX, (int)codeblockY) != RG_SUCCESS)
            LogError("OCLBPC::run: %s\n", rg_last_error());
    }

private:
    KernelInitInfoBase initInfo;
    OCLMemoryManager<T> *memoryManager;
};
