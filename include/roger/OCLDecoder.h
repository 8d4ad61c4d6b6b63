// Decoder facade (reference: src/OCLDecoder.h:25-34, src/OCLDecoder.cpp:24-38):
// upload the Mallat-packed coefficient planes, run the inverse DWT; the mapped
// output holds the reconstructed (level-shifted) samples, interleaved.
#pragma once
#include <vector>
#include "OCLDWTRev.h"
#include "OCLEncodeDecode.h"

template <typename T> class OCLDecoder : public OCLEncodeDecode<T> {
public:
    OCLDecoder(ocl_args_d_t *ocl, bool isLossy)
        : OCLEncodeDecode<T>(ocl, isLossy, false),
          dwt(new OCLDWTRev<T>(KernelInitInfoBase(ocl->commandQueue, "-D WIN_SIZE_X=128 -D WIN_SIZE_Y=8"),
                               this->memoryManager))
    {
        this->memoryManager->setRole(OCLMemoryManager<T>::DECODE);
    }
    ~OCLDecoder() { delete dwt; }
    void run(std::vector<T *> components, size_t w, size_t h, size_t levels, size_t precision)
    {
        OCLEncodeDecode<T>::run(components, w, h, levels, precision);
        dwt->run(this->lossy, w, h, 128, 8);
    }

private:
    OCLDWTRev<T> *dwt;
};
