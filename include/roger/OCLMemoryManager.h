// Device memory owner of one encode or decode (reference:
// src/OCLMemoryManager.h:24-88, src/OCLMemoryManager.cpp:56-272).  The
// buffers live in the library's rg_encoder / rg_decoder objects, which keep
// them resident across runs and reallocate only when the shape grows.
#pragma once
#include <cstdint>
#include <type_traits>
#include <vector>
#include "OCLUtil.h"

template <typename T> class OCLMemoryManager {
public:
    enum Role { ENCODE, DECODE };

    OCLMemoryManager(ocl_args_d_t *ocl, bool lossy, bool outputDwt)
        : ocl(ocl), width(0), height(0), _levels(0), _precision(0), numComponents(0), lossy(lossy),
          onlyDwtOut(outputDwt && lossy), role(ENCODE), enc(nullptr), dec(nullptr), quantStep(1.0 / 64)
    {
        if (rg_encoder_create(ocl->context, &enc) != RG_SUCCESS) LogError("%s\n", rg_last_error());
        if (rg_decoder_create(ocl->context, &dec) != RG_SUCCESS) LogError("%s\n", rg_last_error());
    }
    ~OCLMemoryManager()
    {
        rg_encoder_destroy(enc);
        rg_decoder_destroy(dec);
    }

    size_t getWidth() { return width; }
    size_t getHeight() { return height; }
    size_t getNumLevels() { return _levels; }
    size_t getPrecision() { return _precision; }
    size_t getNumComponents() { return numComponents; }
    bool isOnlyDwtOut() { return onlyDwtOut; }
    bool isLossy() { return lossy; }

    // Mallat-packed coefficient planes (int32 indices, or float when outputDwt).
    tDeviceMem getDWTOut()
    {
        if (onlyDwtOut) {
            float *p = nullptr;
            return rg_encoder_float_coefficients(enc, &p) == RG_SUCCESS ? (tDeviceMem)p : nullptr;
        }
        int32_t *q = nullptr;
        return rg_encoder_coefficients(enc, &q, nullptr, nullptr, nullptr) == RG_SUCCESS ? (tDeviceMem)q : nullptr;
    }
    // Per-channel int16 planes produced by OCLRGBtoPlanar::run.
    tDeviceMem getDWTOutByChannel(size_t channel)
    {
        int16_t *p = nullptr;
        return rg_encoder_channel_i16(enc, (int)channel, &p) == RG_SUCCESS ? (tDeviceMem)p : nullptr;
    }

    // Input of DWT level `level` (reference src/OCLMemoryManager.h:55-59):
    // level 0 the transform's input planes, level k the LL band after k
    // levels; nullptr for levels the transform does not have.
    tDeviceMem getDwtIn(size_t level)
    {
        void *p = nullptr;
        return rg_encoder_dwt_in(enc, (int)level, &p, nullptr, nullptr) == RG_SUCCESS ? (tDeviceMem)p : nullptr;
    }

    void init(std::vector<T *> components, size_t w, size_t h, size_t levels, size_t precision)
    {
        width = w; height = h; _levels = levels; _precision = precision; numComponents = components.size();
        std::vector<const void *> ptrs(components.begin(), components.end());
        int rc = role == ENCODE
                     ? rg_encoder_load_planar(enc, ptrs.data(), (int)ptrs.size(), sampleType(), (int)w, (int)h,
                                              (int)precision, lossy)
                     : rg_decoder_load_planar(dec, ptrs.data(), (int)ptrs.size(), sampleType(), (int)w, (int)h,
                                              (int)precision, lossy);
        if (rc != RG_SUCCESS) LogError("OCLMemoryManager::init: %s\n", rg_last_error());
    }

    tDeviceRC mapImage(tDeviceMem, void **mappedPtr)
    {
        int rc = role == ENCODE ? rg_encoder_map_dwt_out(enc, mappedPtr) : rg_decoder_map_out(dec, mappedPtr);
        if (rc != RG_SUCCESS) LogError("mapImage: %s\n", rg_last_error());
        return rc;
    }
    // Blocking read of a w x h int16 device buffer (e.g. a getDWTOutByChannel
    // plane) into host memory (reference src/OCLMemoryManager.cpp:195-214,
    // which maps width*height*sizeof(short) bytes); valid until unmapMemory.
    tDeviceRC mapBuffer(tDeviceMem buffer, void **mappedPtr)
    {
        if (!mappedPtr || !buffer) return RG_ERR_INVALID;
        bufMap.resize(width * height);
        int rc = rg_context_sync(ocl->context);
        if (rc == RG_SUCCESS)
            rc = rg_memcpy_d2h(ocl->context, bufMap.data(), (const void *)buffer, bufMap.size() * sizeof(short));
        if (rc != RG_SUCCESS) {
            LogError("mapBuffer: %s\n", rg_last_error());
            *mappedPtr = nullptr;
            return rc;
        }
        *mappedPtr = bufMap.data();
        return DeviceSuccess;
    }
    tDeviceRC unmapMemory(tDeviceMem, void *mappedPtr)
    {
        if (!mappedPtr) return RG_ERR_INVALID;
        if (!bufMap.empty() && mappedPtr == (void *)bufMap.data()) return DeviceSuccess;
        return role == ENCODE ? rg_encoder_unmap_dwt_out(enc, mappedPtr) : rg_decoder_unmap_out(dec, mappedPtr);
    }

    // Beyond the reference class.
    void setRole(Role r) { role = r; }
    void setQuantStep(double delta) { quantStep = delta; }
    double getQuantStep() const { return quantStep; }
    rg_encoder *encoder() { return enc; }
    rg_decoder *decoder() { return dec; }
    ocl_args_d_t *args() { return ocl; }

    static int sampleType()
    {
        return std::is_same<T, float>::value ? RG_F32 : (std::is_same<T, short>::value ? RG_I16 : RG_I32);
    }

private:
    ocl_args_d_t *ocl;
    size_t width, height, _levels, _precision, numComponents;
    bool lossy, onlyDwtOut;
    Role role;
    rg_encoder *enc;
    rg_decoder *dec;
    double quantStep;
    std::vector<short> bufMap;   // host copy behind mapBuffer
};
