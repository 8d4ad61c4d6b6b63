// Device discovery (reference: src/OCLDeviceManager.h:27-39).  The reference
// tries a GPU then a CPU OpenCL device; here there is exactly one kind of
// device, the HIP GPU, and the CPU slot stays empty.
#pragma once
#include "OCLUtil.h"

enum eDeviceType { CPU, GPU, NUM_DEVICE_TYPES };

class OCLDeviceManager {
public:
    OCLDeviceManager() : ocl_gpu(nullptr) {}
    ~OCLDeviceManager() { delete ocl_gpu; }
    int init(int device = 0)
    {
        if (ocl_gpu) return DeviceSuccess;
        ocl_args_d_t *a = new ocl_args_d_t();
        int rc = rg_context_create(device, &a->context);
        if (rc != RG_SUCCESS) {
            LogError("OCLDeviceManager::init: %s\n", rg_last_error());
            delete a;
            return rc;
        }
        a->device = device;
        a->commandQueue = a->context;
        ocl_gpu = a;
        return DeviceSuccess;
    }
    ocl_args_d_t *getInfo(eDeviceType type) { return type == GPU ? ocl_gpu : nullptr; }
    ocl_args_d_t *getInfo() { return ocl_gpu; }

private:
    ocl_args_d_t *ocl_gpu;
};
