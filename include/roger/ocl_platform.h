// Platform types of the drop-in classes (reference: src/ocl_platform.h:22-65).
// The OpenCL handles become the roger_amd C-ABI handles: a "queue" is an
// rg_context (one HIP device + stream) and device memory is a plain pointer.
#pragma once
#include <string>
#include "../roger_amd.h"

typedef void *tDeviceMem;
typedef int tDeviceRC;
#define DeviceSuccess RG_SUCCESS

struct QueueInfo {
    QueueInfo(rg_context *queue) : cmd_queue(queue) {}
    rg_context *cmd_queue;
};

// Build options are kept for interface compatibility; the HIP kernels are
// compiled ahead of time, so "-D WIN_SIZE_X/Y" etc. are launch-shape hints only.
struct KernelInitInfoBase : QueueInfo {
    KernelInitInfoBase(rg_context *queue, std::string bldOptions) : QueueInfo(queue), buildOptions(bldOptions) {}
    KernelInitInfoBase(const KernelInitInfoBase &other) : QueueInfo(other.cmd_queue), buildOptions(other.buildOptions) {}
    std::string buildOptions;
};

struct KernelInitInfo : KernelInitInfoBase {
    KernelInitInfo(rg_context *queue, std::string progName, std::string knlName, std::string bldOptions)
        : KernelInitInfoBase(queue, bldOptions), programName(progName), kernelName(knlName) {}
    KernelInitInfo(KernelInitInfoBase info, std::string progName, std::string knlName)
        : KernelInitInfoBase(info), programName(progName), kernelName(knlName) {}
    std::string programName;
    std::string kernelName;
};
