// Device handle bundle and logging (reference: src/OCLUtil.h:31-39, :60-74).
#pragma once
#include <cstdarg>
#include <cstdio>
#include "ocl_platform.h"

// The reference's {context, device, commandQueue}: here all three views share
// one rg_context (device + stream).
struct ocl_args_d_t {
    ocl_args_d_t() : context(nullptr), device(0), commandQueue(nullptr) {}
    ~ocl_args_d_t() { if (context) rg_context_destroy(context); }
    rg_context *context;
    int device;
    rg_context *commandQueue;
};

inline bool &rg_quiet() { static bool q = false; return q; }

inline void LogInfo(const char *fmt, ...)
{
    if (rg_quiet()) return;
    va_list ap; va_start(ap, fmt); vfprintf(stdout, fmt, ap); va_end(ap);
}

inline void LogError(const char *fmt, ...)
{
    va_list ap; va_start(ap, fmt); vfprintf(stderr, fmt, ap); va_end(ap);
}
