// Compiled shader programs: the vertex kernel and the raster kernels of one
// (vertex shader, fragment shader) pair, either ahead-of-time instantiations
// in this library or a hiprtc-compiled code object.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

struct KernelRef {
    const void* host_fn = nullptr;        // ahead-of-time kernel (hipLaunchKernel)
    hipFunction_t module_fn = nullptr;    // kernel of a loaded code object (hipModuleLaunchKernel)
};

struct cpv_program {
    hipModule_t module = nullptr;
    KernelRef vertex;
    KernelRef raster[15];                 // tile configurations 8x1, 16x1, 16x2, 16x4, 32x1, 32x2, 32x4,
                                          // then 16x4 with point and line records, then the seven
                                          // configurations for the fixed formats
    int vs_outputs = 0;
    int psize = -1;                       // vertex output component holding gl_PointSize, -1: none
    int fs_inputs = 0;
    std::string source;                   // generated HIP source
    std::string log;                      // translator / hiprtc log
};
