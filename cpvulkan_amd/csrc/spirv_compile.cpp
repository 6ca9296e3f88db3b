// SPIR-V -> HIP translation and hiprtc compilation of shader programs.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cpv.h"
#include "cpv_program.h"

int cpv_spirv_compile(cpv_context*, const uint32_t*, size_t, const char*, const uint32_t*, size_t, const char*, cpv_program** out,
                      std::string* err)
{
    *out = nullptr;
    *err = "SPIR-V programs are not built yet";
    return CPV_ERROR_UNSUPPORTED;
}
