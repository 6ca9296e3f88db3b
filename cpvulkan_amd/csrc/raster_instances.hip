// Ahead-of-time raster kernels of one built-in fragment shader, the one
// CPV_RASTER_FS names (FsColor, FsTexLambert, ...).  The build compiles this
// file once per shader, so the kernel variants of different shaders compile
// in parallel; the library's program table reaches them through
// raster_table<FS>.
#include <hip/hip_runtime.h>

#include "cpv_internal.h"
#include "programs_builtin.h"
#include "raster_kernels.h"

#ifndef CPV_RASTER_FS
#error "CPV_RASTER_FS names the fragment shader whose raster kernels this unit instantiates"
#endif

using namespace cpv;

namespace cpvrt {

template <class FS>
const void* raster_table(int i)
{
    static const void* fns[kRasterVariants] = {
        (const void*)&raster_kernel<FS, 8, 1>,  (const void*)&raster_kernel<FS, 16, 1>,
        (const void*)&raster_kernel<FS, 16, 2>, (const void*)&raster_kernel<FS, 16, 4>,
        (const void*)&raster_kernel<FS, 32, 1>, (const void*)&raster_kernel<FS, 32, 2>,
        (const void*)&raster_kernel<FS, 32, 4>, (const void*)&raster_kernel<FS, 16, 4, true>,
        (const void*)&raster_kernel<FS, 8, 1, false, true>,  (const void*)&raster_kernel<FS, 16, 1, false, true>,
        (const void*)&raster_kernel<FS, 16, 2, false, true>, (const void*)&raster_kernel<FS, 16, 4, false, true>,
        (const void*)&raster_kernel<FS, 32, 1, false, true>, (const void*)&raster_kernel<FS, 32, 2, false, true>,
        (const void*)&raster_kernel<FS, 32, 4, false, true>};
    static_assert(kRasterAnyPrim == 7 && kRasterFixedFormats == 8, "variant order");
    return fns[i];
}

template const void* raster_table<CPV_RASTER_FS>(int);

}  // namespace cpvrt
