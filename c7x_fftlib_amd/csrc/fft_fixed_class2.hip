// Class kernels of the octaves [2, 5, 8] (see fft_fixed_class.h).
#include <hip/hip_runtime.h>

#include "fft_fixed_class.h"
#include "fft_fixed_impl.h"

namespace fftlib {

hipError_t fixed_class_launch_tu2(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   switch (cls) {
   case 2:
      return fixedk::launch_class<36, 40, 45, 48, 50, 54, 60, 64>(c, grid, stream);
   case 5:
      return fixedk::launch_class<270, 288, 300, 320, 324, 360, 375, 384, 400, 405, 432, 450, 480, 486, 500, 512>(c, grid, stream);
   case 8:
      return fixedk::launch_class<2160, 2187, 2250, 2304, 2400, 2430, 2500, 2560, 2592, 2700, 2880, 2916, 3000, 3072, 3125,
                                  3200, 3240, 3375, 3456, 3600, 3645, 3750, 3840, 3888, 4000, 4050, 4096>(c, grid, stream);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace fftlib
