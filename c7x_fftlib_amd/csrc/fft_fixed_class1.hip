// Class kernels of the octaves [0, 4, 9] (see fft_fixed_class.h).
#include <hip/hip_runtime.h>

#include "fft_fixed_class.h"
#include "fft_fixed_impl.h"

namespace fftlib {

hipError_t fixed_class_launch_tu1(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   switch (cls) {
   case 0:
      return fixedk::launch_class<16>(c, grid, stream);
   case 4:
      return fixedk::launch_class<135, 144, 150, 160, 162, 180, 192, 200, 216, 225, 240, 243, 250, 256>(c, grid, stream);
   case 9:
      return fixedk::launch_class<4320, 4374, 4500, 4608, 4800, 4860, 5000, 5120, 5184, 5400, 5625, 5760, 5832, 6000, 6075,
                                  6144, 6250, 6400, 6480, 6561, 6750, 6912, 7200, 7290, 7500, 7680, 7776, 8000, 8100, 8192>(c, grid, stream);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace fftlib
