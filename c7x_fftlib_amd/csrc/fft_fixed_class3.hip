// Class kernels of the octaves [3, 6, 7] (see fft_fixed_class.h).
#include <hip/hip_runtime.h>

#include "fft_fixed_class.h"
#include "fft_fixed_impl.h"

namespace fftlib {

hipError_t fixed_class_launch_tu3(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   switch (cls) {
   case 3:
      return fixedk::launch_class<72, 75, 80, 81, 90, 96, 100, 108, 120, 125, 128>(c, grid, stream);
   case 6:
      return fixedk::launch_class<540, 576, 600, 625, 640, 648, 675, 720, 729, 750, 768, 800, 810, 864, 900, 960, 972, 1000,
                                  1024>(c, grid, stream);
   case 7:
      return fixedk::launch_class<1080, 1125, 1152, 1200, 1215, 1250, 1280, 1296, 1350, 1440, 1458, 1500, 1536, 1600, 1620,
                                  1728, 1800, 1875, 1920, 1944, 2000, 2025, 2048>(c, grid, stream);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace fftlib
