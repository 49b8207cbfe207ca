// Class kernels of the octaves [1, 10] (see fft_fixed_class.h).
#include <hip/hip_runtime.h>

#include "fft_fixed_class.h"
#include "fft_fixed_impl.h"

namespace fftlib {

hipError_t fixed_class_launch_tu0(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   switch (cls) {
   case 1:
      return fixedk::launch_class<18, 20, 24, 25, 27, 30, 32>(c, grid, stream);
   case 10:
      return fixedk::launch_class<8640, 8748, 9000, 9216, 9375, 9600, 9720, 10000, 10125, 10240, 10368, 10800, 10935, 11250,
                                  11520, 11664, 12000, 12150, 12288, 12500, 12800, 12960, 13122, 13500, 13824, 14400, 14580,
                                  15000, 15360, 15552, 15625, 16000, 16200, 16384>(c, grid, stream);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace fftlib
