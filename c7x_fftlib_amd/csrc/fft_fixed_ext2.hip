// Specialised complex FFT kernels, part 2 of the extra lengths (see
// fft_fixed_ext.h). Lengths: 41.
#include <hip/hip_runtime.h>

#include "fft_fixed_ext.h"
#include "fft_fixed_impl.h"

#define FFTLIB_FIXED_EXT_LENGTHS(X) \
   X(20) X(30) X(50) X(75) X(100) X(135) X(162) X(225) X(288) X(360) X(405) X(500) X(625) X(720) X(800) \
   X(972) X(1215) X(1350) X(1600) X(1875) X(2187) X(2430) X(2700) X(3072) X(3456) X(3888) X(4500) \
   X(5000) X(5625) X(6075) X(6561) X(7290) X(8000) X(9000) X(9720) X(10368) X(11520) X(12288) X(13122) \
   X(14580) X(15625)

namespace fftlib {
namespace {

bool has_here(uint32_t n)
{
   switch (n) {
#define FFTLIB_FIXED_CASE(L) case L:
      FFTLIB_FIXED_EXT_LENGTHS(FFTLIB_FIXED_CASE)
#undef FFTLIB_FIXED_CASE
      return true;
   default: return false;
   }
}

hipError_t launch_here(uint32_t n, const fixedk::FixedArgs &a, uint32_t maxGrid, hipStream_t stream)
{
   switch (n) {
#define FFTLIB_FIXED_CASE(L)                                                                                        \
   case L: return fixedk::launch_fixed<L, fixedk::KIND_C2C>(a, stream, maxGrid);
      FFTLIB_FIXED_EXT_LENGTHS(FFTLIB_FIXED_CASE)
#undef FFTLIB_FIXED_CASE
   default: return hipErrorInvalidValue;
   }
}

}  // namespace

hipError_t fixed_ext_part2_launch(uint32_t n, const fixedk::FixedArgs &a, uint32_t maxGrid, hipStream_t stream)
{
   return launch_here(n, a, maxGrid, stream);
}
bool fixed_ext_part2_has(uint32_t n) { return has_here(n); }

}  // namespace fftlib
