// Specialised complex FFT kernels, part 0 of the extra lengths (see
// fft_fixed_ext.h). Lengths: 41.
#include <hip/hip_runtime.h>

#include "fft_fixed_ext.h"
#include "fft_fixed_impl.h"

#define FFTLIB_FIXED_EXT_LENGTHS(X) \
   X(25) X(40) X(60) X(81) X(120) X(150) X(200) X(250) X(320) X(384) X(450) X(576) X(648) X(750) X(864) \
   X(1125) X(1280) X(1458) X(1728) X(2025) X(2304) X(2560) X(2916) X(3240) X(3645) X(4320) X(4800) \
   X(5184) X(5832) X(6250) X(6912) X(7680) X(8640) X(9375) X(10125) X(10935) X(12000) X(12800) X(13824) \
   X(15360) X(16200)

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

hipError_t fixed_ext_part0_launch(uint32_t n, const fixedk::FixedArgs &a, uint32_t maxGrid, hipStream_t stream)
{
   return launch_here(n, a, maxGrid, stream);
}
bool fixed_ext_part0_has(uint32_t n) { return has_here(n); }

}  // namespace fftlib
