// Specialised complex FFT kernels, part 3 of the extra lengths (see
// fft_fixed_ext.h). Lengths: 41.
#include <hip/hip_runtime.h>

#include "fft_fixed_ext.h"
#include "fft_fixed_impl.h"

#define FFTLIB_FIXED_EXT_LENGTHS(X) \
   X(18) X(27) X(45) X(72) X(90) X(125) X(160) X(216) X(270) X(324) X(400) X(486) X(600) X(675) X(768) \
   X(900) X(1152) X(1296) X(1500) X(1800) X(2160) X(2400) X(2592) X(3000) X(3375) X(3750) X(4374) \
   X(4860) X(5400) X(6000) X(6480) X(7200) X(7776) X(8748) X(9600) X(10240) X(11250) X(12150) X(12960) \
   X(14400) X(15552)

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

hipError_t fixed_ext_part3_launch(uint32_t n, const fixedk::FixedArgs &a, uint32_t maxGrid, hipStream_t stream)
{
   return launch_here(n, a, maxGrid, stream);
}
bool fixed_ext_part3_has(uint32_t n) { return has_here(n); }

}  // namespace fftlib
