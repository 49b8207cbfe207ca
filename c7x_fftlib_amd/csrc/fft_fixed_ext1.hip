// Specialised complex FFT kernels, part 1 of the extra lengths (see
// fft_fixed_ext.h). Lengths: 41.
#include <hip/hip_runtime.h>

#include "fft_fixed_ext.h"
#include "fft_fixed_impl.h"

#define FFTLIB_FIXED_EXT_LENGTHS(X) \
   X(24) X(36) X(54) X(80) X(108) X(144) X(180) X(243) X(300) X(375) X(432) X(540) X(640) X(729) X(810) \
   X(1080) X(1250) X(1440) X(1620) X(1944) X(2250) X(2500) X(2880) X(3125) X(3600) X(4050) X(4608) \
   X(5120) X(5760) X(6144) X(6750) X(7500) X(8100) X(9216) X(10000) X(10800) X(11664) X(12500) X(13500) \
   X(15000) X(16000)

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

hipError_t fixed_ext_part1_launch(uint32_t n, const fixedk::FixedArgs &a, uint32_t maxGrid, hipStream_t stream)
{
   return launch_here(n, a, maxGrid, stream);
}
bool fixed_ext_part1_has(uint32_t n) { return has_here(n); }

}  // namespace fftlib
