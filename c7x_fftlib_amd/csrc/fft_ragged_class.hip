// The persistent ragged kernel (ragged_class_kernel in fft_fixed_impl.h) over
// the lengths of the class kernels of octaves 0..8 (16 to 4096 points, one
// 256-item workgroup size): one launch runs a whole ragged direction, longer
// transforms as column and row groups of these lengths.
#include <hip/hip_runtime.h>

#include "fft_fixed_class.h"
#include "fft_fixed_impl.h"

namespace fftlib {

#define FFTLIB_RAGGED_CLASS_LENGTHS                                                                                      \
   16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 81, 90, 96, 100, 108, 120, 125, 128, 135,  \
      144, 150, 160, 162, 180, 192, 200, 216, 225, 240, 243, 250, 256, 270, 288, 300, 320, 324, 360, 375, 384, 400, 405, \
      432, 450, 480, 486, 500, 512, 540, 576, 600, 625, 640, 648, 675, 720, 729, 750, 768, 800, 810, 864, 900, 960, 972,  \
      1000, 1024, 1080, 1125, 1152, 1200, 1215, 1250, 1280, 1296, 1350, 1440, 1458, 1500, 1536, 1600, 1620, 1728, 1800,  \
      1875, 1920, 1944, 2000, 2025, 2048, 2160, 2187, 2250, 2304, 2400, 2430, 2500, 2560, 2592, 2700, 2880, 2916, 3000,   \
      3072, 3125, 3200, 3240, 3375, 3456, 3600, 3645, 3750, 3840, 3888, 4000, 4050, 4096

hipError_t ragged_class_launch(const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   return fixedk::launch_ragged_class<FFTLIB_RAGGED_CLASS_LENGTHS>(c, grid, stream);
}

uint32_t ragged_class_resident() { return fixedk::ragged_resident_groups<FFTLIB_RAGGED_CLASS_LENGTHS>(); }

#undef FFTLIB_RAGGED_CLASS_LENGTHS

}  // namespace fftlib
