// Multi-length ("class") complex FFT kernels: every 2/3/5-smooth length from
// 16 to 16384, grouped by octave (2^k, 2^(k+1)] so that the lengths of a class
// share the work-item count and transforms per workgroup. The classes are
// instantiated in fft_fixed_class0.hip .. fft_fixed_class3.hip so that they
// compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftlib {
namespace fixedk {
struct ClassArgs;
}

constexpr int kFixedClasses = 11;

// Class of length n (-1 when n has none).
int fixed_class_of(uint32_t n);

// Launches class cls over c's items; maxGrid > 0 caps the persistent grid.
hipError_t fixed_class_launch(int cls, const fixedk::ClassArgs &c, uint32_t maxGrid, hipStream_t stream);

// Per translation unit: launch of the classes it instantiates
// (hipErrorInvalidValue for others).
hipError_t fixed_class_launch_tu0(int cls, const fixedk::ClassArgs &c, uint32_t maxGrid, hipStream_t stream);
hipError_t fixed_class_launch_tu1(int cls, const fixedk::ClassArgs &c, uint32_t maxGrid, hipStream_t stream);
hipError_t fixed_class_launch_tu2(int cls, const fixedk::ClassArgs &c, uint32_t maxGrid, hipStream_t stream);
hipError_t fixed_class_launch_tu3(int cls, const fixedk::ClassArgs &c, uint32_t maxGrid, hipStream_t stream);

}  // namespace fftlib
