// Further lengths of the specialised complex kernels, instantiated in
// fft_fixed_ext0.hip .. fft_fixed_ext3.hip (every 2/3/5-smooth length from
// 16 to 16384 not listed in fft_fixed.hip), so that ragged batches of
// arbitrary lengths up to 16384 run on specialised kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftlib {
namespace fixedk {
struct FixedArgs;
}

// Each part: whether it has length n, and the launch (hipErrorInvalidValue
// when it does not). maxGrid = 0 sizes the persistent grid to the resident
// workgroup slots.
bool       fixed_ext_has(int part, uint32_t n);
hipError_t fixed_ext_launch(int part, uint32_t n, const fixedk::FixedArgs &a, uint32_t maxGrid, hipStream_t stream);
constexpr int kFixedExtParts = 4;

}  // namespace fftlib
