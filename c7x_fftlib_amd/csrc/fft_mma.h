// Launchers of the fixed-point DFT kernels (fft_mma.hip) used by the
// fft_c7xmma C-ABI layer (fftlib_mma.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftlib {

constexpr int kMmaInt16 = 1;  // FFTLIB_INT16 data, 64-bit accumulation, 32 elements per MMA row
constexpr int kMmaInt32 = 2;  // FFTLIB_INT32 data, 128-bit accumulation, 16 elements per MMA row

// y[l][j] = round(sum_k x[l][k] w[k*ldw + j] >> shift), l < rows, j, k < k2
// (dftSmall: ldw = MMA; dftLarge: ldw = k2).
hipError_t launch_mma_matrix(int dtype, const void *x, const void *w, void *y, uint32_t rows, uint32_t k2,
                             uint32_t ldw, uint32_t shift, hipStream_t stream);

// highRadixDecompositions of B channels of N points; x is work space, w1 the
// internal decomposition matrices (MMA x MMA), w2 the caller's table, shifts[5].
hipError_t launch_mma_high_radix(int dtype, void *x, const void *w1, const void *w2, void *y, uint32_t N, uint32_t B,
                                 const uint32_t *shifts, uint32_t interleave, hipStream_t stream);

}  // namespace fftlib
