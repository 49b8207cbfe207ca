/*
 * Top-level header (src/fftlib.h:5-16 of the reference): common types, the
 * fft_c7x kernels, the fixed-point DFT (fft_c7xmma) kernels and the GPU
 * controls of this build.
 */
#ifndef FFTLIB_H_
#define FFTLIB_H_

#include "FFTLIB_bufParams.h"
#include "FFTLIB_types.h"
#include "c7x_fft.h"
#include "fftlib_fft.h"
#include "fftlib_gpu.h"

#endif
