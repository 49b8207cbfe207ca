/*
 * Umbrella of the fixed-point DFT kernels (src/fft_c7xmma/fftlib_fft.h:8-13 of
 * the reference, which includes them for C7100/C7120 only; here always).
 */
#ifndef FFTLIB_FFT_MMA_H_
#define FFTLIB_FFT_MMA_H_ 1

#include "FFTLIB_FFT_dftLarge_ixX_cxX_oxX.h"
#include "FFTLIB_FFT_dftSmall_ixX_cxX_oxX.h"
#include "FFTLIB_FFT_fft_ixX_cxX_oxX.h"
#include "FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX.h"

#endif
