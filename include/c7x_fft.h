/*
 * Umbrella header of the fft_c7x kernels (src/fft_c7x/c7x_fft.h:7-14 of the
 * reference). The reference omits the two single-channel inverse headers
 * (their PBLOCK macro names clash with the forward ones there); both are
 * included here.
 */
#ifndef C7X_FFT_H_
#define C7X_FFT_H_

#include "FFTLIB_fft1dBatched_i16sc_c16sc_o16sc.h"
#include "FFTLIB_fft1dBatched_i32f_c32fc_o32fc.h"
#include "FFTLIB_fft1dBatched_i32fc_c32fc_o32fc.h"
#include "FFTLIB_fft1d_i16sc_c16sc_o16sc.h"
#include "FFTLIB_fft1d_i32f_c32fc_o32fc.h"
#include "FFTLIB_fft1d_i32fc_c32fc_o32fc.h"
#include "FFTLIB_ifft1dBatched_i32fc_c32fc_o32f.h"
#include "FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc.h"
#include "FFTLIB_ifft1d_i32fc_c32fc_o32f.h"
#include "FFTLIB_ifft1d_i32fc_c32fc_o32fc.h"

#endif
