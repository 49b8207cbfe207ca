/*
 * fp32 batched complex-to-complex forward FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/
 * FFTLIB_fft1dBatched_i32fc_c32fc_o32fc.h:27 (PBLOCK_SIZE), :59 (init),
 * :92 (kernel), :127 (checkParams). Semantics follow the reference natural-C
 * kernel (..._cn.c:210-495): unscaled DFT Y[k] = sum_n x[n] e^{-2 pi i kn/N}
 * per channel, N = 2^a 3^b 5^c, channel l at floats [2Nl, 2N(l+1)), twiddles
 * from the caller's pW table (mixed-radix layout of the reference test
 * driver's tw_gen, exported here as FFTLIB_tw_gen_c2c).
 *
 * On this platform _kernel runs on the GPU (see FFTLIB_FUNCTION_STYLE); pX is
 * not clobbered.
 */
#ifndef FFTLIB_FFT1DBATCHED_I32FC_C32FC_O32FC_H_
#define FFTLIB_FFT1DBATCHED_I32FC_C32FC_O32FC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_FFT1DBATCHED_I32FC_C32FC_O32FC_PBLOCK_SIZE (31 * SE_PARAM_SIZE + 18 * sizeof(uint32_t))

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         uint32_t numPoints, uint32_t numChannels, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                           FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                           FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                           uint32_t numPoints, uint32_t numChannels, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                uint32_t numPoints, uint32_t numChannels,
                                                                void *pBlock);

#ifdef __cplusplus
}
#endif
#endif
