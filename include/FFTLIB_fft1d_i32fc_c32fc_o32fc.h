/*
 * fp32 single-channel complex-to-complex forward FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_fft1d_i32fc_c32fc_o32fc/
 * FFTLIB_fft1d_i32fc_c32fc_o32fc.h:27 (PBLOCK_SIZE), :55 (init), :83 (kernel),
 * :113 (checkParams). N = bufParamsX->dim_x / 2; semantics as the batched
 * kernel with one channel (..._cn.c:203-489).
 */
#ifndef FFTLIB_FFT1D_I32FC_C32FC_O32FC_H_
#define FFTLIB_FFT1D_I32FC_C32FC_O32FC_H_ 1

#include "FFTLIB_types.h"

#ifndef FFTLIB_FFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE
#define FFTLIB_FFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE (25 * SE_PARAM_SIZE + 16 * sizeof(uint32_t))
#endif

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                  FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                  FFTLIB_bufParams1D_t *bufParamsY, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                    FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                    FFTLIB_bufParams1D_t *bufParamsY, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         void *pBlock);

#ifdef __cplusplus
}
#endif
#endif
