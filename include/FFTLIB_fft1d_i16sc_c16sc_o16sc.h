/*
 * int16 single-channel complex FFT, bit-exact fixed point.
 *
 * Drop-in for src/fft_c7x/FFTLIB_fft1d_i16sc_c16sc_o16sc/
 * FFTLIB_fft1d_i16sc_c16sc_o16sc.h:26 (PBLOCK_SIZE), :59 (init), :94 (kernel),
 * :129 (checkParams). N = bufParamsX->dim_x / 2; arithmetic identical to the
 * batched int16 kernel.
 */
#ifndef FFTLIB_FFT1D_I16SC_C16SC_O16SC_H_
#define FFTLIB_FFT1D_I16SC_C16SC_O16SC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_FFT1D_I16SC_C16SC_O16SC_PBLOCK_SIZE (9 * SE_PARAM_SIZE)

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_init(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                  FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                  FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                  FFTLIB_bufParams1D_t *bufParamsShift, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_kernel(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                    FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                    FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                    FFTLIB_bufParams1D_t *bufParamsShift, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_checkParams(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                         FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                         FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                         FFTLIB_bufParams1D_t *bufParamsShift, void *pBlock);

#ifdef __cplusplus
}
#endif
#endif
