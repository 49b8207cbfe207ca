/*
 * int16 batched complex FFT, bit-exact fixed point.
 *
 * Drop-in for src/fft_c7x/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/
 * FFTLIB_fft1dBatched_i16sc_c16sc_o16sc.h:29 (PBLOCK_SIZE), :64 (init),
 * :102 (kernel), :142 (checkParams). Radix-4 DIF stages (final radix-2 when
 * N = 2*4^k) with 16-bit wrap-around adds, VCROT90H-style saturation,
 * round-half-up Q15 twiddle products and a per-stage rounding right shift
 * pShift[s] (one entry shared by the 8-point and final 2-point stages);
 * bit-reversed output (..._cn.c:191-338). Output is bit-identical to the
 * reference natural-C kernel.
 */
#ifndef FFTLIB_FFT1DBATCHED_I16SC_C16SC_O16SC_H_
#define FFTLIB_FFT1DBATCHED_I16SC_C16SC_O16SC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_FFT1DBATCHED_I16SC_C16SC_O16SC_PBLOCK_SIZE (13 * SE_PARAM_SIZE)

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_init(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                         FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                         FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                         FFTLIB_bufParams1D_t *bufParamsShift, uint32_t numPoints,
                                                         uint32_t numChannels, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_kernel(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                           FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                           FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                           FFTLIB_bufParams1D_t *bufParamsShift, uint32_t numPoints,
                                                           uint32_t numChannels, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_checkParams(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                int16_t *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                int16_t *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                uint32_t *pShift, FFTLIB_bufParams1D_t *bufParamsShift,
                                                                uint32_t numPoints, uint32_t numChannels,
                                                                void *pBlock);

#ifdef __cplusplus
}
#endif
#endif
