/*
 * fp32 batched complex-to-complex inverse FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc/
 * FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc.h:27 (PBLOCK_SIZE), :59 (init),
 * :92 (kernel), :127 (checkParams). Semantics follow ..._cn.c:232-512:
 * y = conj(FFT(conj(x) * (1.0f/N))) using the FORWARD twiddle table.
 */
#ifndef FFTLIB_IFFT1DBATCHED_I32FC_C32FC_O32FC_H_
#define FFTLIB_IFFT1DBATCHED_I32FC_C32FC_O32FC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_IFFT1DBATCHED_I32FC_C32FC_O32FC_PBLOCK_SIZE (31 * SE_PARAM_SIZE + 18 * sizeof(uint32_t))

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                          FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                          FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                          uint32_t numPoints, uint32_t numChannels, void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                            uint32_t numPoints, uint32_t numChannels, void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                 FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                 FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                 uint32_t numPoints, uint32_t numChannels,
                                                                 void *pBlock);

#ifdef __cplusplus
}
#endif
#endif
