/*
 * fp32 single-channel complex-to-complex inverse FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_ifft1d_i32fc_c32fc_o32fc/
 * FFTLIB_ifft1d_i32fc_c32fc_o32fc.h:27 (PBLOCK_SIZE), :55 (init), :83 (kernel),
 * :113 (checkParams). The reference header reuses the forward macro name
 * FFTLIB_FFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE (with a larger value); that name
 * is kept when no forward header defined it first, and the inverse-specific
 * name is added. The GPU plan fits in either size.
 */
#ifndef FFTLIB_IFFT1D_I32FC_C32FC_O32FC_H_
#define FFTLIB_IFFT1D_I32FC_C32FC_O32FC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_IFFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE (27 * SE_PARAM_SIZE + 16 * sizeof(uint32_t))
#ifndef FFTLIB_FFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE
#define FFTLIB_FFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE FFTLIB_IFFT1D_I32FC_C32FC_O32FC_PBLOCK_SIZE
#endif

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                   FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                   FFTLIB_bufParams1D_t *bufParamsY, void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                     FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                     FFTLIB_bufParams1D_t *bufParamsY, void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                          FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                          FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                          void *pBlock);

#ifdef __cplusplus
}
#endif
#endif
