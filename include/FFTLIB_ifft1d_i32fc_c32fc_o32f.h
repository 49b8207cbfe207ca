/*
 * fp32 single-channel complex-to-real inverse FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_ifft1d_i32fc_c32fc_o32f/
 * FFTLIB_ifft1d_i32fc_c32fc_o32f.h:32 (PBLOCK_SIZE), :67 (init), :106 (kernel),
 * :141 (checkParams), :154 ifft_c2r_split_factor_gen, :159 ifft_c2r_tw_gen.
 * N = bufParamsX->dim_x; the input is the full N-bin spectrum (the batched
 * kernel's inMode 1) (..._cn.c:8-78).
 */
#ifndef FFTLIB_IFFT1D_I32FC_C32FC_O32F_H_
#define FFTLIB_IFFT1D_I32FC_C32FC_O32F_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_IFFT1D_I32FC_C32FC_O32F_PBLOCK_SIZE (34 * SE_PARAM_SIZE + 16 * sizeof(uint32_t))

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                  FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                                  FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                                  FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                                  FFTLIB_bufParams1D_t *bufParamsY, void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                    FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pSf,
                                                    FFTLIB_bufParams1D_t *bufParamsSf,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                    void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         void *pBlock);

void ifft_c2r_split_factor_gen(FFTLIB_F32 *pSFactor, int32_t numPoints);
void ifft_c2r_tw_gen(FFTLIB_F32 *pW, int32_t numPoints);

#ifdef __cplusplus
}
#endif
#endif
