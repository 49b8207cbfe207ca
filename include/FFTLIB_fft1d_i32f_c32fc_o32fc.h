/*
 * fp32 single-channel real-to-complex FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_fft1d_i32f_c32fc_o32fc/
 * FFTLIB_fft1d_i32f_c32fc_o32fc.h:32 (PBLOCK_SIZE), :66 (init), :104 (kernel),
 * :138 (checkParams), :152 split_factor_gen_r2c, :160 tw_gen_r2c. N =
 * bufParamsX->dim_x real samples; always writes the full N-bin spectrum
 * (..._cn.c:8-154).
 */
#ifndef FFTLIB_FFT1D_I32F_C32FC_O32FC_H_
#define FFTLIB_FFT1D_I32F_C32FC_O32FC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_FFT1D_I32F_C32FC_O32FC_PBLOCK_SIZE (35 * SE_PARAM_SIZE + 18 * sizeof(uint32_t))

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                 FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                                 FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                                 FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                                 FFTLIB_bufParams1D_t *bufParamsY, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                   FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                   void *pBlock);

FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                        FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                        FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                        void *pBlock);

void split_factor_gen_r2c(FFTLIB_F32 *pSFactor, int32_t numPoints);
void tw_gen_r2c(FFTLIB_F32 *pW, int32_t numPoints);

#ifdef __cplusplus
}
#endif
#endif
