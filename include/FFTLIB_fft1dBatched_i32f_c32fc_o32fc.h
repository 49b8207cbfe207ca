/*
 * fp32 batched real-to-complex FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_fft1dBatched_i32f_c32fc_o32fc/
 * FFTLIB_fft1dBatched_i32f_c32fc_o32fc.h:34 (PBLOCK_SIZE), :72 (init),
 * :116 (kernel), :156 (checkParams), :172 split_factor_gen_r2c, :180 tw_gen_r2c.
 * N real samples per channel are taken as N/2 interleaved complex points,
 * transformed by the N/2-point C2C FFT and split with the factors of pSf
 * (..._cn.c:10-125). outmode 0 writes N/2+1 bins per channel (channel stride
 * N/2+1 complex), outmode 1 writes all N bins. pW is the N/2-point table from
 * tw_gen_r2c, pSf (2N floats) from split_factor_gen_r2c. The GPU kernel fuses
 * the FFT and the split, so pXFFT is accepted but not written.
 */
#ifndef FFTLIB_FFT1DBATCHED_I32F_C32FC_O32FC_H_
#define FFTLIB_FFT1DBATCHED_I32F_C32FC_O32FC_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_FFT1DBATCHED_I32F_C32FC_O32FC_PBLOCK_SIZE (41 * SE_PARAM_SIZE + 18 * sizeof(uint32_t))

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                        FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                        FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                        FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                        FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                        uint32_t numPoints, uint32_t numChannels, uint8_t outmode,
                                                        void *pBlock);

FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX,
                                                          FFTLIB_bufParams1D_t *bufParamsX,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pW,
                                                          FFTLIB_bufParams1D_t *bufParamsW,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                          FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pSf,
                                                          FFTLIB_bufParams1D_t *bufParamsSf,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pY,
                                                          FFTLIB_bufParams1D_t *bufParamsY, uint32_t numPoints,
                                                          uint32_t numChannels, uint8_t outmode, void *pBlock);

FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                               FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                               FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                               uint32_t numPoints, uint32_t numChannels,
                                                               uint8_t outmode, void *pBlock);

/* Split factors A[i] at floats [2i, 2i+1], B[i] at [2N-2i-2, 2N-2i-1], i < N/2
 * (src/fft_c7x/FFTLIB_fft1dBatched_i32f_c32fc_o32fc/..._tw_gen.cpp:80-107). */
void split_factor_gen_r2c(FFTLIB_F32 *pSFactor, int32_t numPoints);
/* N/2-point mixed-radix twiddle table (..._tw_gen.cpp:9-78). */
void tw_gen_r2c(FFTLIB_F32 *pW, int32_t numPoints);

#ifdef __cplusplus
}
#endif
#endif
