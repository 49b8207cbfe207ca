/*
 * fp32 batched complex-to-real inverse FFT.
 *
 * Drop-in for src/fft_c7x/FFTLIB_ifft1dBatched_i32fc_c32fc_o32f/
 * FFTLIB_ifft1dBatched_i32fc_c32fc_o32f.h:34 (PBLOCK_SIZE), :73 (init),
 * :118 (kernel), :159 (checkParams), :174 ifft_c2r_split_factor_gen,
 * :179 ifft_c2r_tw_gen. Per channel: Z[i] = X[i]A'[i] + V[i]B'[i], i < N/2,
 * with V[i] = conj(X[N/2-i]) for inMode 0 (input N/2+1 bins, channel stride
 * N/2+1) or X[N/2+i] for inMode 1 (input N bins); then the N/2-point C2C IFFT
 * (scale 1/(N/2)) yields N reals (..._cn.c:8-91). The GPU kernel fuses the
 * split and the IFFT, so pXFFT is accepted but not written.
 */
#ifndef FFTLIB_IFFT1DBATCHED_I32FC_C32FC_O32F_H_
#define FFTLIB_IFFT1DBATCHED_I32FC_C32FC_O32F_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_IFFT1DBATCHED_I32FC_C32FC_O32F_PBLOCK_SIZE (41 * SE_PARAM_SIZE + 18 * sizeof(uint32_t))

#ifdef __cplusplus
extern "C" {
#endif

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                         FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         uint32_t numPoints, uint32_t numChannels, uint8_t inMode,
                                                         void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX,
                                                           FFTLIB_bufParams1D_t *bufParamsX,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pW,
                                                           FFTLIB_bufParams1D_t *bufParamsW,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                           FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pSf,
                                                           FFTLIB_bufParams1D_t *bufParamsSf,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pY,
                                                           FFTLIB_bufParams1D_t *bufParamsY, uint32_t numPoints,
                                                           uint32_t numChannels, uint8_t inMode, void *pBlock);

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                uint32_t numPoints, uint32_t numChannels,
                                                                uint8_t inMode, void *pBlock);

/* A'[i] at floats [2i, 2i+1], B'[i] at [N+2i, N+2i+1], i < N/2
 * (src/fft_c7x/FFTLIB_ifft1dBatched_i32fc_c32fc_o32f/..._tw_gen.cpp:81-95). */
void ifft_c2r_split_factor_gen(FFTLIB_F32 *pSFactor, int32_t numPoints);
/* N/2-point mixed-radix twiddle table (..._tw_gen.cpp:9-79). */
void ifft_c2r_tw_gen(FFTLIB_F32 *pW, int32_t numPoints);

#ifdef __cplusplus
}
#endif
#endif
