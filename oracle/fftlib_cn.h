/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the natural-C ("_cn") kernels of FFTLIB's fft_c7x hot
 * path, used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline as
 * the correctness checker. The product library (c7x_fftlib_amd) never links,
 * loads or calls anything in oracle/.
 *
 * Function names are the reference's own _cn names so that the tests read like
 * the reference's test drivers (which compare _kernel against _cn).
 */
#ifndef ORACLE_FFTLIB_CN_H_
#define ORACLE_FFTLIB_CN_H_

#include <stdint.h>

#include "../include/FFTLIB_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fp32 C2C, batched and single channel, forward and inverse. */
FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                       FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                       FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                       uint32_t numPoints, uint32_t numChannels);
FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                   FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                   FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                   uint32_t numPoints, uint32_t numChannels);
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                        FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                        FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                        uint32_t numPoints, uint32_t numChannels);
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                    FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                    FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                    uint32_t numPoints, uint32_t numChannels);
FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                FFTLIB_bufParams1D_t *bufParamsY);
FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY);
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                 FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                 FFTLIB_bufParams1D_t *bufParamsY);
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                             FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                             FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY);

/* int16 fixed point. */
FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                       FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                       FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                       FFTLIB_bufParams1D_t *bufParamsShift, uint32_t numPoints,
                                                       uint32_t numChannels);
FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_checkParams_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                   int16_t *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                   int16_t *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                   uint32_t *pShift,
                                                                   FFTLIB_bufParams1D_t *bufParamsShift,
                                                                   uint32_t numPoints, uint32_t numChannels);
FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                FFTLIB_bufParams1D_t *bufParamsShift);
FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_checkParams_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            int16_t *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            int16_t *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                            uint32_t *pShift, FFTLIB_bufParams1D_t *bufParamsShift);

/* fp32 real-to-complex and complex-to-real. */
FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                      FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                      FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                      FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                      FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                      uint32_t numPoints, uint32_t numChannels, uint8_t outmode);
FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                  FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                  FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                  uint32_t numPoints, uint32_t numChannels,
                                                                  uint8_t outmode);
FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                               FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                               FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                               FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                               FFTLIB_bufParams1D_t *bufParamsY);
FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                           FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                           FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY);
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                       FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                       FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                       FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                       FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                       uint32_t numPoints, uint32_t numChannels, uint8_t inMode);
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                   FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                   FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                   uint32_t numPoints, uint32_t numChannels,
                                                                   uint8_t inMode);
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                                FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                                FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                                FFTLIB_bufParams1D_t *bufParamsY);
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY);

/* Table generators (test drivers and _tw_gen.cpp files of the reference). */
void oracle_tw_gen_c2c(FFTLIB_F32 *pW, int32_t numPoints);
void oracle_tw_gen_i16sc(int16_t *pW, uint32_t numPoints);
void oracle_tw_gen_r2c(FFTLIB_F32 *pW, int32_t numPoints);
void oracle_split_factor_gen_r2c(FFTLIB_F32 *pSf, int32_t numPoints);
void oracle_ifft_c2r_tw_gen(FFTLIB_F32 *pW, int32_t numPoints);
void oracle_ifft_c2r_split_factor_gen(FFTLIB_F32 *pSf, int32_t numPoints);

/* Multi-threaded batch split used only as the all-cores CPU baseline: the
 * channels are partitioned over nThreads POSIX threads, each calling the
 * single-threaded _cn on its contiguous sub-batch. */
FFTLIB_STATUS oracle_fft1dBatched_c2c_threads(FFTLIB_F32 *pX, FFTLIB_F32 *pW, FFTLIB_F32 *pY, uint32_t numPoints,
                                              uint32_t numChannels, int inverse, int nThreads);
FFTLIB_STATUS oracle_fft1dBatched_i16_threads(int16_t *pX, int16_t *pW, int16_t *pY, uint32_t *pShift,
                                              uint32_t numShifts, uint32_t numPoints, uint32_t numChannels,
                                              int nThreads);

#ifdef __cplusplus
}
#endif

#endif
