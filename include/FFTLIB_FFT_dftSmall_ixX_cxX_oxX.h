/*
 * DFT of fftSize <= MMA/2 points as one matrix product per channel:
 * Y[l][j] = round(sum_k X[l][k] W[k*MMA + j] >> shiftVector[0]) over the
 * 2*fftSize real values of each channel, W the MMA x MMA block-diagonal
 * real-form DFT matrix of twGen (natural C: ..._cn.cpp:19-46).
 *
 * Drop-in for src/fft_c7xmma/FFTLIB_FFT_dftSmall_ixX_cxX_oxX/FFTLIB_FFT_dftSmall_ixX_cxX_oxX.h:35-269: same InitArgs layout and the same seven functions.
 * Data are complex INT16 or INT32 values, interleaved (R0 I0 R1 I1 ...) or
 * not (R0 R1 ... I0 I1 ...) per channel; channels follow each other.
 * InitArgs.funcStyle selects FFTLIB_FUNCTION_GPU (buffers in device memory,
 * used in place on the library stream) or FFTLIB_FUNCTION_OPTIMIZED (host
 * buffers staged through device memory); FFTLIB_FUNCTION_NATC has no
 * implementation here (init returns FFTLIB_ERR_NOT_IMPLEMENTED). twGen
 * writes host memory. The handle is caller-allocated host memory of
 * getHandleSize bytes.
 */
#ifndef FFTLIB_FFT_DFTSMALL_IXX_CXX_OXX_H_
#define FFTLIB_FFT_DFTSMALL_IXX_CXX_OXX_H_ 1

#include "FFTLIB_types.h"

#define FFTLIB_FFT_DFTSMALL_IXX_CXX_OXX_NUMSHIFTS 1

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
   int8_t   funcStyle;  /* FFTLIB_FUNCTION_STYLE */
   uint32_t fftSize;    /* complex points per channel */
   uint32_t batchSize;  /* channels */
   uint32_t interleave; /* 1: R0 I0 R1 I1 ..., 0: R0 R1 ... I0 I1 ... */
   uint32_t shiftVector[FFTLIB_FFT_DFTSMALL_IXX_CXX_OXX_NUMSHIFTS];
} FFTLIB_FFT_dftSmall_ixX_cxX_oxX_InitArgs;

int32_t FFTLIB_FFT_dftSmall_ixX_cxX_oxX_getHandleSize(FFTLIB_FFT_dftSmall_ixX_cxX_oxX_InitArgs *pKerInitArgs);

FFTLIB_STATUS FFTLIB_FFT_dftSmall_ixX_cxX_oxX_getSizes(const FFTLIB_FFT_dftSmall_ixX_cxX_oxX_InitArgs *pKerInitArgs, FFTLIB_bufParams1D_t *bufParamsX,
      FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_bufParams1D_t *bufParamsY);

FFTLIB_STATUS FFTLIB_FFT_dftSmall_ixX_cxX_oxX_twGen(const FFTLIB_FFT_dftSmall_ixX_cxX_oxX_InitArgs *pKerInitArgs, void *pW, const FFTLIB_bufParams1D_t *bufParamsW);

FFTLIB_STATUS FFTLIB_FFT_dftSmall_ixX_cxX_oxX_init(FFTLIB_kernelHandle handle, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_bufParams1D_t *bufParamsW,
      FFTLIB_bufParams1D_t *bufParamsY, const FFTLIB_FFT_dftSmall_ixX_cxX_oxX_InitArgs *pKerInitArgs);

FFTLIB_STATUS FFTLIB_FFT_dftSmall_ixX_cxX_oxX_init_checkParams(FFTLIB_kernelHandle handle, const FFTLIB_bufParams1D_t *bufParamsX,
      const FFTLIB_bufParams1D_t *bufParamsW, const FFTLIB_bufParams1D_t *bufParamsY,
      const FFTLIB_FFT_dftSmall_ixX_cxX_oxX_InitArgs *pKerInitArgs);

FFTLIB_STATUS FFTLIB_FFT_dftSmall_ixX_cxX_oxX_exec(FFTLIB_kernelHandle handle, void *pX, void *pW, void *pY);

FFTLIB_STATUS FFTLIB_FFT_dftSmall_ixX_cxX_oxX_exec_checkParams(FFTLIB_kernelHandle handle, const void *pX, const void *pW, const void *pY);

#ifdef __cplusplus
}
#endif

#endif
