/*
 * GPU controls of the MI355X build: the function-style selector, the HIP
 * stream and completion mode used by every _kernel/_exec entry point, and the
 * twiddle generators the reference keeps in its test drivers.
 *
 * None of these has a counterpart function in the reference library; the
 * selector extends FFTLIB_FUNCTION_STYLE (src/common/FFTLIB_types.h:216-221),
 * which in the reference only the MMA kernels' InitArgs consult.
 */
#ifndef FFTLIB_GPU_H_
#define FFTLIB_GPU_H_ 1

#include "FFTLIB_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Style recorded by every subsequent fft_c7x _init call (process wide;
 * default FFTLIB_FUNCTION_OPTIMIZED). Returns FFTLIB_ERR_INVALID_VALUE for an
 * unknown style. */
FFTLIB_STATUS FFTLIB_setFunctionStyle(FFTLIB_FUNCTION_STYLE style);
FFTLIB_FUNCTION_STYLE FFTLIB_getFunctionStyle(void);

/* HIP stream (a hipStream_t, NULL = the null stream) on which kernels are
 * enqueued. */
FFTLIB_STATUS FFTLIB_GPU_setStream(void *hipStream);
void *FFTLIB_GPU_getStream(void);

/* 1 (default): _kernel returns after the GPU finished, like the reference's
 * synchronous API. 0: device-pointer (FFTLIB_FUNCTION_GPU) calls return right
 * after enqueueing; the caller synchronizes the stream. Host-pointer calls
 * always complete before returning. */
FFTLIB_STATUS FFTLIB_GPU_setSynchronous(int32_t synchronous);

/* 1 (default): complex FFT lengths with a compile-time specialised kernel use
 * it; 0: every length runs the generic (runtime stage table) kernels. Both
 * give results within fp32 rounding of each other; the switch exists for
 * tuning and for testing the generic kernels at every length. */
FFTLIB_STATUS FFTLIB_GPU_setSpecializedKernels(int32_t enable);

/* Arithmetic of the fp32 complex FFT kernels (batched and single, forward and
 * inverse) for lengths that have a DFT-as-GEMM matrix-core kernel (N = 960);
 * other lengths always use FFTLIB_C2C_VARIANT_FP32.
 *   FFTLIB_C2C_VARIANT_FP32   (0, default) mixed-radix fp32, the reference's
 *                             algorithm and twiddle table;
 *   FFTLIB_C2C_VARIANT_BF16X3 (1) radix-15 and radix-64 DFT stages as MFMA
 *                             GEMMs on bf16 hi+mid+lo operand parts (6
 *                             products), FP32 accumulation;
 *   FFTLIB_C2C_VARIANT_FP16X2 (2) same with fp16 hi+lo parts (3 products);
 *   FFTLIB_C2C_VARIANT_BF16X2 (3) same with bf16 hi+lo parts (3 products).
 * The MFMA variants generate their own DFT matrices and do not read pW.
 * Process wide; returns FFTLIB_ERR_INVALID_VALUE for an unknown value. */
#define FFTLIB_C2C_VARIANT_FP32   0
#define FFTLIB_C2C_VARIANT_BF16X3 1
#define FFTLIB_C2C_VARIANT_FP16X2 2
#define FFTLIB_C2C_VARIANT_BF16X2 3
FFTLIB_STATUS FFTLIB_GPU_setC2CVariant(int32_t variant);
int32_t       FFTLIB_GPU_getC2CVariant(void);

/* Waits for all work the library enqueued on its stream. */
FFTLIB_STATUS FFTLIB_GPU_synchronize(void);

/* Text of the last HIP failure seen by the calling thread ("" if none). */
const char *FFTLIB_GPU_getLastError(void);

/* Library version string. */
const char *FFTLIB_GPU_version(void);

/* Mixed-radix fp32 C2C twiddle table of 2*numPoints floats, the layout every
 * fp32 C2C kernel reads (test/fft_c7x/FFTLIB_fft1d_i32fc_c32fc_o32fc/
 * FFTLIB_fft1d_i32fc_c32fc_o32fc_d.c:31-97, where it is a test-driver helper). */
void FFTLIB_tw_gen_c2c(FFTLIB_F32 *pW, int32_t numPoints);

/* Radix-4 int16 Q15 twiddle table of 2*numPoints int16 for the int16 kernels
 * (test/fft_c7x/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/
 * FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_d.c:29-61). */
void FFTLIB_tw_gen_i16sc(int16_t *pW, uint32_t numPoints);

#ifdef __cplusplus
}
#endif

#endif
