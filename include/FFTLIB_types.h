/*
 * FFTLIB common ABI types for the MI355X (gfx950) build of the FFTLIB batched
 * 1-D FFT hot path.
 *
 * Every value and type below is part of the binary interface a caller of the
 * reference library compiles against, so names and numeric values follow
 * src/common/FFTLIB_types.h:164-221 of TexasInstruments/c7x_fftlib exactly.
 * The C7x intrinsic headers the reference pulls in (<c7x.h>) have no role on
 * a GPU and are not included.
 */
#ifndef FFTLIB_TYPES_H_
#define FFTLIB_TYPES_H_ 1

#include <inttypes.h>
#include <stddef.h>
#include <stdint.h>

#include "FFTLIB_bufParams.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Floating-point element types (src/common/FFTLIB_types.h:164-165). */
typedef double FFTLIB_D64;
typedef float  FFTLIB_F32;

/* Status codes (src/common/FFTLIB_types.h:168-209). Values are ABI. */
typedef enum {
   FFTLIB_SUCCESS                      = 0,
   FFTLIB_ERR_FAILURE                  = 1,
   FFTLIB_ERR_INVALID_VALUE            = 2,
   FFTLIB_ERR_INVALID_TYPE             = 3,
   FFTLIB_ERR_INVALID_DIMENSION        = 4,
   FFTLIB_ERR_NULL_POINTER             = 5,
   FFTLIB_ERR_NOT_IMPLEMENTED          = 6,
   FFTLIB_ERR_NOT_EQUAL_WIDTH_STRIDE   = 7,
   FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES = 8,
   FFTLIB_ERR_NOT_ALIGNED_WIDTHS       = 9,
   FFTLIB_ERR_BUFFER_TOO_SMALL         = 10,
   FFTLIB_ERR_INVALID_ACTIVATION       = 11,
   FFTLIB_ERR_INVALID_SHIFT            = 12,
   FFTLIB_ERROR_MAX                    = 13,
   FFTLIB_ERR_INVALID_OUTMODE          = 14,
   FFTLIB_ERR_MATRIX_MATRIX_MULTIPLY_UNSUPPORTED_INPUT_TYPE                  = 10000,
   FFTLIB_ERR_MATRIX_MATRIX_MULTIPLY_UNSUPPORTED_OUTPUT_TYPE                 = 10001,
   FFTLIB_ERR_MATRIX_MATRIX_MULTIPLY_UNSUPPORTED_TYPE_COMBINATION            = 10002,
   FFTLIB_ERR_MATRIX_MATRIX_MULTIPLY_ACCUMULATE_UNSUPPORTED_INPUT_TYPE       = 11000,
   FFTLIB_ERR_MATRIX_MATRIX_MULTIPLY_ACCUMULATE_UNSUPPORTED_OUTPUT_TYPE      = 11001,
   FFTLIB_ERR_MATRIX_MATRIX_MULTIPLY_ACCUMULATE_UNSUPPORTED_TYPE_COMBINATION = 11002,
   FFTLIB_ERR_MATRIX_TRANSPOSE_UNSUPPORTED_INPUT_TYPE                        = 12000,
   FFTLIB_ERR_MATRIX_TRANSPOSE_UNSUPPORTED_OUTPUT_TYPE                       = 12001,
   FFTLIB_ERR_MATRIX_TRANSPOSE_UNSUPPORTED_TYPE_COMBINATION                  = 12002,
   FFTLIB_ERR_POINTWISE_MATRIX_MATRIX_MULTIPLY_UNSUPPORTED_INPUT_TYPE        = 13000,
   FFTLIB_ERR_POINTWISE_MATRIX_MATRIX_MULTIPLY_UNSUPPORTED_OUTPUT_TYPE       = 13001,
   FFTLIB_ERR_POINTWISE_MATRIX_MATRIX_MULTIPLY_UNSUPPORTED_TYPE_COMBINATION  = 13002
} FFTLIB_STATUS_NAME;

typedef FFTLIB_STATUS_NAME FFTLIB_STATUS;

/* Opaque handle of the MMA-style coordinator API (src/common/FFTLIB_types.h:213). */
typedef void *FFTLIB_kernelHandle;

/*
 * Function style (src/common/FFTLIB_types.h:216-221). NATC and OPTIMIZED keep
 * their reference values; FFTLIB_FUNCTION_GPU is the one addition of this
 * build. On this platform every computing style runs on the GPU:
 *   FFTLIB_FUNCTION_OPTIMIZED  buffers are ordinary host memory; the library
 *                              stages them through device memory (default, so
 *                              unmodified host callers keep working).
 *   FFTLIB_FUNCTION_GPU        buffers are device memory (hipMalloc, managed
 *                              or hipHostMalloc'd); kernels run in place on
 *                              the configured HIP stream with no copies.
 *   FFTLIB_FUNCTION_NATC       has no implementation in the product library
 *                              (the natural-C kernels are a test oracle); the
 *                              compute entry points return
 *                              FFTLIB_ERR_NOT_IMPLEMENTED for it.
 */
typedef enum {
   FFTLIB_FUNCTION_NATC      = 0,
   FFTLIB_FUNCTION_OPTIMIZED = 1,
   FFTLIB_FUNCTION_GPU       = 2,
   FFTLIB_FUNCTION_MAX       = 128
} FFTLIB_FUNCTION_STYLE;

/*
 * Size in bytes of one C7x streaming-engine parameter block. The reference
 * takes it from TI's <c7x.h>; the PBLOCK_SIZE macros of the kernel headers are
 * expressed in it, so callers that size pBlock arrays with those macros get
 * the same byte counts as on C7x (64-byte SE templates). The GPU plan stored
 * in pBlock is far smaller than the smallest of them.
 */
#ifndef SE_PARAM_SIZE
#define SE_PARAM_SIZE 64
#endif

#ifdef __cplusplus
#define FFTLIB_RESTRICT __restrict__
#else
#define FFTLIB_RESTRICT restrict
#endif

#ifdef __cplusplus
}
#endif

#endif /* FFTLIB_TYPES_H_ */
