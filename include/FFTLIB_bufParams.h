/*
 * Buffer descriptors and element-type enumeration.
 *
 * Layout and enumerator values are ABI and follow
 * src/common/FFTLIB_bufParams.h:22-152 of TexasInstruments/c7x_fftlib.
 * dim_x is counted in ELEMENTS (floats or int16s), so an interleaved complex
 * buffer of n points has dim_x == 2*n.
 */
#ifndef FFTLIB_BUFPARAMS_H_
#define FFTLIB_BUFPARAMS_H_ 1

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum _FFTLIB_data_type_e {
   FFTLIB_INT8    = 0,
   FFTLIB_INT16   = 1,
   FFTLIB_INT32   = 2,
   FFTLIB_INT64   = 3,
   FFTLIB_INT128  = 4,
   FFTLIB_UINT8   = 5,
   FFTLIB_UINT16  = 6,
   FFTLIB_UINT24  = 7,
   FFTLIB_UINT32  = 8,
   FFTLIB_UINT64  = 9,
   FFTLIB_UINT128 = 10,
   FFTLIB_FLOAT16 = 11,
   FFTLIB_FLOAT32 = 12,
   FFTLIB_FLOAT64 = 13
} FFTLIB_data_type_e;

#define FFTLIB_ISSIGNED(type) ((type) < FFTLIB_UINT8 || (type) > FFTLIB_UINT128)
#define FFTLIB_SIGNEDNESS(type) (FFTLIB_ISSIGNED((type)) ? "signed" : "unsigned")

typedef enum _FFTLIB_dimension_e {
   FFTLIB_2D = 0,
   FFTLIB_3D = 1,
   FFTLIB_4D = 2,
   FFTLIB_5D = 3,
   FFTLIB_6D = 4
} FFTLIB_dimension_e;

typedef struct {
   uint32_t data_type; /* FFTLIB_data_type_e */
   uint32_t dim_x;     /* width in elements */
} FFTLIB_bufParams1D_t;

typedef struct {
   uint32_t data_type;
   uint32_t dim_x;
   uint32_t dim_y;
   int32_t  stride_y; /* bytes */
} FFTLIB_bufParams2D_t;

typedef struct {
   uint32_t data_type;
   uint32_t dim_x;
   uint32_t dim_y;
   int32_t  stride_y;
   uint32_t dim_z;
   int32_t  stride_z;
} FFTLIB_bufParams3D_t;

typedef struct {
   uint32_t data_type;
   uint32_t dim_0;
   uint32_t dim_1;
   int32_t  stride_1;
   uint32_t dim_2;
   int32_t  stride_2;
   uint32_t dim_3;
   int32_t  stride_3;
   uint32_t dim_4;
   int32_t  stride_4;
   uint32_t dim_5;
   int32_t  stride_5;
   uint32_t dimension; /* FFTLIB_dimension_e */
} FFTLIB_bufParamsND_t;

/* Bytes per element of a FFTLIB_data_type_e; unknown types count as 1 byte. */
static inline int32_t FFTLIB_sizeof(uint32_t type)
{
   switch (type) {
   case FFTLIB_INT16:
   case FFTLIB_UINT16:
   case FFTLIB_FLOAT16: return 2;
   case FFTLIB_UINT24: return 3;
   case FFTLIB_INT32:
   case FFTLIB_UINT32:
   case FFTLIB_FLOAT32: return 4;
   case FFTLIB_INT64:
   case FFTLIB_UINT64:
   case FFTLIB_FLOAT64: return 8;
   case FFTLIB_INT128:
   case FFTLIB_UINT128: return 16;
   default: return 1;
   }
}

#ifdef __cplusplus
}
#endif

#endif /* FFTLIB_BUFPARAMS_H_ */
