// C ABI of the fixed-point DFT kernels (the reference's src/fft_c7xmma):
// FFTLIB_FFT_fft_ixX_cxX_oxX_* (the coordinator) and the three kernels it
// chooses from, FFTLIB_FFT_dftSmall/dftLarge/highRadixDecompositions_ixX_cxX_oxX_*,
// with the reference's sizes, table layouts, parameter checks and status
// codes; exec runs the gfx950 kernels of fft_mma.hip.
//
// References (paths relative to the reference root, $MMA = src/fft_c7xmma):
//   kernel choice        $MMA/FFTLIB_FFT_fft_ixX_cxX_oxX/FFTLIB_FFT_fft_ixX_cxX_oxX.cpp:13-49
//   dftSmall sizes/table $MMA/FFTLIB_FFT_dftSmall_ixX_cxX_oxX/c71/..._ci.cpp:73-207, checks ..._ixX_cxX_oxX.cpp:13-75
//   dftLarge sizes/table $MMA/FFTLIB_FFT_dftLarge_ixX_cxX_oxX/c71/..._ci.cpp:43-156, checks ..._ixX_cxX_oxX.cpp:13-89
//                        (+ src/linalg_c7xmma/FFTLIB_LINALG_matrixMatrixMultiply_ixX_ixX_oxX/..._ixX_ixX_oxX.cpp:15-62)
//   highRadix sizes/tables $MMA/FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX/c71/..._ci.cpp:1975-2168,
//                        checks ..._ixX_cxX_oxX.cpp:14-84
// The reference's highRadix twGen also fills a library-global table of
// decomposition matrices; here each exec uses a device copy of the matrices
// of its own size (built on first use per device and kept).
#include <math.h>
#include <stddef.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_mma.h"
#include "fftlib_fft.h"
#include "fftlib_runtime.h"

using namespace fftlib;

namespace {

constexpr uint32_t kMmaMagic = 0x464d4d41u;  // "AMMF"

enum MmaKind : int32_t { MK_FFT = 0, MK_SMALL = 1, MK_LARGE = 2, MK_HIGH = 3 };
enum MmaVariant : int32_t { MV_SMALL = 0, MV_LARGE = 1, MV_HIGH = 2 };

// All InitArgs types share this layout (the coordinator passes its own to the
// kernel it chooses, as the reference does).
struct Args {
   int8_t   funcStyle;
   uint32_t fftSize;
   uint32_t batchSize;
   uint32_t interleave;
   uint32_t shiftVector[5];
};

struct MmaHandle {
   uint32_t magic;
   int32_t  kind;
   int32_t  variant;
   int32_t  dtype;  // kMmaInt16 / kMmaInt32
   int32_t  style;
   uint32_t fft, batch, interleave;
   uint32_t shifts[5];
};

Args args_of(const void *p, int nShifts)
{
   Args a{};
   memcpy(&a, p, offsetof(Args, shiftVector) + sizeof(uint32_t) * (size_t) nShifts);
   return a;
}

int nshifts(int32_t kind) { return (kind == MK_SMALL || kind == MK_LARGE) ? 1 : 5; }

uint32_t mma_of(uint32_t dataType) { return dataType == FFTLIB_INT16 ? 32u : 16u; }
bool     int_type(uint32_t dataType) { return dataType == FFTLIB_INT16 || dataType == FFTLIB_INT32; }

int32_t choose(const Args &a, uint32_t mma)
{
   if (a.fftSize <= mma / 2) return MV_SMALL;
   if (a.fftSize == mma) return MV_LARGE;
   if (a.fftSize == mma * 2) {
      if (a.batchSize < mma / 2) return MV_HIGH;
      if (a.batchSize == mma / 2) return a.interleave ? MV_HIGH : MV_LARGE;
      return MV_LARGE;
   }
   return MV_HIGH;
}

uint32_t stage0_factor(uint32_t fft, uint32_t mma)
{
   uint32_t f = fft;
   while (f > mma / 2) f /= mma / 2;
   return f;
}

uint32_t hr_wsize(uint32_t fft, uint32_t mma)
{
   const uint32_t d0 = stage0_factor(fft, mma);
   uint32_t       w = 2 * fft, d = fft / d0;
   while (d > mma / 2) {
      w += 2 * d;
      d /= mma / 2;
   }
   return w + mma * mma;
}

// Sizes of one kernel's buffers for data type dt; a kernel that cannot take
// the size returns FFTLIB_ERR_INVALID_DIMENSION and leaves the sizes alone.
FFTLIB_STATUS sizes(int32_t variant, const Args &a, uint32_t dt, FFTLIB_bufParams1D_t *x, FFTLIB_bufParams1D_t *w,
                    FFTLIB_bufParams1D_t *y)
{
   if (!int_type(dt)) return FFTLIB_ERR_NOT_IMPLEMENTED;
   const uint32_t mma = mma_of(dt);
   uint32_t       wn  = 0;
   if (variant == MV_SMALL) {
      if (a.fftSize > (mma >> 1)) return FFTLIB_ERR_INVALID_DIMENSION;
      wn = mma * mma;
   }
   else if (variant == MV_LARGE) {
      if (a.fftSize <= mma / 2) return FFTLIB_ERR_INVALID_DIMENSION;
      wn = a.fftSize * 2 * a.fftSize * 2;
   }
   else {
      wn = hr_wsize(a.fftSize, mma);
   }
   x->dim_x = a.fftSize * a.batchSize * 2;
   y->dim_x = a.fftSize * a.batchSize * 2;
   w->dim_x = wn;
   return FFTLIB_SUCCESS;
}

// ---- table generation (host, in the reference's arithmetic) ----------------
template <typename T>
T conv(double x)
{
   const int64_t maxv = ((int64_t) 1 << (sizeof(T) * 8 - 1)) - 1;
   const int64_t v    = (int64_t) floor(0.5 + x);
   return (T) (v >= maxv ? maxv : v <= -maxv - 1 ? -maxv - 1 : v);
}
template <typename T>
double tw_scale()
{
   return (double) (((uint32_t) 1 << (sizeof(T) * 8 - 2)) - 1) + (double) ((uint32_t) 1 << (sizeof(T) * 8 - 2)) + 0.5;
}
const double kPi = 3.141592654;

template <typename T>
void gen_small(T *w, const Args &a)
{
   const uint32_t mma = sizeof(T) == 2 ? 32u : 16u, N = a.fftSize, bpr = (mma >> 1) / N;
   const double   sc  = tw_scale<T>();
   memset(w, 0, sizeof(T) * mma * mma);
   for (uint32_t j = 0; j < bpr; j++)
      for (uint32_t k = 0; k < N; k++)
         for (uint32_t n = 0; n < N; n++) {
            const double ang = 2 * kPi * k * n / N;
            const T      c = conv<T>(sc * cos(ang)), s = conv<T>(sc * sin(ang));
            if (a.interleave) {  // block j of the diagonal
               const uint32_t idx = (j * N + k) * 2 * mma + j * N * 2;
               w[idx + n * 2]           = c;
               w[idx + n * 2 + 1]       = (T) -s;
               w[idx + mma + n * 2]     = s;
               w[idx + mma + n * 2 + 1] = c;
            }
            else if (j < a.batchSize) {
               const uint32_t idx = (j * N * 2 + k) * mma + j * N * 2;
               w[idx + n]                 = c;
               w[idx + N + n]             = (T) -s;
               w[idx + N * mma + n]       = s;
               w[idx + N * mma + N + n]   = c;
            }
         }
}

template <typename T>
void gen_large(T *w, const Args &a)
{
   const uint32_t N  = a.fftSize;
   const double   sc = tw_scale<T>();
   for (uint32_t j = 0; j < N; j++)
      for (uint32_t k = 0; k < N; k++) {
         const double ang = 2 * kPi * j * k / N;
         const T      c = conv<T>(sc * cos(ang)), s = conv<T>(sc * sin(ang));
         if (a.interleave) {
            w[j * 2 * N * 2 + k * 2]           = c;
            w[j * 2 * N * 2 + k * 2 + 1]       = (T) -s;
            w[(j * 2 + 1) * N * 2 + k * 2]     = s;
            w[(j * 2 + 1) * N * 2 + k * 2 + 1] = c;
         }
         else {
            w[j * N * 2 + k]                 = c;
            w[j * N * 2 + N + k]             = (T) -s;
            w[N * N * 2 + j * N * 2 + k]     = s;
            w[N * N * 2 + j * N * 2 + N + k] = c;
         }
      }
}

// Caller table of highRadixDecompositions: per-stage twiddles, then the final
// MMA/2-point DFT matrix.
template <typename T>
void gen_high(T *p, uint32_t fft)
{
   const uint32_t mma = sizeof(T) == 2 ? 32u : 16u, d0 = stage0_factor(fft, mma);
   const double   sc  = tw_scale<T>();
   uint32_t       d   = fft;
   for (uint32_t j = 0; j < (d * 2) / (mma * d0); j++)
      for (uint32_t l = 0; l < d0; l++)
         for (uint32_t k = 0; k < mma / 2; k++) {
            const double   ang = (2 * kPi * (j * mma / 2 + k) * l) / d;
            const uint32_t i   = j * mma * d0 + l * mma + k * 2;
            p[i + 1]           = conv<T>(sc * cos(ang));
            p[i]               = (T) -conv<T>(sc * sin(ang));
         }
   p += d * 2;
   d /= d0;
   while (d > mma / 2) {
      for (uint32_t j = 0; j < (d * 4) / (mma * mma); j++)
         for (uint32_t l = 0; l < mma / 2; l++)
            for (uint32_t k = 0; k < mma / 2; k++) {
               const double   ang = (2 * kPi * (j * mma / 2 + k) * l) / d;
               const uint32_t i   = j * mma * mma / 2 + l * mma + k * 2;
               p[i + 1]           = conv<T>(sc * cos(ang));
               p[i]               = (T) -conv<T>(sc * sin(ang));
            }
      p += d * 2;
      d /= mma / 2;
   }
   for (uint32_t k = 0; k < mma / 2; k++)
      for (uint32_t l = 0; l < mma / 2; l++) {
         const double ang = (2 * kPi * k * l * 2) / mma;
         const T      c = conv<T>(sc * cos(ang)), s = conv<T>(sc * sin(ang));
         p[k * 2 * mma + l * 2]           = s;
         p[k * 2 * mma + l * 2 + 1]       = c;
         p[(k * 2 + 1) * mma + l * 2]     = c;
         p[(k * 2 + 1) * mma + l * 2 + 1] = (T) -s;
      }
}

// Decomposition matrices of highRadixDecompositions (mma x mma elements): the
// block-diagonal stage-0 DFT of fft's first factor, then the MMA/2-point DFT.
template <typename T>
std::vector<T> gen_high_w1(uint32_t fft)
{
   const uint32_t mma = sizeof(T) == 2 ? 32u : 16u, d0 = stage0_factor(fft, mma);
   const double   sc  = tw_scale<T>();
   std::vector<T> w(mma * mma, (T) 0);
   for (uint32_t j = 0; j < mma / (2 * d0); j++)
      for (uint32_t k = 0; k < d0; k++)
         for (uint32_t l = 0; l < d0; l++) {
            const double   ang = (2 * kPi * k * l) / d0;
            const uint32_t i   = j * d0 * mma + j * d0 * 2 + k * mma + l * 2;
            w[i]               = conv<T>(sc * cos(ang));
            w[i + 1]           = (T) -conv<T>(sc * sin(ang));
         }
   T *b = w.data() + mma * mma / 2;
   for (uint32_t k = 0; k < mma / 2; k++)
      for (uint32_t l = 0; l < mma / 2; l++) {
         const double ang     = (2 * kPi * k * l * 2) / mma;
         b[k * mma + l * 2]     = conv<T>(sc * cos(ang));
         b[k * mma + l * 2 + 1] = (T) -conv<T>(sc * sin(ang));
      }
   return w;
}

std::mutex                                   g_w1Mu;
std::map<std::tuple<int, int, uint32_t>, void *> g_w1;   // (device, dtype, fft) -> device matrices

hipError_t device_w1(int dtype, uint32_t fft, const void **out)
{
   int        dev = 0;
   hipError_t e   = hipGetDevice(&dev);
   if (e != hipSuccess) return e;
   std::lock_guard<std::mutex> lock(g_w1Mu);
   auto                        key = std::make_tuple(dev, dtype, fft);
   auto                        it  = g_w1.find(key);
   if (it != g_w1.end()) {
      *out = it->second;
      return hipSuccess;
   }
   void  *d     = nullptr;
   size_t bytes = 0;
   const void *src = nullptr;
   std::vector<int16_t> w16;
   std::vector<int32_t> w32;
   if (dtype == kMmaInt16) {
      w16 = gen_high_w1<int16_t>(fft), src = w16.data(), bytes = w16.size() * sizeof(int16_t);
   }
   else {
      w32 = gen_high_w1<int32_t>(fft), src = w32.data(), bytes = w32.size() * sizeof(int32_t);
   }
   if ((e = hipMalloc(&d, bytes)) != hipSuccess) return e;
   if ((e = hipMemcpy(d, src, bytes, hipMemcpyHostToDevice)) != hipSuccess) {
      (void) hipFree(d);
      return e;
   }
   g_w1[key] = d;
   *out      = d;
   return hipSuccess;
}

// ---- entry-point bodies ----------------------------------------------------
FFTLIB_STATUS do_get_sizes(int32_t kind, const void *initArgs, FFTLIB_bufParams1D_t *x, FFTLIB_bufParams1D_t *w,
                           FFTLIB_bufParams1D_t *y)
{
   if (initArgs == nullptr || x == nullptr || w == nullptr || y == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const Args a = args_of(initArgs, nshifts(kind));
   if (!int_type(x->data_type)) return FFTLIB_ERR_NOT_IMPLEMENTED;
   const int32_t v = kind == MK_FFT ? choose(a, mma_of(x->data_type)) : kind - 1;
   return sizes(v, a, x->data_type, x, w, y);
}

FFTLIB_STATUS do_tw_gen(int32_t kind, const void *initArgs, void *pW, const FFTLIB_bufParams1D_t *bw)
{
   if (initArgs == nullptr || pW == nullptr || bw == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const Args a = args_of(initArgs, nshifts(kind));
   if (!int_type(bw->data_type)) return FFTLIB_ERR_NOT_IMPLEMENTED;
   const int32_t        v = kind == MK_FFT ? choose(a, mma_of(bw->data_type)) : kind - 1;
   FFTLIB_bufParams1D_t lx{bw->data_type, 0}, lw{bw->data_type, 0}, ly{bw->data_type, 0};
   FFTLIB_STATUS        st = sizes(v, a, bw->data_type, &lx, &lw, &ly);
   if (st != FFTLIB_SUCCESS) return st;
   if (bw->dim_x != lw.dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bw->data_type == FFTLIB_INT16) {
      if (v == MV_SMALL) gen_small((int16_t *) pW, a);
      else if (v == MV_LARGE) gen_large((int16_t *) pW, a);
      else gen_high((int16_t *) pW, a.fftSize);
   }
   else {
      if (v == MV_SMALL) gen_small((int32_t *) pW, a);
      else if (v == MV_LARGE) gen_large((int32_t *) pW, a);
      else gen_high((int32_t *) pW, a.fftSize);
   }
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS do_init_check(int32_t kind, FFTLIB_kernelHandle handle, const FFTLIB_bufParams1D_t *bx,
                            const FFTLIB_bufParams1D_t *bw, const FFTLIB_bufParams1D_t *by, const void *initArgs)
{
   if (handle == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (bx == nullptr || bw == nullptr || by == nullptr || initArgs == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const Args a = args_of(initArgs, nshifts(kind));
   if (!int_type(bx->data_type)) return FFTLIB_ERR_NOT_IMPLEMENTED;
   const uint32_t       mma = mma_of(bx->data_type);
   const int32_t        v   = kind == MK_FFT ? choose(a, mma) : kind - 1;
   FFTLIB_bufParams1D_t lx{bx->data_type, 0}, lw{bx->data_type, 0}, ly{bx->data_type, 0};
   FFTLIB_STATUS        st = sizes(v, a, bx->data_type, &lx, &lw, &ly);
   if (st != FFTLIB_SUCCESS) return st;
   if (v == MV_HIGH && a.fftSize < mma * 2) return FFTLIB_ERR_INVALID_VALUE;
   if (bx->dim_x != lx.dim_x || by->dim_x != ly.dim_x || bw->dim_x != lw.dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bx->data_type != bw->data_type || bx->data_type != by->data_type) return FFTLIB_ERR_INVALID_TYPE;
   if (v == MV_LARGE) {
      // the matrix-product kernel's shift check (an int8_t shift against the
      // accumulator width: 64 bits for INT16 data, 128 for INT32)
      const int8_t shift = (int8_t) a.shiftVector[0];
      if (shift > (bx->data_type == FFTLIB_INT16 ? 64 : 128)) return FFTLIB_ERR_INVALID_SHIFT;
   }
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS do_init(int32_t kind, FFTLIB_kernelHandle handle, const FFTLIB_bufParams1D_t *bx,
                      const FFTLIB_bufParams1D_t *bw, const FFTLIB_bufParams1D_t *by, const void *initArgs)
{
   if (handle == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (bx == nullptr || bw == nullptr || by == nullptr || initArgs == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const Args a = args_of(initArgs, nshifts(kind));
   if (!int_type(bx->data_type)) return FFTLIB_ERR_NOT_IMPLEMENTED;
   if (a.funcStyle == FFTLIB_FUNCTION_NATC) {
      set_last_error_text("FFTLIB_FUNCTION_NATC: the natural-C kernels are not part of this library");
      return FFTLIB_ERR_NOT_IMPLEMENTED;
   }
   const uint32_t       mma = mma_of(bx->data_type);
   const int32_t        v   = kind == MK_FFT ? choose(a, mma) : kind - 1;
   FFTLIB_bufParams1D_t lx{bx->data_type, 0}, lw{bx->data_type, 0}, ly{bx->data_type, 0};
   FFTLIB_STATUS        st = sizes(v, a, bx->data_type, &lx, &lw, &ly);
   if (st != FFTLIB_SUCCESS) return st;
   if (v == MV_HIGH) {
      // a power of two of at least 2*MMA points that the first stage and at
      // most three MMA/2-point stages reduce to the final MMA/2-point DFT
      uint32_t rest = a.fftSize / stage0_factor(a.fftSize, mma);
      for (int s = 0; s < 3 && rest > mma / 2; ++s) rest /= mma / 2;
      if (a.fftSize < mma * 2 || (a.fftSize & (a.fftSize - 1)) != 0 || rest > mma / 2) {
         set_last_error_text("highRadixDecompositions: fftSize must be a power of two from 2*MMA up to the "
                             "first factor times (MMA/2)^4");
         return FFTLIB_ERR_INVALID_DIMENSION;
      }
   }
   MmaHandle h{};
   h.magic = kMmaMagic, h.kind = kind, h.variant = v;
   h.dtype      = bx->data_type == FFTLIB_INT16 ? kMmaInt16 : kMmaInt32;
   h.style      = a.funcStyle == FFTLIB_FUNCTION_GPU ? FFTLIB_FUNCTION_GPU : FFTLIB_FUNCTION_OPTIMIZED;
   h.fft        = a.fftSize, h.batch = a.batchSize, h.interleave = a.interleave;
   for (int i = 0; i < 5; ++i) h.shifts[i] = a.shiftVector[i];
   if (kind == MK_SMALL || kind == MK_LARGE)
      for (int i = 1; i < 5; ++i) h.shifts[i] = 0;
   memcpy(handle, &h, sizeof h);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS do_exec(int32_t kind, FFTLIB_kernelHandle handle, void *pX, void *pW, void *pY)
{
   if (handle == nullptr || pX == nullptr || pW == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
   MmaHandle h;
   memcpy(&h, handle, sizeof h);
   if (h.magic != kMmaMagic || h.kind != kind) return FFTLIB_ERR_INVALID_VALUE;
   const size_t   es   = h.dtype == kMmaInt16 ? 2 : 4;
   const uint32_t mma  = h.dtype == kMmaInt16 ? 32u : 16u;
   const size_t   data = (size_t) h.fft * h.batch * 2 * es;
   const size_t   wn   = h.variant == MV_SMALL   ? (size_t) mma * mma
                         : h.variant == MV_LARGE ? (size_t) h.fft * 2 * h.fft * 2
                                                 : (size_t) hr_wsize(h.fft, mma);
   const void *w1 = nullptr;
   if (h.variant == MV_HIGH) {
      const hipError_t e = device_w1(h.dtype, h.fft, &w1);
      if (e != hipSuccess) {
         set_last_error("decomposition matrices", e);
         return FFTLIB_ERR_FAILURE;
      }
   }
   CallBuffers b{};
   b.in  = IoBuffer{pX, data};
   b.tw  = IoBuffer{pW, wn * es};
   b.out = pY, b.outBytes = data;
   return run_buffers((uint32_t) h.style, b, [&](const DeviceBuffers &d, hipStream_t s) -> hipError_t {
      if (h.variant == MV_HIGH)
         return launch_mma_high_radix(h.dtype, const_cast<void *>(d.in), w1, d.tw, d.out, h.fft, h.batch, h.shifts,
                                      h.interleave, s);
      const uint32_t ldw = h.variant == MV_SMALL ? mma : 2 * h.fft;
      return launch_mma_matrix(h.dtype, d.in, d.tw, d.out, h.batch, 2 * h.fft, ldw, h.shifts[0], s);
   });
}

FFTLIB_STATUS do_exec_check(const void *pX, const void *pW, const void *pY)
{
   return (pX == nullptr || pW == nullptr || pY == nullptr) ? FFTLIB_ERR_NULL_POINTER : FFTLIB_SUCCESS;
}

}  // namespace

extern "C" {

#define FFTLIB_MMA_ENTRY_POINTS(K, KIND)                                                                            \
   int32_t K##_getHandleSize(K##_InitArgs *pKerInitArgs)                                                             \
   {                                                                                                                \
      return (int32_t) sizeof(MmaHandle);                                                                           \
   }                                                                                                                \
   FFTLIB_STATUS K##_getSizes(const K##_InitArgs *pKerInitArgs, FFTLIB_bufParams1D_t *bufParamsX,                    \
                              FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_bufParams1D_t *bufParamsY)                    \
   {                                                                                                                \
      return do_get_sizes(KIND, pKerInitArgs, bufParamsX, bufParamsW, bufParamsY);                                  \
   }                                                                                                                \
   FFTLIB_STATUS K##_twGen(const K##_InitArgs *pKerInitArgs, void *pW, const FFTLIB_bufParams1D_t *bufParamsW)       \
   {                                                                                                                \
      return do_tw_gen(KIND, pKerInitArgs, pW, bufParamsW);                                                         \
   }                                                                                                                \
   FFTLIB_STATUS K##_init(FFTLIB_kernelHandle handle, FFTLIB_bufParams1D_t *bufParamsX,                              \
                          FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_bufParams1D_t *bufParamsY,                        \
                          const K##_InitArgs *pKerInitArgs)                                                          \
   {                                                                                                                \
      return do_init(KIND, handle, bufParamsX, bufParamsW, bufParamsY, pKerInitArgs);                               \
   }                                                                                                                \
   FFTLIB_STATUS K##_init_checkParams(FFTLIB_kernelHandle handle, const FFTLIB_bufParams1D_t *bufParamsX,            \
                                      const FFTLIB_bufParams1D_t *bufParamsW, const FFTLIB_bufParams1D_t *bufParamsY,\
                                      const K##_InitArgs *pKerInitArgs)                                              \
   {                                                                                                                \
      return do_init_check(KIND, handle, bufParamsX, bufParamsW, bufParamsY, pKerInitArgs);                         \
   }                                                                                                                \
   FFTLIB_STATUS K##_exec(FFTLIB_kernelHandle handle, void *pX, void *pW, void *pY)                                 \
   {                                                                                                                \
      return do_exec(KIND, handle, pX, pW, pY);                                                                     \
   }                                                                                                                \
   FFTLIB_STATUS K##_exec_checkParams(FFTLIB_kernelHandle handle, const void *pX, const void *pW, const void *pY)  \
   {                                                                                                                \
      return do_exec_check(pX, pW, pY);                                                                             \
   }

FFTLIB_MMA_ENTRY_POINTS(FFTLIB_FFT_fft_ixX_cxX_oxX, MK_FFT)
FFTLIB_MMA_ENTRY_POINTS(FFTLIB_FFT_dftSmall_ixX_cxX_oxX, MK_SMALL)
FFTLIB_MMA_ENTRY_POINTS(FFTLIB_FFT_dftLarge_ixX_cxX_oxX, MK_LARGE)
FFTLIB_MMA_ENTRY_POINTS(FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX, MK_HIGH)

}  // extern "C"
