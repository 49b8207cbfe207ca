// Fixed-point DFT kernels of the fft_c7xmma family on gfx950: INT16 data
// with 64-bit accumulation and INT32 data with 128-bit accumulation, the
// reference's rounding right shifts and saturation, so results are
// bit-identical to its natural-C kernels
// (src/fft_c7xmma/FFTLIB_FFT_dftSmall_ixX_cxX_oxX/..._cn.cpp:19-46,
// src/linalg_c7xmma/FFTLIB_LINALG_matrixMatrixMultiply_ixX_ixX_oxX/..._cn.cpp:27-65,
// src/fft_c7xmma/FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX/..._cn.cpp:19-455).
//
// These kernels are small (the reference's sizes: up to 8192 points, 2048
// channels) and are bound by launch latency rather than by HBM or ALU: each
// is one work item per complex output, computing its dot products with exact
// integer arithmetic (v_mad_i64_i32 for INT16 data, 128-bit adds for INT32).
// The MFMA integer path (int8 operands) would need the data split into four
// byte planes per INT32 value and is not used.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "fft_mma.h"

namespace fftlib {
namespace {

constexpr int kThreads = 256;

template <typename T>
struct Acc;
template <>
struct Acc<int16_t> {
   using type                    = int64_t;
   static constexpr uint32_t mma = 32;  // FFTLIB_MMA_SIZE_16_BIT
   static constexpr uint32_t nm  = 15;  // twiddle products are shifted by the data width minus one
};
template <>
struct Acc<int32_t> {
   using type                    = __int128;
   static constexpr uint32_t mma = 16;  // FFTLIB_MMA_SIZE_32_BIT
   static constexpr uint32_t nm  = 31;
};

// FFTLIB_UTIL_saturate (src/common/FFTLIB_utility.h:518-530 and :546-566).
__device__ __forceinline__ int16_t sat(int64_t x, int16_t) { return x > 0x7FFF ? 0x7FFF : x < -0x8000 ? -0x8000 : (int16_t) x; }
__device__ __forceinline__ int32_t sat(__int128 v, int32_t)
{
   const uint64_t hi = (uint64_t) (v >> 64), lo = (uint64_t) v;
   if (hi & 0x8000000000000000ull) {
      if (hi != 0xFFFFFFFFFFFFFFFFull || (~lo & 0xFFFFFFFF80000000ull) != 0) return INT32_MIN;
      return (int32_t) lo;
   }
   if (lo & 0xFFFFFFFF80000000ull) return INT32_MAX;
   return (int32_t) lo;
}

// FFTLIB_UTIL_shiftAndRound (FFTLIB_utility.h:576-596, :634-656): the shift is
// an 8-bit value; shift 0 saturates only; otherwise ((v >> (s-1)) + 1) >> 1
// with arithmetic shifts (a count beyond the accumulator shifts in sign bits).
template <typename T>
__device__ __forceinline__ T shift_round(typename Acc<T>::type v, uint32_t shift)
{
   using A            = typename Acc<T>::type;
   constexpr uint32_t W = sizeof(A) * 8 - 1;
   shift &= 0xFFu;
   if (shift == 0) return sat(v, T());
   const uint32_t s = shift - 1 > W ? W : shift - 1;
   A              t = (v >> s) + 1;
   return sat(t >> 1, T());
}

// FFTLIB_UTIL_negate (FFTLIB_utility.h:229-242).
__device__ __forceinline__ int16_t neg(int16_t a) { return a == -32768 ? 32767 : (int16_t) -a; }
__device__ __forceinline__ int32_t neg(int32_t a) { return a == INT32_MIN ? INT32_MAX : -a; }

template <typename T>
__device__ __forceinline__ typename Acc<T>::type mul(T a, T b)
{
   return (typename Acc<T>::type) ((int64_t) a * (int64_t) b);
}

// Y[l][j] = shift_round(sum_k X[l][k] W[k*ldw + j]), l < rows, j, k < k2.
template <typename T>
__global__ void __launch_bounds__(kThreads) matrix_kernel(const T *__restrict__ x, const T *__restrict__ w,
                                                          T *__restrict__ y, uint32_t rows, uint32_t k2,
                                                          uint32_t ldw, uint32_t shift)
{
   const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
   if (t >= rows * k2) return;
   const uint32_t l = t / k2, j = t - l * k2;
   const T       *xr = x + (size_t) l * k2;
   typename Acc<T>::type acc = 0;
   for (uint32_t k = 0; k < k2; ++k) acc += mul(xr[k], w[(size_t) ldw * k + j]);
   y[(size_t) l * k2 + j] = shift_round<T>(acc, shift);
}

// Re-layout between non-interleaved (R0..R(N-1) I0..I(N-1)) and interleaved
// channels: toInterleaved 1 reads the former and writes the latter.
template <typename T>
__global__ void __launch_bounds__(kThreads) relayout_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                            uint32_t N, uint32_t B, int toInterleaved)
{
   const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
   if (t >= N * B) return;
   const uint32_t i = t / N, j = t - i * N;
   const size_t   c = (size_t) i * N * 2;
   if (toInterleaved) {
      dst[c + 2 * j]     = src[c + j];
      dst[c + 2 * j + 1] = src[c + N + j];
   }
   else {
      dst[c + j]     = src[c + 2 * j];
      dst[c + N + j] = src[c + 2 * j + 1];
   }
}

// Twiddle product of a stage result (re, im) with table entries (w0, w1),
// stored as the reference does (..._cn.cpp:107-126).
template <typename T>
__device__ __forceinline__ void twiddle_store(T *dst, size_t o, T ra, T ia, T w0, T w1)
{
   constexpr uint32_t nm = Acc<T>::nm;
   dst[o]     = shift_round<T>(mul(w0, ra) + mul(w1, ia), nm);
   dst[o + 1] = shift_round<T>(mul(neg(w0), ia) + mul(w1, ra), nm);
}

// Stage 0 (..._cn.cpp:76-128): the d0-point DFT of points p + m*(N/d0),
// then the twiddle. One work item per (channel j, column group n, k, l).
template <typename T>
__global__ void __launch_bounds__(kThreads) hr_stage0_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                             const T *__restrict__ w1, const T *__restrict__ w2,
                                                             uint32_t N, uint32_t B, uint32_t d0, uint32_t shift)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   using A                = typename Acc<T>::type;
   const uint32_t nGroups = (N * 2) / (MMA * d0);
   const uint32_t t       = blockIdx.x * kThreads + threadIdx.x;
   if (t >= B * nGroups * (MMA / 2) * d0) return;
   uint32_t       r = t;
   const uint32_t l = r % d0;
   r /= d0;
   const uint32_t k = r % (MMA / 2);
   r /= MMA / 2;
   const uint32_t n = r % nGroups;
   const uint32_t j = r / nGroups;
   A              re = 0, im = 0;
   for (uint32_t m = 0; m < d0; ++m) {
      const size_t b  = (size_t) j * N * 2 + (n * MMA + k * 2) + (m * N * 2) / d0;
      const T      c  = w1[l * MMA + m * 2], s = w1[l * MMA + m * 2 + 1];
      im += mul(c, src[b + 1]);
      im += mul(s, src[b]);
      re += mul(c, src[b]);
      re += mul(s, neg(src[b + 1]));
   }
   const T        ra = shift_round<T>(re, shift), ia = shift_round<T>(im, shift);
   const uint32_t wi = n * MMA * d0 + k * 2 + l * MMA;
   twiddle_store(dst, (size_t) j * N * 2 + (n * MMA + k * 2) + (l * N * 2) / d0, ra, ia, w2[wi], w2[wi + 1]);
}

// Stages 1-3 (..._cn.cpp:138-203): the MMA/2-point DFT over points spaced by
// d/(MMA/2) within sub-transforms of d points, then the twiddle.
template <typename T>
__global__ void __launch_bounds__(kThreads) hr_stage_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                            const T *__restrict__ w1, const T *__restrict__ w2,
                                                            uint32_t N, uint32_t B, uint32_t d, uint32_t shift)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   using A                = typename Acc<T>::type;
   const uint32_t nSub    = N / d;
   const uint32_t nL      = (d * 4) / (MMA * MMA);
   const uint32_t t       = blockIdx.x * kThreads + threadIdx.x;
   if (t >= B * nSub * nL * (MMA / 2) * (MMA / 2)) return;
   uint32_t       r = t;
   const uint32_t m = r % (MMA / 2);
   r /= MMA / 2;
   const uint32_t i = r % (MMA / 2);
   r /= MMA / 2;
   const uint32_t l = r % nL;
   r /= nL;
   const uint32_t k = r % nSub;
   const uint32_t j = r / nSub;
   A              re = 0, im = 0;
   for (uint32_t n = 0; n < MMA / 2; ++n) {
      const size_t b = (size_t) j * N * 2 + k * d * 2 + (l * MMA + i * 2) + n * (d * 4) / MMA;
      const T      c = w1[m * MMA + n * 2], s = w1[m * MMA + n * 2 + 1];
      im += mul(c, src[b]);
      im += mul(s, src[b + 1]);
      re += mul(c, src[b + 1]);
      re += mul(s, neg(src[b]));
   }
   const T        ra = shift_round<T>(re, shift), ia = shift_round<T>(im, shift);
   const uint32_t wi = l * MMA * MMA / 2 + i * 2 + m * MMA;
   twiddle_store(dst, (size_t) j * N * 2 + k * d * 2 + l * MMA + i * 2 + m * (d * 4) / MMA, ra, ia, w2[wi],
                 w2[wi + 1]);
}

// Final stage (..._cn.cpp:383-433): MMA/2-point DFTs of consecutive points,
// stored in digit-reversed (natural output) order.
template <typename T>
__global__ void __launch_bounds__(kThreads) hr_dft_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                          const T *__restrict__ wd, uint32_t N, uint32_t B,
                                                          uint32_t d0, uint32_t mC, uint32_t iC, uint32_t hC,
                                                          uint32_t shift)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   using A                = typename Acc<T>::type;
   const uint32_t t       = blockIdx.x * kThreads + threadIdx.x;
   if (t >= B * d0 * mC * iC * hC * (MMA / 2)) return;
   uint32_t       r = t;
   const uint32_t l = r % (MMA / 2);
   r /= MMA / 2;
   const uint32_t h = r % hC;
   r /= hC;
   const uint32_t i = r % iC;
   r /= iC;
   const uint32_t m = r % mC;
   r /= mC;
   const uint32_t k = r % d0;
   const uint32_t j = r / d0;
   A              re = 0, im = 0;
   for (uint32_t n = 0; n < MMA / 2; ++n) {
      const size_t b = (size_t) j * N * 2 + k * (N * 2) / d0 + m * (N * 4) / (d0 * MMA) +
                       i * (N * 8) / (d0 * MMA * MMA) + h * MMA + n * 2;
      re += mul(wd[l * 2 * MMA + n * 2], src[b]);
      re += mul(wd[l * 2 * MMA + n * 2 + 1], src[b + 1]);
      im += mul(wd[(l * 2 + 1) * MMA + n * 2], src[b]);
      im += mul(wd[(l * 2 + 1) * MMA + n * 2 + 1], src[b + 1]);
   }
   const size_t o = (size_t) j * N * 2 + l * d0 * mC * iC * hC * 2 + h * d0 * mC * iC * 2 + i * d0 * mC * 2 +
                    m * d0 * 2 + k * 2;
   dst[o]     = shift_round<T>(re, shift);
   dst[o + 1] = shift_round<T>(im, shift);
}

inline dim3 grid_of(uint64_t items) { return dim3((unsigned) ((items + kThreads - 1) / kThreads)); }

template <typename T>
hipError_t matrix_t(const void *x, const void *w, void *y, uint32_t rows, uint32_t k2, uint32_t ldw, uint32_t shift,
                    hipStream_t s)
{
   if ((uint64_t) rows * k2 == 0) return hipSuccess;
   hipLaunchKernelGGL(matrix_kernel<T>, grid_of((uint64_t) rows * k2), dim3(kThreads), 0, s, (const T *) x,
                      (const T *) w, (T *) y, rows, k2, ldw, shift);
   return hipGetLastError();
}

template <typename T>
hipError_t high_radix_t(void *xv, const void *w1v, const void *w2v, void *yv, uint32_t N, uint32_t B,
                        const uint32_t *shifts, uint32_t interleave, hipStream_t s)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   T *x = (T *) xv, *y = (T *) yv;
   const T *w1 = (const T *) w1v, *w2 = (const T *) w2v;
   hipError_t e;
   T *src = x, *dst = y;
   if (!interleave) {
      hipLaunchKernelGGL(relayout_kernel<T>, grid_of((uint64_t) N * B), dim3(kThreads), 0, s, (const T *) x, y, N, B,
                         1);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      src = y, dst = x;
   }
   uint32_t d0 = N;
   while (d0 > MMA / 2) d0 /= MMA / 2;
   uint32_t d = N, cnt = 0;
   const T *p2 = w2;
   hipLaunchKernelGGL(hr_stage0_kernel<T>, grid_of((uint64_t) N * B), dim3(kThreads), 0, s, (const T *) src, dst, w1,
                      p2, N, B, d0, shifts[0]);
   if ((e = hipGetLastError()) != hipSuccess) return e;
   T *tmp = src;
   src = dst, dst = tmp;
   p2 += d * 2;
   d /= d0;
   ++cnt;
   for (uint32_t st = 1; st <= 3 && d > MMA / 2; ++st) {
      hipLaunchKernelGGL(hr_stage_kernel<T>, grid_of((uint64_t) N * B), dim3(kThreads), 0, s, (const T *) src, dst,
                         w1 + MMA * MMA / 2, p2, N, B, d, shifts[st]);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      tmp = src, src = dst, dst = tmp;
      p2 += d * 2;
      d /= MMA / 2;
      ++cnt;
   }
   const uint32_t q  = (N * 2) / (MMA * d0), r = (N * 4) / (MMA * MMA * d0);
   const uint32_t mC = q > MMA / 2 ? MMA / 2 : q;
   const uint32_t iC = q > MMA / 2 ? (r > MMA / 2 ? MMA / 2 : r) : 1;
   const uint32_t hC = r > MMA / 2 ? (N * 8) / (MMA * MMA * MMA * d0) : 1;
   hipLaunchKernelGGL(hr_dft_kernel<T>, grid_of((uint64_t) B * d0 * mC * iC * hC * (MMA / 2)), dim3(kThreads), 0, s,
                      (const T *) src, dst, p2, N, B, d0, mC, iC, hC, shifts[4]);
   if ((e = hipGetLastError()) != hipSuccess) return e;
   if (!interleave) {
      const T *s2 = (cnt & 1u) ? y : x;
      T       *d2 = (cnt & 1u) ? x : y;
      hipLaunchKernelGGL(relayout_kernel<T>, grid_of((uint64_t) N * B), dim3(kThreads), 0, s, s2, d2, N, B, 0);
      if ((e = hipGetLastError()) != hipSuccess) return e;
   }
   if (cnt & 1u) return hipMemcpyAsync(y, x, sizeof(T) * (size_t) N * B * 2, hipMemcpyDeviceToDevice, s);
   return hipSuccess;
}

}  // namespace

hipError_t launch_mma_matrix(int dtype, const void *x, const void *w, void *y, uint32_t rows, uint32_t k2,
                             uint32_t ldw, uint32_t shift, hipStream_t stream)
{
   return dtype == kMmaInt16 ? matrix_t<int16_t>(x, w, y, rows, k2, ldw, shift, stream)
                             : matrix_t<int32_t>(x, w, y, rows, k2, ldw, shift, stream);
}

hipError_t launch_mma_high_radix(int dtype, void *x, const void *w1, const void *w2, void *y, uint32_t N, uint32_t B,
                                 const uint32_t *shifts, uint32_t interleave, hipStream_t stream)
{
   return dtype == kMmaInt16 ? high_radix_t<int16_t>(x, w1, w2, y, N, B, shifts, interleave, stream)
                             : high_radix_t<int32_t>(x, w1, w2, y, N, B, shifts, interleave, stream);
}

}  // namespace fftlib
