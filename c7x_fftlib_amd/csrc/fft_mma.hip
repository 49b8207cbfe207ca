// Fixed-point DFT kernels of the fft_c7xmma family on gfx950: INT16 data
// with 64-bit accumulation and INT32 data with 128-bit accumulation, the
// reference's rounding right shifts and saturation, so results are
// bit-identical to its natural-C kernels
// (src/fft_c7xmma/FFTLIB_FFT_dftSmall_ixX_cxX_oxX/..._cn.cpp:19-46,
// src/linalg_c7xmma/FFTLIB_LINALG_matrixMatrixMultiply_ixX_ixX_oxX/..._cn.cpp:27-65,
// src/fft_c7xmma/FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX/..._cn.cpp:19-455).
//
// The DFT-matrix products of dftSmall and dftLarge (Y = X W over the batch)
// run on the FP64 matrix cores (v_mfma_f64_16x16x4f64), exactly: every
// product of two INT16 values (|p| <= 2^30) and every sum of up to 2^22 of
// them is an integer below 2^53, so the FP64 accumulation is the exact
// integer sum. INT32 values are split into a signed high and an unsigned low
// 16-bit part, x = xh 2^16 + xl; the three accumulators sum xh wh, xh wl +
// xl wh and xl wl (each term below 2^32, each sum below 2^53 for the K of
// these kernels), and the 128-bit accumulator is rebuilt from them exactly.
// The rounding shift and saturation follow as in the reference.
//
// The stages of highRadixDecompositions are the same kind of product (a small
// DFT matrix times many complex vectors, with the reference's saturating
// negations applied to the gathered operands) on the same exact tiles. These
// kernels are small (the reference's sizes: up to 8192 points, 2048
// channels) and bound by launch latency rather than by HBM or ALU.
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

// ---- DFT-matrix product on the FP64 matrix cores --------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mfma_f64(double a, double b, f64x4 c)
{
   return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Largest K (dot-product length) whose exact sum the FP64 accumulators hold
// (INT32 parts: |term| < 2^32, two terms per K in the middle accumulator).
constexpr uint32_t kMfmaMaxK = 1u << 20;

// One 16 x 16 tile of Y = X W per wavefront: operand layouts of
// v_mfma_f64_16x16x4f64 for lane l = 16 q + r (checked on the GPU with a
// one-hot probe): A[r][q] (row r, k = q of the step), B[q][r] (k = q,
// column r); element v of the result is D[q + 4 v][r].
template <typename T>
__global__ void __launch_bounds__(kThreads) matrix_mfma_kernel(const T *__restrict__ x, const T *__restrict__ w,
                                                               T *__restrict__ y, uint32_t rows, uint32_t k2,
                                                               uint32_t ldw, uint32_t shift)
{
   const uint32_t lane = threadIdx.x % 64u, q = lane / 16u, r = lane % 16u;
   const uint32_t tilesJ = (k2 + 15u) / 16u;
   const uint32_t tile   = blockIdx.x * (kThreads / 64u) + threadIdx.x / 64u;
   const uint32_t ti = tile / tilesJ, tj = tile - ti * tilesJ;
   if (ti * 16u >= rows) return;   // the whole wavefront
   const uint32_t l = ti * 16u + r, j = tj * 16u + r;
   f64x4          acc[3];
#pragma unroll
   for (int a = 0; a < 3; ++a) acc[a] = f64x4{0.0, 0.0, 0.0, 0.0};
   for (uint32_t k0 = 0; k0 < k2; k0 += 4u) {
      const uint32_t k  = k0 + q;
      const int32_t  av = (l < rows && k < k2) ? (int32_t) x[(size_t) l * k2 + k] : 0;
      const int32_t  bv = (j < k2 && k < k2) ? (int32_t) w[(size_t) k * ldw + j] : 0;
      if constexpr (sizeof(T) == 2) {
         acc[0] = mfma_f64((double) av, (double) bv, acc[0]);
      }
      else {
         const double ah = (double) (av >> 16), al = (double) (av & 0xFFFF);
         const double bh = (double) (bv >> 16), bl = (double) (bv & 0xFFFF);
         acc[0] = mfma_f64(ah, bh, acc[0]);
         acc[1] = mfma_f64(ah, bl, acc[1]);
         acc[1] = mfma_f64(al, bh, acc[1]);
         acc[2] = mfma_f64(al, bl, acc[2]);
      }
   }
#pragma unroll
   for (int v = 0; v < 4; ++v) {
      const uint32_t row = ti * 16u + q + 4u * (uint32_t) v, col = tj * 16u + r;
      if (row >= rows || col >= k2) continue;
      typename Acc<T>::type s;
      if constexpr (sizeof(T) == 2) {
         s = (int64_t) acc[0][v];
      }
      else {
         s = ((__int128) (int64_t) acc[0][v] << 32) + ((__int128) (int64_t) acc[1][v] << 16) + (__int128) (int64_t) acc[2][v];
      }
      y[(size_t) row * k2 + col] = shift_round<T>(s, shift);
   }
}

// Exact integer sums over K terms of a(k) b(k) on the FP64 matrix cores for
// one 16 x 16 tile: a(k) is the lane's A operand (its row r = lane % 16 at k),
// b(k) its B operand (its column r at k); element v of the result holds
// row q + 4 v, column r (q = lane / 16). INT32 operands are split as in
// matrix_mfma_kernel.
template <typename T>
struct ExactTile {
   f64x4 p[sizeof(T) == 2 ? 1 : 3];
};

template <typename T, typename FA, typename FB>
__device__ __forceinline__ ExactTile<T> exact_tile(uint32_t K, FA a, FB b)
{
   const uint32_t q = (threadIdx.x % 64u) / 16u;
   ExactTile<T>   t;
#pragma unroll
   for (auto &x : t.p) x = f64x4{0.0, 0.0, 0.0, 0.0};
   for (uint32_t k0 = 0; k0 < K; k0 += 4u) {
      const uint32_t k  = k0 + q;
      const int32_t  av = k < K ? a(k) : 0;
      const int32_t  bv = k < K ? b(k) : 0;
      if constexpr (sizeof(T) == 2) {
         t.p[0] = mfma_f64((double) av, (double) bv, t.p[0]);
      }
      else {
         const double ah = (double) (av >> 16), al = (double) (av & 0xFFFF);
         const double bh = (double) (bv >> 16), bl = (double) (bv & 0xFFFF);
         t.p[0] = mfma_f64(ah, bh, t.p[0]);
         t.p[1] = mfma_f64(ah, bl, t.p[1]);
         t.p[1] = mfma_f64(al, bh, t.p[1]);
         t.p[2] = mfma_f64(al, bl, t.p[2]);
      }
   }
   return t;
}

template <typename T>
__device__ __forceinline__ typename Acc<T>::type exact_value(const ExactTile<T> &t, int v)
{
   if constexpr (sizeof(T) == 2)
      return (int64_t) t.p[0][v];
   else
      return ((__int128) (int64_t) t.p[0][v] << 32) + ((__int128) (int64_t) t.p[1][v] << 16) +
             (__int128) (int64_t) t.p[2][v];
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

// ---- the high-radix stages as matrix products on the FP64 matrix cores --------
// Each stage is a small DFT matrix times many complex vectors (the stages of
// ..._cn.cpp): rows of the product are the stage's independent vectors,
// columns its DFT outputs, K the real and imaginary parts of the inputs. The
// reference's saturating negations are applied to the gathered operands, and
// each output is rounded by the stage shift, multiplied by its twiddle and
// rounded again (twiddle_store). One 16-row tile per wavefront.

// Stage 0 (..._cn.cpp:76-128): the d0-point DFT of points p + m (N / d0),
// then the twiddle. Rows (j, n, k), k < MMA/2; columns l < d0; K = 2 d0.
template <typename T>
__global__ void __launch_bounds__(kThreads) hr_stage0_mfma(const T *__restrict__ src, T *__restrict__ dst,
                                                           const T *__restrict__ w1, const T *__restrict__ w2,
                                                           uint32_t N, uint32_t B, uint32_t d0, uint32_t shift)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   const uint32_t lane = threadIdx.x % 64u, q = lane / 16u, r = lane % 16u;
   const uint32_t nGroups = (N * 2) / (MMA * d0);
   const uint32_t rows    = B * nGroups * (MMA / 2);
   const uint32_t tile    = blockIdx.x * (kThreads / 64u) + threadIdx.x / 64u;
   if (tile * 16u >= rows) return;
   auto base = [&](uint32_t R) {
      const uint32_t k = R % (MMA / 2), n = (R / (MMA / 2)) % nGroups, j = R / (MMA / 2) / nGroups;
      return (size_t) j * N * 2 + (n * MMA + k * 2);
   };
   const uint32_t Ra = tile * 16u + r;
   const bool     va = Ra < rows;
   const size_t   ba = va ? base(Ra) : 0;
   auto           at = [&](uint32_t kk) { return ba + ((kk >> 1) * N * 2) / d0; };
   auto           wb = [&](uint32_t kk) -> int32_t { return r < d0 ? w1[r * MMA + kk] : 0; };
   const ExactTile<T> re = exact_tile<T>(2 * d0, [&](uint32_t kk) -> int32_t {
      if (!va) return 0;
      const size_t b = at(kk);
      return (kk & 1u) ? neg(src[b + 1]) : src[b];
   }, wb);
   const ExactTile<T> im = exact_tile<T>(2 * d0, [&](uint32_t kk) -> int32_t {
      if (!va) return 0;
      const size_t b = at(kk);
      return (kk & 1u) ? src[b] : src[b + 1];
   }, wb);
#pragma unroll
   for (int v = 0; v < 4; ++v) {
      const uint32_t R = tile * 16u + q + 4u * (uint32_t) v, l = r;
      if (R >= rows || l >= d0) continue;
      const uint32_t k = R % (MMA / 2), n = (R / (MMA / 2)) % nGroups, j = R / (MMA / 2) / nGroups;
      const T        ra = shift_round<T>(exact_value(re, v), shift), ia = shift_round<T>(exact_value(im, v), shift);
      const uint32_t wi = n * MMA * d0 + k * 2 + l * MMA;
      twiddle_store(dst, (size_t) j * N * 2 + (n * MMA + k * 2) + (l * N * 2) / d0, ra, ia, w2[wi], w2[wi + 1]);
   }
}

// Stages 1-3 (..._cn.cpp:138-203): the MMA/2-point DFT over points spaced by
// d / (MMA/2) within sub-transforms of d points, then the twiddle. Rows
// (j, k, l, i), i < MMA/2; columns m < MMA/2; K = MMA.
template <typename T>
__global__ void __launch_bounds__(kThreads) hr_stage_mfma(const T *__restrict__ src, T *__restrict__ dst,
                                                          const T *__restrict__ w1, const T *__restrict__ w2,
                                                          uint32_t N, uint32_t B, uint32_t d, uint32_t shift)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   const uint32_t lane = threadIdx.x % 64u, q = lane / 16u, r = lane % 16u;
   const uint32_t nSub = N / d, nL = (d * 4) / (MMA * MMA);
   const uint32_t rows = B * nSub * nL * (MMA / 2);
   const uint32_t tile = blockIdx.x * (kThreads / 64u) + threadIdx.x / 64u;
   if (tile * 16u >= rows) return;
   auto base = [&](uint32_t R) {
      const uint32_t i = R % (MMA / 2), l = (R / (MMA / 2)) % nL, k = (R / (MMA / 2) / nL) % nSub,
                     j = R / (MMA / 2) / nL / nSub;
      return (size_t) j * N * 2 + k * d * 2 + (l * MMA + i * 2);
   };
   const uint32_t Ra = tile * 16u + r;
   const bool     va = Ra < rows;
   const size_t   ba = va ? base(Ra) : 0;
   auto           at = [&](uint32_t kk) { return ba + (kk >> 1) * (d * 4) / MMA; };
   auto           wb = [&](uint32_t kk) -> int32_t { return r < MMA / 2 ? w1[r * MMA + kk] : 0; };
   const ExactTile<T> re = exact_tile<T>(MMA, [&](uint32_t kk) -> int32_t {
      if (!va) return 0;
      const size_t b = at(kk);
      return (kk & 1u) ? neg(src[b]) : src[b + 1];
   }, wb);
   const ExactTile<T> im = exact_tile<T>(MMA, [&](uint32_t kk) -> int32_t {
      if (!va) return 0;
      const size_t b = at(kk);
      return (kk & 1u) ? src[b + 1] : src[b];
   }, wb);
#pragma unroll
   for (int v = 0; v < 4; ++v) {
      const uint32_t R = tile * 16u + q + 4u * (uint32_t) v, m = r;
      if (R >= rows || m >= MMA / 2) continue;
      const uint32_t i = R % (MMA / 2), l = (R / (MMA / 2)) % nL, k = (R / (MMA / 2) / nL) % nSub,
                     j = R / (MMA / 2) / nL / nSub;
      const T        ra = shift_round<T>(exact_value(re, v), shift), ia = shift_round<T>(exact_value(im, v), shift);
      const uint32_t wi = l * MMA * MMA / 2 + i * 2 + m * MMA;
      twiddle_store(dst, (size_t) j * N * 2 + k * d * 2 + l * MMA + i * 2 + m * (d * 4) / MMA, ra, ia, w2[wi],
                    w2[wi + 1]);
   }
}

// Final stage (..._cn.cpp:383-433): MMA/2-point DFTs of consecutive points,
// stored in digit-reversed (natural output) order. Rows (j, k, m, i, h);
// columns l < MMA/2; K = MMA.
template <typename T>
__global__ void __launch_bounds__(kThreads) hr_dft_mfma(const T *__restrict__ src, T *__restrict__ dst,
                                                        const T *__restrict__ wd, uint32_t N, uint32_t B, uint32_t d0,
                                                        uint32_t mC, uint32_t iC, uint32_t hC, uint32_t shift)
{
   constexpr uint32_t MMA = Acc<T>::mma;
   const uint32_t lane = threadIdx.x % 64u, q = lane / 16u, r = lane % 16u;
   const uint32_t rows = B * d0 * mC * iC * hC;
   const uint32_t tile = blockIdx.x * (kThreads / 64u) + threadIdx.x / 64u;
   if (tile * 16u >= rows) return;
   const uint32_t Ra = tile * 16u + r;
   const bool     va = Ra < rows;
   size_t         ba = 0;
   if (va) {
      uint32_t t = Ra;
      const uint32_t h = t % hC;
      t /= hC;
      const uint32_t i = t % iC;
      t /= iC;
      const uint32_t m = t % mC;
      t /= mC;
      const uint32_t k = t % d0, j = t / d0;
      ba = (size_t) j * N * 2 + k * (N * 2) / d0 + m * (N * 4) / (d0 * MMA) + i * (N * 8) / (d0 * MMA * MMA) + h * MMA;
   }
   auto a = [&](uint32_t kk) -> int32_t { return va ? src[ba + kk] : 0; };
   const ExactTile<T> re =
      exact_tile<T>(MMA, a, [&](uint32_t kk) -> int32_t { return r < MMA / 2 ? wd[r * 2 * MMA + kk] : 0; });
   const ExactTile<T> im =
      exact_tile<T>(MMA, a, [&](uint32_t kk) -> int32_t { return r < MMA / 2 ? wd[(r * 2 + 1) * MMA + kk] : 0; });
#pragma unroll
   for (int v = 0; v < 4; ++v) {
      const uint32_t R = tile * 16u + q + 4u * (uint32_t) v, l = r;
      if (R >= rows || l >= MMA / 2) continue;
      uint32_t t = R;
      const uint32_t h = t % hC;
      t /= hC;
      const uint32_t i = t % iC;
      t /= iC;
      const uint32_t m = t % mC;
      t /= mC;
      const uint32_t k = t % d0, j = t / d0;
      const size_t   o = (size_t) j * N * 2 + l * d0 * mC * iC * hC * 2 + h * d0 * mC * iC * 2 + i * d0 * mC * 2 +
                       m * d0 * 2 + k * 2;
      dst[o]     = shift_round<T>(exact_value(re, v), shift);
      dst[o + 1] = shift_round<T>(exact_value(im, v), shift);
   }
}

inline dim3 grid_of(uint64_t items) { return dim3((unsigned) ((items + kThreads - 1) / kThreads)); }
// One wavefront per 16 rows.
inline dim3 grid_of_tiles(uint64_t rows)
{
   const uint64_t tiles = (rows + 15u) / 16u;
   return dim3((unsigned) ((tiles + kThreads / 64 - 1) / (kThreads / 64)));
}

template <typename T>
hipError_t matrix_t(const void *x, const void *w, void *y, uint32_t rows, uint32_t k2, uint32_t ldw, uint32_t shift,
                    hipStream_t s)
{
   if ((uint64_t) rows * k2 == 0) return hipSuccess;
   if (k2 > kMfmaMaxK) return hipErrorInvalidValue;   // beyond the exact FP64 sums (far above the API's sizes)
   // one wavefront per 16 x 16 output tile
   const uint64_t tiles = (uint64_t) ((rows + 15u) / 16u) * ((k2 + 15u) / 16u);
   hipLaunchKernelGGL(matrix_mfma_kernel<T>, dim3((unsigned) ((tiles + kThreads / 64 - 1) / (kThreads / 64))),
                      dim3(kThreads), 0, s, (const T *) x, (const T *) w, (T *) y, rows, k2, ldw, shift);
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
   hipLaunchKernelGGL(hr_stage0_mfma<T>, grid_of_tiles((uint64_t) B * ((N * 2) / (MMA * d0)) * (MMA / 2)),
                      dim3(kThreads), 0, s, (const T *) src, dst, w1, p2, N, B, d0, shifts[0]);
   if ((e = hipGetLastError()) != hipSuccess) return e;
   T *tmp = src;
   src = dst, dst = tmp;
   p2 += d * 2;
   d /= d0;
   ++cnt;
   for (uint32_t st = 1; st <= 3 && d > MMA / 2; ++st) {
      hipLaunchKernelGGL(hr_stage_mfma<T>,
                         grid_of_tiles((uint64_t) B * (N / d) * ((d * 4) / (MMA * MMA)) * (MMA / 2)), dim3(kThreads),
                         0, s, (const T *) src, dst, w1 + MMA * MMA / 2, p2, N, B, d, shifts[st]);
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
   hipLaunchKernelGGL(hr_dft_mfma<T>, grid_of_tiles((uint64_t) B * d0 * mC * iC * hC), dim3(kThreads), 0, s,
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
