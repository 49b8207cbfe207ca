/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in C (gcc, with __int128), of the natural-C kernels of
 * TexasInstruments/c7x_fftlib's fft_c7xmma directory: the DFT-as-matrix
 * kernels dftSmall and dftLarge, the highRadixDecompositions FFT and the
 * coordinator's kernel choice, on INT16 data with 64-bit accumulation and
 * INT32 data with 128-bit accumulation. Only tests/ load it, as the checker
 * of the GPU product.
 *
 * Reference paths are relative to the reference root; "$MMA" abbreviates
 * src/fft_c7xmma. The table generators restate the twGen functions of the
 * kernels' c71 files, which are host C code (cos/sin in double, PI =
 * 3.141592654, rounded and saturated to the data type).
 *
 * Pinning: the reference's committed vectors
 * test/fft_c7xmma/<kernel>/<kernel>_idat.c (static inputs, expected outputs,
 * shift vectors; converted to tests/golden/mma_*.npz), compared with the
 * reference driver's tolerance of 50 (tests/test_oracle_mma.py), and the
 * FP64 DFT scaled by the shifts.
 *
 * Assumption (parity unpinned): a rounding shift whose count minus one
 * reaches the accumulator width shifts in sign bits only (count clamped to
 * 63 / 127); the reference's intrinsic is not defined here and no committed
 * vector uses such shifts.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define MMA16 32u /* FFTLIB_MMA_SIZE_16_BIT: 64-byte vectors of int16 */
#define MMA32 16u /* FFTLIB_MMA_SIZE_32_BIT: 64-byte vectors of int32 */

typedef __int128 i128;

/* ---- fixed-point helpers: src/common/FFTLIB_utility.h ------------------- */
/* FFTLIB_UTIL_saturate(int64_t) -> int16, FFTLIB_utility.h:518-530 */
static int16_t sat16(int64_t x) { return x > 0x7FFF ? 0x7FFF : x < -0x8000 ? -0x8000 : (int16_t) x; }

/* FFTLIB_UTIL_saturate(hi, lo) -> int32, FFTLIB_utility.h:546-566: a
 * negative value saturates unless hi is all ones and lo's top 33 bits are
 * ones; a non-negative one saturates when lo's top 33 bits are not zero
 * (hi is not inspected). */
static int32_t sat32(i128 v)
{
   const uint64_t hi = (uint64_t) (v >> 64), lo = (uint64_t) v;
   if (hi & 0x8000000000000000ull) {
      if (hi != 0xFFFFFFFFFFFFFFFFull || (~lo & 0xFFFFFFFF80000000ull) != 0) return INT32_MIN;
      return (int32_t) lo;
   }
   if (lo & 0xFFFFFFFF80000000ull) return INT32_MAX;
   return (int32_t) lo;
}

/* FFTLIB_UTIL_shiftAndRound<int64_t, int16_t> (FFTLIB_utility.h:576-596) and
 * <FFTLIB_int128_t, int32_t> (:634-656): shift 0 saturates; otherwise
 * ((v >> (shift-1)) + 1) >> 1 with arithmetic shifts, then saturation. The
 * shift travels as uint8_t. */
static int16_t sr16(int64_t v, uint32_t shift)
{
   shift &= 0xFFu;
   if (shift == 0) return sat16(v);
   const uint32_t s = shift - 1 > 63 ? 63 : shift - 1;
   int64_t        t = (v >> s) + 1;
   return sat16(t >> 1);
}
static int32_t sr32(i128 v, uint32_t shift)
{
   shift &= 0xFFu;
   if (shift == 0) return sat32(v);
   const uint32_t s = shift - 1 > 127 ? 127 : shift - 1;
   i128           t = (v >> s) + 1;
   return sat32(t >> 1);
}

/* FFTLIB_UTIL_negate, FFTLIB_utility.h:229-242 */
static int16_t neg16(int16_t a) { return a == -32768 ? 32767 : (int16_t) -a; }
static int32_t neg32(int32_t a) { return a == INT32_MIN ? INT32_MAX : -a; }

/* FFTLIB_UTIL_typeConv_i64f_oxX (FFTLIB_utility.h:666-690): floor(0.5 + x)
 * saturated to the type; FFTLIB_UTIL_cos/sin_i64f_oxX (:704-719) scale first. */
static int64_t conv(double x, int bits)
{
   const int64_t maxv = ((int64_t) 1 << (bits - 1)) - 1;
   const int64_t v    = (int64_t) floor(0.5 + x);
   if (v >= maxv) return maxv;
   if (v <= -maxv - 1) return -maxv - 1;
   return v;
}
static double tw_scale(int bits) { return (double) (((uint32_t) 1 << (bits - 2)) - 1) + (double) ((uint32_t) 1 << (bits - 2)) + 0.5; }

static const double kPi = 3.141592654;

/* ---- kernel choice: $MMA/FFTLIB_FFT_fft_ixX_cxX_oxX/FFTLIB_FFT_fft_ixX_cxX_oxX.cpp:13-49 ---- */
/* 0 = dftSmall, 1 = dftLarge, 2 = highRadixDecompositions */
int oracle_mma_variant(uint32_t fftSize, uint32_t batchSize, uint32_t interleave, uint32_t mma)
{
   if (fftSize <= mma / 2) return 0;
   if (fftSize == mma) return 1;
   if (fftSize == mma * 2) {
      if (batchSize < mma / 2) return 2;
      if (batchSize == mma / 2) return interleave ? 2 : 1;
      return 1;
   }
   return 2;
}

/* Stage-0 decomposition factor: fftSize divided by MMA/2 while larger
 * ($MMA/FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX/..._cn.cpp:69-74). */
static uint32_t stage0_factor(uint32_t fftSize, uint32_t mma)
{
   uint32_t f = fftSize;
   while (f > mma / 2) f /= mma / 2;
   return f;
}

/* Size in elements of highRadix's caller table (c71/..._ci.cpp:1975-2020). */
uint32_t oracle_mma_hr_wsize(uint32_t fftSize, uint32_t mma)
{
   const uint32_t d0 = stage0_factor(fftSize, mma);
   uint32_t       w = 2 * fftSize, d = fftSize / d0;
   while (d > mma / 2) {
      w += 2 * d;
      d /= mma / 2;
   }
   return w + mma * mma;
}

/* The generators and kernels below are written once as macros over
 * (data type T, accumulator A, bit width, helpers) and instantiated for
 * int16/int64 and int32/int128. */

/* ---- dftSmall table: $MMA/FFTLIB_FFT_dftSmall_ixX_cxX_oxX/c71/..._ci.cpp:146-205 ---- */
#define DEF_DFTSMALL_TW(SFX, T, BITS, MMA)                                                                         \
   void oracle_mma_dftsmall_tw_##SFX(T *w, uint32_t fftSize, uint32_t batchSize, uint32_t interleave)              \
   {                                                                                                              \
      const double   sc  = tw_scale(BITS);                                                                        \
      const uint32_t bpr = (MMA >> 1) / fftSize;                                                                  \
      memset(w, 0, sizeof(T) * MMA * MMA);                                                                         \
      for (uint32_t j = 0; j < bpr; j++)                                                                          \
         for (uint32_t k = 0; k < fftSize; k++)                                                                   \
            for (uint32_t l = 0; l < (interleave ? bpr : batchSize); l++) {                                       \
               if (l != j) continue;                                                                              \
               for (uint32_t n = 0; n < fftSize; n++) {                                                           \
                  const double a = 2 * kPi * k * n / fftSize;                                                     \
                  const T      c = (T) conv(sc * cos(a), BITS), s = (T) conv(sc * sin(a), BITS);                  \
                  if (interleave) {                                                                               \
                     const uint32_t idx = (j * fftSize + k) * 2 * MMA + l * fftSize * 2;                          \
                     w[idx + n * 2]           = c;                                                                \
                     w[idx + n * 2 + 1]       = (T) -s;                                                           \
                     w[idx + MMA + n * 2]     = s;                                                                \
                     w[idx + MMA + n * 2 + 1] = c;                                                                \
                  }                                                                                               \
                  else {                                                                                          \
                     const uint32_t idx = (j * fftSize * 2 + k) * MMA + l * fftSize * 2;                          \
                     w[idx + n]                         = c;                                                      \
                     w[idx + fftSize + n]               = (T) -s;                                                 \
                     w[idx + fftSize * MMA + n]         = s;                                                      \
                     w[idx + fftSize * MMA + fftSize + n] = c;                                                    \
                  }                                                                                               \
               }                                                                                                  \
            }                                                                                                     \
   }

/* ---- dftLarge table: $MMA/FFTLIB_FFT_dftLarge_ixX_cxX_oxX/c71/..._ci.cpp:113-154 ---- */
#define DEF_DFTLARGE_TW(SFX, T, BITS)                                                                              \
   void oracle_mma_dftlarge_tw_##SFX(T *w, uint32_t fftSize, uint32_t interleave)                                  \
   {                                                                                                              \
      const double   sc = tw_scale(BITS);                                                                         \
      const uint32_t N  = fftSize;                                                                                \
      for (uint32_t j = 0; j < N; j++)                                                                            \
         for (uint32_t k = 0; k < N; k++) {                                                                       \
            const double a = 2 * kPi * j * k / N;                                                                 \
            const T      c = (T) conv(sc * cos(a), BITS), s = (T) conv(sc * sin(a), BITS);                        \
            if (interleave) {                                                                                     \
               w[j * 2 * N * 2 + k * 2]       = c;                                                                \
               w[j * 2 * N * 2 + k * 2 + 1]   = (T) -s;                                                           \
               w[(j * 2 + 1) * N * 2 + k * 2] = s;                                                                \
               w[(j * 2 + 1) * N * 2 + k * 2 + 1] = c;                                                            \
            }                                                                                                     \
            else {                                                                                                \
               w[j * N * 2 + k]                 = c;                                                              \
               w[j * N * 2 + N + k]             = (T) -s;                                                         \
               w[N * N * 2 + j * N * 2 + k]     = s;                                                              \
               w[N * N * 2 + j * N * 2 + N + k] = c;                                                              \
            }                                                                                                     \
         }                                                                                                        \
   }

/* ---- matrix kernels ------------------------------------------------------
 * dftSmall: Y[l][j] = sr(sum_k X[l][k] W[k*MMA + j]) over the 2N x 2N leading
 * block ($MMA/FFTLIB_FFT_dftSmall_ixX_cxX_oxX/..._cn.cpp:19-46); dftLarge:
 * the same with W rows of 2N elements, through
 * src/linalg_c7xmma/FFTLIB_LINALG_matrixMatrixMultiply_ixX_ixX_oxX/..._cn.cpp:27-65
 * (M = batchSize, K = N = 2 fftSize, element strides 2 fftSize). */
#define DEF_MATRIX(SFX, T, A, SR)                                                                                  \
   void oracle_mma_matrix_##SFX(const T *x, const T *w, T *y, uint32_t rows, uint32_t k2, uint32_t ldw,            \
                                uint32_t shift)                                                                   \
   {                                                                                                              \
      for (uint32_t l = 0; l < rows; l++)                                                                         \
         for (uint32_t j = 0; j < k2; j++) {                                                                      \
            A acc = 0;                                                                                            \
            for (uint32_t k = 0; k < k2; k++) acc += (A) x[l * k2 + k] * (A) w[ldw * k + j];                      \
            y[l * k2 + j] = SR(acc, shift);                                                                       \
         }                                                                                                        \
   }

/* ---- highRadixDecompositions tables: c71/..._ci.cpp:2038-2168 ----
 * w1 (MMA x MMA elements): the stage-0 block-diagonal DFT matrix of the
 * decomposition factor, then the MMA/2-point DFT matrix used by stages 1-3
 * (the reference writes it to a library-global buffer, FFTLIB_FFT_twData);
 * w2 (oracle_mma_hr_wsize elements): the caller's table of per-stage
 * twiddles and the final DFT matrix. */
#define DEF_HR_TW(SFX, T, BITS, MMA)                                                                               \
   void oracle_mma_hr_tw_##SFX(T *w1, T *w2, uint32_t fftSize)                                                    \
   {                                                                                                              \
      const double   sc = tw_scale(BITS);                                                                         \
      const uint32_t d0 = stage0_factor(fftSize, MMA);                                                            \
      memset(w1, 0, sizeof(T) * MMA * MMA / 2);                                                                    \
      for (uint32_t j = 0; j < MMA / (2 * d0); j++)                                                               \
         for (uint32_t k = 0; k < d0; k++)                                                                        \
            for (uint32_t l = 0; l < d0; l++) {                                                                   \
               const double   a = (2 * kPi * k * l) / d0;                                                         \
               const uint32_t i = j * d0 * MMA + j * d0 * 2 + k * MMA + l * 2;                                    \
               w1[i]            = (T) conv(sc * cos(a), BITS);                                                    \
               w1[i + 1]        = (T) -(T) conv(sc * sin(a), BITS);                                               \
            }                                                                                                     \
      T *w1b = w1 + MMA * MMA / 2;                                                                                \
      for (uint32_t k = 0; k < MMA / 2; k++)                                                                      \
         for (uint32_t l = 0; l < MMA / 2; l++) {                                                                 \
            const double a         = (2 * kPi * k * l * 2) / MMA;                                                 \
            w1b[k * MMA + l * 2]     = (T) conv(sc * cos(a), BITS);                                               \
            w1b[k * MMA + l * 2 + 1] = (T) -(T) conv(sc * sin(a), BITS);                                          \
         }                                                                                                        \
      T       *p = w2;                                                                                            \
      uint32_t d = fftSize;                                                                                       \
      for (uint32_t j = 0; j < (d * 2) / (MMA * d0); j++)                                                         \
         for (uint32_t l = 0; l < d0; l++)                                                                        \
            for (uint32_t k = 0; k < MMA / 2; k++) {                                                              \
               const double   a = (2 * kPi * (j * MMA / 2 + k) * l) / d;                                          \
               const uint32_t i = j * MMA * d0 + l * MMA + k * 2;                                                 \
               p[i + 1]         = (T) conv(sc * cos(a), BITS);                                                    \
               p[i]             = (T) -(T) conv(sc * sin(a), BITS);                                               \
            }                                                                                                     \
      p += d * 2;                                                                                                 \
      d /= d0;                                                                                                    \
      while (d > MMA / 2) {                                                                                       \
         for (uint32_t j = 0; j < (d * 4) / (MMA * MMA); j++)                                                     \
            for (uint32_t l = 0; l < MMA / 2; l++)                                                                \
               for (uint32_t k = 0; k < MMA / 2; k++) {                                                           \
                  const double   a = (2 * kPi * (j * MMA / 2 + k) * l) / d;                                       \
                  const uint32_t i = j * MMA * MMA / 2 + l * MMA + k * 2;                                         \
                  p[i + 1]         = (T) conv(sc * cos(a), BITS);                                                 \
                  p[i]             = (T) -(T) conv(sc * sin(a), BITS);                                            \
               }                                                                                                  \
         p += d * 2;                                                                                              \
         d /= MMA / 2;                                                                                            \
      }                                                                                                           \
      for (uint32_t k = 0; k < MMA / 2; k++)                                                                      \
         for (uint32_t l = 0; l < MMA / 2; l++) {                                                                 \
            const double a = (2 * kPi * k * l * 2) / MMA;                                                         \
            const T      c = (T) conv(sc * cos(a), BITS), s = (T) conv(sc * sin(a), BITS);                        \
            p[k * 2 * MMA + l * 2]           = s;                                                                 \
            p[k * 2 * MMA + l * 2 + 1]       = c;                                                                 \
            p[(k * 2 + 1) * MMA + l * 2]     = c;                                                                 \
            p[(k * 2 + 1) * MMA + l * 2 + 1] = (T) -s;                                                            \
         }                                                                                                        \
   }

/* ---- highRadixDecompositions: $MMA/FFTLIB_FFT_highRadixDecompositions_ixX_cxX_oxX/..._cn.cpp:19-455 ----
 * x is used as work space (as in the reference). Stage 0 applies the
 * decomposition-factor DFT and the twiddles; stages 1-3 (while the remaining
 * sub-length exceeds MMA/2) the MMA/2-point DFT and twiddles; the last stage
 * the MMA/2-point DFT with the digit-reversed store. Non-interleaved data is
 * interleaved first and de-interleaved at the end. */
#define DEF_HR(SFX, T, A, MMA, SR, NEG, NBITS)                                                                     \
   static void hr_stage_##SFX(const T *src, T *dst, const T *w1, const T *w2, uint32_t N, uint32_t B, uint32_t d,  \
                              uint32_t shift)                                                                     \
   {                                                                                                              \
      const uint32_t nm = NBITS - 1;                                                                              \
      for (uint32_t j = 0; j < B; j++)                                                                            \
         for (uint32_t k = 0; k < N / d; k++)                                                                     \
            for (uint32_t l = 0; l < (d * 4) / (MMA * MMA); l++)                                                  \
               for (uint32_t i = 0; i < MMA / 2; i++)                                                             \
                  for (uint32_t m = 0; m < MMA / 2; m++) {                                                        \
                     A re = 0, im = 0;                                                                            \
                     for (uint32_t n = 0; n < MMA / 2; n++) {                                                     \
                        const uint32_t b = j * N * 2 + k * d * 2 + (l * MMA + i * 2) + n * (d * 4) / MMA;         \
                        im += (A) w1[m * MMA + n * 2] * (A) src[b];                                               \
                        im += (A) w1[m * MMA + n * 2 + 1] * (A) src[b + 1];                                       \
                        re += (A) w1[m * MMA + n * 2] * (A) src[b + 1];                                           \
                        re += (A) w1[m * MMA + n * 2 + 1] * (A) NEG(src[b]);                                      \
                     }                                                                                            \
                     const T        ra = SR(re, shift), ia = SR(im, shift);                                       \
                     const uint32_t wi = l * MMA * MMA / 2 + i * 2 + m * MMA;                                     \
                     const uint32_t o  = j * N * 2 + k * d * 2 + l * MMA + i * 2 + m * (d * 4) / MMA;             \
                     dst[o]     = SR((A) w2[wi] * (A) ra + (A) w2[wi + 1] * (A) ia, nm);                          \
                     dst[o + 1] = SR((A) NEG(w2[wi]) * (A) ia + (A) w2[wi + 1] * (A) ra, nm);                     \
                  }                                                                                               \
   }                                                                                                              \
   void oracle_mma_hr_exec_##SFX(T *x, const T *w1, const T *w2, T *y, uint32_t N, uint32_t B,                    \
                                 const uint32_t *shifts, uint32_t interleave)                                     \
   {                                                                                                              \
      const uint32_t nm = NBITS - 1;                                                                              \
      T             *src = x, *dst = y;                                                                           \
      if (!interleave) {                                                                                          \
         for (uint32_t i = 0; i < B; i++)                                                                         \
            for (uint32_t j = 0; j < N; j++) {                                                                    \
               y[i * N * 2 + j * 2]     = x[i * N * 2 + j];                                                       \
               y[i * N * 2 + j * 2 + 1] = x[i * N * 2 + N + j];                                                   \
            }                                                                                                     \
         src = y, dst = x;                                                                                        \
      }                                                                                                           \
      const uint32_t d0 = stage0_factor(N, MMA);                                                                  \
      uint32_t       d = N, cnt = 0;                                                                              \
      const T       *p2 = w2;                                                                                     \
      for (uint32_t j = 0; j < B; j++)                                                                            \
         for (uint32_t n = 0; n < (N * 2) / (MMA * d0); n++)                                                      \
            for (uint32_t k = 0; k < MMA / 2; k++)                                                                \
               for (uint32_t l = 0; l < d0; l++) {                                                                \
                  A re = 0, im = 0;                                                                               \
                  for (uint32_t m = 0; m < d0; m++) {                                                             \
                     const uint32_t b = j * N * 2 + (n * MMA + k * 2) + (m * N * 2) / d0;                         \
                     im += (A) w1[l * MMA + m * 2] * (A) src[b + 1];                                              \
                     im += (A) w1[l * MMA + m * 2 + 1] * (A) src[b];                                              \
                     re += (A) w1[l * MMA + m * 2] * (A) src[b];                                                  \
                     re += (A) w1[l * MMA + m * 2 + 1] * (A) NEG(src[b + 1]);                                     \
                  }                                                                                               \
                  const T        ra = SR(re, shifts[0]), ia = SR(im, shifts[0]);                                  \
                  const uint32_t wi = n * MMA * d0 + k * 2 + l * MMA;                                             \
                  const uint32_t o  = j * N * 2 + (n * MMA + k * 2) + (l * N * 2) / d0;                           \
                  dst[o]     = SR((A) p2[wi] * (A) ra + (A) p2[wi + 1] * (A) ia, nm);                             \
                  dst[o + 1] = SR((A) NEG(p2[wi]) * (A) ia + (A) p2[wi + 1] * (A) ra, nm);                        \
               }                                                                                                  \
      { T *t = src; src = dst; dst = t; }                                                                         \
      p2 += d * 2;                                                                                                \
      d /= d0;                                                                                                    \
      cnt++;                                                                                                      \
      for (uint32_t s = 1; s <= 3 && d > MMA / 2; s++) {                                                          \
         hr_stage_##SFX(src, dst, w1 + MMA * MMA / 2, p2, N, B, d, shifts[s]);                                    \
         { T *t = src; src = dst; dst = t; }                                                                      \
         p2 += d * 2;                                                                                             \
         d /= MMA / 2;                                                                                            \
         cnt++;                                                                                                   \
      }                                                                                                           \
      const T       *wd = p2;                                                                                     \
      const uint32_t q  = (N * 2) / (MMA * d0), r = (N * 4) / (MMA * MMA * d0);                                   \
      const uint32_t mC = q > MMA / 2 ? MMA / 2 : q;                                                              \
      const uint32_t iC = q > MMA / 2 ? (r > MMA / 2 ? MMA / 2 : r) : 1;                                          \
      const uint32_t hC = r > MMA / 2 ? (N * 8) / (MMA * MMA * MMA * d0) : 1;                                     \
      for (uint32_t j = 0; j < B; j++)                                                                            \
         for (uint32_t k = 0; k < d0; k++)                                                                        \
            for (uint32_t m = 0; m < mC; m++)                                                                     \
               for (uint32_t i = 0; i < iC; i++)                                                                  \
                  for (uint32_t h = 0; h < hC; h++)                                                               \
                     for (uint32_t l = 0; l < MMA / 2; l++) {                                                     \
                        A re = 0, im = 0;                                                                         \
                        for (uint32_t n = 0; n < MMA / 2; n++) {                                                  \
                           const uint32_t b = j * N * 2 + k * (N * 2) / d0 + m * (N * 4) / (d0 * MMA) +           \
                                              i * (N * 8) / (d0 * MMA * MMA) + h * MMA + n * 2;                   \
                           re += (A) wd[l * 2 * MMA + n * 2] * (A) src[b];                                        \
                           re += (A) wd[l * 2 * MMA + n * 2 + 1] * (A) src[b + 1];                                \
                           im += (A) wd[(l * 2 + 1) * MMA + n * 2] * (A) src[b];                                  \
                           im += (A) wd[(l * 2 + 1) * MMA + n * 2 + 1] * (A) src[b + 1];                          \
                        }                                                                                         \
                        const uint32_t o = j * N * 2 + l * d0 * mC * iC * hC * 2 + h * d0 * mC * iC * 2 +         \
                                           i * d0 * mC * 2 + m * d0 * 2 + k * 2;                                  \
                        dst[o]     = SR(re, shifts[4]);                                                           \
                        dst[o + 1] = SR(im, shifts[4]);                                                           \
                     }                                                                                            \
      if (!interleave) {                                                                                          \
         const T *s2 = (cnt & 1u) ? y : x;                                                                        \
         T       *d2 = (cnt & 1u) ? x : y;                                                                        \
         for (uint32_t i = 0; i < B; i++)                                                                         \
            for (uint32_t j = 0; j < N; j++) {                                                                    \
               d2[i * N * 2 + j]     = s2[i * N * 2 + j * 2];                                                     \
               d2[i * N * 2 + N + j] = s2[i * N * 2 + j * 2 + 1];                                                 \
            }                                                                                                     \
      }                                                                                                           \
      if (cnt & 1u) memcpy(y, x, sizeof(T) * N * B * 2);                                                          \
   }

DEF_DFTSMALL_TW(i16, int16_t, 16, MMA16)
DEF_DFTSMALL_TW(i32, int32_t, 32, MMA32)
DEF_DFTLARGE_TW(i16, int16_t, 16)
DEF_DFTLARGE_TW(i32, int32_t, 32)
DEF_MATRIX(i16, int16_t, int64_t, sr16)
DEF_MATRIX(i32, int32_t, i128, sr32)
DEF_HR_TW(i16, int16_t, 16, MMA16)
DEF_HR_TW(i32, int32_t, 32, MMA32)
DEF_HR(i16, int16_t, int64_t, MMA16, sr16, neg16, 16)
DEF_HR(i32, int32_t, i128, MMA32, sr32, neg32, 32)
