/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in plain C99, of the natural-C ("_cn") kernels of
 * TexasInstruments/c7x_fftlib's fft_c7x directory. It is the correctness
 * checker for the GPU product: tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py are the only code that loads it. The product
 * library never links or calls it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so float expressions are
 * evaluated in the same order and without fused multiply-adds, as the
 * reference's host build does on x86-64).
 *
 * Every function cites the reference lines it restates. Reference paths are
 * relative to the reference repository root; "$C2CB" abbreviates
 * src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc and likewise for the other
 * kernel directories ($C2C1, $IC2CB, $IC2C1, $I16B, $I16, $R2CB, $R2C1, $C2RB,
 * $C2R1).
 *
 * Pinning (what checks this file, besides the product agreeing with it):
 *   - fp32 C2C fwd/inv: FP64 direct DFT closed form, impulse/tone/linearity/
 *     Parseval/round-trip invariants (tests/test_oracle_fp32.py);
 *   - R2C / C2R / single-channel IFFT: the reference's committed
 *     extended_test_cases fixtures (tests/golden, .npz files) and FP64 rfft/irfft;
 *   - int16: vectors written by the reference's own Python model of the kernel
 *     (tests/golden/gen_i16_vectors.py) for inputs that never wrap, plus
 *     hand-worked wrap/saturation cases and the FP64 DFT within the
 *     fixed-point error bound;
 *   - table generators: closed-form cos/sin values and the DFT checks above.
 * The reference _cn sources themselves cannot be compiled here: every one of
 * them includes src/common/FFTLIB_types.h, which includes TI's <c7x.h>, a
 * compiler header this machine does not have.
 */
#include "fftlib_cn.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* fp32 mixed-radix machinery                                               */
/* ------------------------------------------------------------------------ */

/* cos/sin of 2*pi*q/5 and 2*pi*q/3 rounded to float (the reference spells the
 * same values as float literals, $C2CB/_cn.c:12-17). Index q = (m*j) mod r. */
static const float kCos5[5] = {1.0f, 0.30901699437494745f, -0.80901699437494734f, -0.80901699437494734f,
                               0.30901699437494745f};
static const float kSin5[5] = {0.0f, 0.95105651629515353f, 0.58778525229247325f, -0.58778525229247325f,
                               -0.95105651629515353f};
static const float kCos3[3] = {1.0f, -0.5f, -0.5f};
static const float kSin3[3] = {0.0f, 0.86602540378443871f, -0.86602540378443871f};

typedef struct {
   uint32_t n5, n3, n4, n2;
} stage_counts_t;

/* Greedy factorisation 5 > 3 > 4 > 2 exactly as $C2CB/_cn.c:240-260: all
 * radix-5 stages first, then radix-3, then radix-4, then at most one radix-2.
 * A leftover factor (not 2/3/5-smooth) is ignored, as in the reference. */
static stage_counts_t count_stages(uint32_t n)
{
   stage_counts_t c = {0, 0, 0, 0};
   while (n > 1u) {
      if (n % 5u == 0u) {
         c.n5++;
         n /= 5u;
      }
      else if (n % 3u == 0u) {
         c.n3++;
         n /= 3u;
      }
      else if (n % 4u == 0u) {
         c.n4++;
         n /= 4u;
      }
      else if (n % 2u == 0u) {
         c.n2++;
         n /= 2u;
      }
      else {
         break;
      }
   }
   return c;
}

/* Small-DFT of radix r (3 or 5) on r complex inputs, output m multiplied by
 * twiddle w[m-1] (m >= 1); restates butterfly3x3/butterfly5x5
 * ($C2CB/_cn.c:76-186). The sums run in the reference's order: p0, then for
 * j = 1..r-1 the real-part term then the imaginary-part term. */
static void dft_odd(int r, const float *re, const float *im, const float *wRe, const float *wIm, float *outRe,
                    float *outIm)
{
   const float *cs = (r == 5) ? kCos5 : kCos3;
   const float *sn = (r == 5) ? kSin5 : kSin3;
   int           m, j;
   float         accRe, accIm;

   accRe = re[0];
   accIm = im[0];
   for (j = 1; j < r; j++) {
      accRe = accRe + re[j];
   }
   for (j = 1; j < r; j++) {
      accIm = accIm + im[j];
   }
   outRe[0] = accRe;
   outIm[0] = accIm;
   for (m = 1; m < r; m++) {
      accRe = re[0];
      accIm = im[0];
      for (j = 1; j < r; j++) {
         int q = (m * j) % r;
         accRe = accRe + re[j] * cs[q];
         accRe = accRe + im[j] * sn[q];
      }
      for (j = 1; j < r; j++) {
         int q = (m * j) % r;
         accIm = accIm + re[j] * (-sn[q]);
         accIm = accIm + im[j] * cs[q];
      }
      outRe[m] = accRe * wRe[m - 1] - accIm * wIm[m - 1];
      outIm[m] = accRe * wIm[m - 1] + accIm * wRe[m - 1];
   }
}

/* Radix-4 DFT with output twiddles; restates butterfly4x4 ($C2CB/_cn.c:19-74).
 * Outputs are returned in natural order 0..3; the caller stores 1 and 2
 * swapped ($C2CB/_cn.c:368-381). */
static void dft4(const float *re, const float *im, const float *wRe, const float *wIm, float *outRe, float *outIm)
{
   float a1 = ((re[0] + im[1]) - re[2]) - im[3];
   float b1 = ((im[0] - re[1]) - im[2]) + re[3];
   float a2 = ((re[0] - re[1]) + re[2]) - re[3];
   float b2 = ((im[0] - im[1]) + im[2]) - im[3];
   float a3 = ((re[0] - im[1]) - re[2]) + im[3];
   float b3 = ((im[0] + re[1]) - im[2]) - re[3];

   outRe[0] = ((re[0] + re[1]) + re[2]) + re[3];
   outIm[0] = ((im[0] + im[1]) + im[2]) + im[3];
   outRe[1] = a1 * wRe[0] - b1 * wIm[0];
   outIm[1] = a1 * wIm[0] + b1 * wRe[0];
   outRe[2] = a2 * wRe[1] - b2 * wIm[1];
   outIm[2] = a2 * wIm[1] + b2 * wRe[1];
   outRe[3] = a3 * wRe[2] - b3 * wIm[2];
   outIm[3] = a3 * wIm[2] + b3 * wRe[2];
}

/* One in-place DIF stage of radix r over every channel; restates the stage
 * loops of $C2CB/_cn.c:266-386. M is the current sub-DFT length
 * (numPointsPerDft), pw the stage's twiddle block: for butterfly point index p
 * and output m >= 1 the twiddle is complex element (m-1)*(M/r) + p. */
static void dif_stage(float *pX, const float *pw, uint32_t N, uint32_t B, uint32_t M, int r)
{
   uint32_t q = M / (uint32_t) r;
   uint32_t k, l;
   int      m;

   for (k = 0; k < N / (uint32_t) r; k++) {
      uint32_t dft   = k / q;
      uint32_t point = k % q;
      uint32_t base  = dft * M + point;
      float    wRe[4], wIm[4];

      for (m = 1; m < r; m++) {
         wRe[m - 1] = pw[2u * ((uint32_t) (m - 1) * q + point)];
         wIm[m - 1] = pw[2u * ((uint32_t) (m - 1) * q + point) + 1u];
      }
      for (l = 0; l < B; l++) {
         float   *ch = pX + (size_t) l * N * 2u;
         float    re[5], im[5], oRe[5], oIm[5];
         for (m = 0; m < r; m++) {
            re[m] = ch[2u * (base + (uint32_t) m * q)];
            im[m] = ch[2u * (base + (uint32_t) m * q) + 1u];
         }
         if (r == 4) {
            dft4(re, im, wRe, wIm, oRe, oIm);
            /* bit-swapped store: output 1 -> slot 2, output 2 -> slot 1 */
            float t;
            t      = oRe[1];
            oRe[1] = oRe[2];
            oRe[2] = t;
            t      = oIm[1];
            oIm[1] = oIm[2];
            oIm[2] = t;
         }
         else {
            dft_odd(r, re, im, wRe, wIm, oRe, oIm);
         }
         for (m = 0; m < r; m++) {
            ch[2u * (base + (uint32_t) m * q)]      = oRe[m];
            ch[2u * (base + (uint32_t) m * q) + 1u] = oIm[m];
         }
      }
   }
}

/* Final radix-2 stage on adjacent pairs, no twiddles ($C2CB/_cn.c:388-407). */
static void radix2_pairs(float *pX, uint32_t N, uint32_t B)
{
   uint32_t k, l;
   for (k = 0; k < N / 2u; k++) {
      for (l = 0; l < B; l++) {
         float *a   = pX + (size_t) l * N * 2u + 4u * k;
         float  s0  = a[0] + a[2];
         float  s1  = a[1] + a[3];
         float  d0  = a[0] - a[2];
         float  d1  = a[1] - a[3];
         a[0]       = s0;
         a[1]       = s1;
         a[2]       = d0;
         a[3]       = d1;
      }
   }
}

/* Bit reversal over the number of bits of (maxValue) ($C2CB/_cn.c:188-207). */
static uint32_t bit_reverse(uint32_t in, uint32_t maxValue)
{
   uint32_t bits = 1u, out = 0u, i;
   for (i = 31u; i > 0u; i--) {
      if ((maxValue >> i) & 1u) {
         bits = i + 1u;
         break;
      }
   }
   for (i = 0; i < bits; i++) {
      out |= ((in >> i) & 1u) << (bits - 1u - i);
   }
   return out;
}

/* Position, in the in-place DIF result, of frequency k: mixed-radix digit
 * reversal over the digit sequence [5..., 3..., 2,2 per radix-4, 2]
 * ($C2CB/_cn.c:428-491); for pure powers of two this equals bit reversal
 * ($C2CB/_cn.c:414-427). */
static uint32_t dif_position(uint32_t k, uint32_t N)
{
   uint32_t digitsR[40], nd = 0, t = N, pos = 0, i;
   while (t % 5u == 0u) {
      digitsR[nd++] = 5u;
      t /= 5u;
   }
   while (t % 3u == 0u) {
      digitsR[nd++] = 3u;
      t /= 3u;
   }
   if (nd == 0u) {
      return bit_reverse(k, N - 1u);
   }
   while (t % 4u == 0u) {
      digitsR[nd++] = 2u;
      digitsR[nd++] = 2u;
      t /= 4u;
   }
   if (t == 2u) {
      digitsR[nd++] = 2u;
   }
   /* least significant digit of k (radix digitsR[0]) becomes the most
    * significant digit of the position */
   for (i = 0; i < nd; i++) {
      pos = pos * digitsR[i] + (k % digitsR[i]);
      k /= digitsR[i];
   }
   return pos;
}

/* Direction/scaling variants of the fp32 C2C _cn body. */
enum {
   C2C_FORWARD = 0,
   /* batched inverse: conj, multiply by (1.0f/N) up front, conj on the
    * reorder store ($IC2CB/_cn.c:209-229, :264-265, :442, :506) */
   C2C_INVERSE_PRESCALE = 1,
   /* single-channel inverse: conj up front, divide by (float)N and conj on
    * the reorder store ($IC2C1/_cn.c:205-214, :248, :436-437, :498-499) */
   C2C_INVERSE_POSTDIV = 2
};

/* Shared body of the fp32 C2C forward and inverse _cn kernels. */
static void c2c_core(float *pX, const float *pW, float *pY, uint32_t N, uint32_t B, int mode)
{
   stage_counts_t c  = count_stages(N);
   const float   *pw = pW;
   uint32_t       M  = N, s, k, l;
   int            inverse = (mode != C2C_FORWARD);

   if (mode == C2C_INVERSE_POSTDIV) {
      size_t i;
      for (i = 0; i < (size_t) N * B; i++) {
         pX[2 * i + 1] = (-1.0f) * pX[2 * i + 1];
      }
   }
   if (mode == C2C_INVERSE_PRESCALE) {
      /* conj, then scale by (1.0f/N) ($IC2CB/_cn.c:209-229, :264-265) */
      float scale = 1.0f / (float) N;
      size_t i;
      for (i = 0; i < (size_t) N * B; i++) {
         pX[2 * i + 1] = (-1.0f) * pX[2 * i + 1];
      }
      for (i = 0; i < (size_t) 2 * N * B; i++) {
         pX[i] = scale * pX[i];
      }
   }
   for (s = 0; s < c.n5; s++) {
      dif_stage(pX, pw, N, B, M, 5);
      M /= 5u;
      pw += 2u * 4u * M;
   }
   for (s = 0; s < c.n3; s++) {
      dif_stage(pX, pw, N, B, M, 3);
      M /= 3u;
      pw += 2u * 2u * M;
   }
   for (s = 0; s < c.n4; s++) {
      dif_stage(pX, pw, N, B, M, 4);
      M /= 4u;
      pw += 2u * 3u * M;
   }
   for (s = 0; s < c.n2; s++) {
      radix2_pairs(pX, N, B);
      M /= 2u;
   }
   for (k = 0; k < N; k++) {
      uint32_t p = dif_position(k, N);
      for (l = 0; l < B; l++) {
         const float *src = pX + (size_t) l * N * 2u + 2u * p;
         float       *dst = pY + (size_t) l * N * 2u + 2u * k;
         if (mode == C2C_INVERSE_POSTDIV) {
            dst[0] = src[0] / (float) N;
            dst[1] = -1.0f * (src[1] / (float) N);
         }
         else {
            dst[0] = src[0];
            /* inverse: conj on the reorder store ($IC2CB/_cn.c:442, :506) */
            dst[1] = inverse ? (-1.0f) * src[1] : src[1];
         }
      }
   }
}

/* 2/3/5-smoothness test of the checkParams functions ($C2CB/_cn.c:523-540). */
static int is_235_smooth(uint32_t n)
{
   uint32_t i;
   for (i = 0; i < 32u && n % 5u == 0u; i++) n /= 5u;
   for (i = 0; i < 32u && n % 3u == 0u; i++) n /= 3u;
   for (i = 0; i < 32u && n % 2u == 0u; i++) n /= 2u;
   return n == 1u;
}

/* $C2CB/_cn.c:210-495 */
FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                       FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                       FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                       uint32_t numPoints, uint32_t numChannels)
{
   (void) bufParamsX;
   (void) bufParamsW;
   (void) bufParamsY;
   c2c_core(pX, pW, pY, numPoints, numChannels, C2C_FORWARD);
   return FFTLIB_SUCCESS;
}

/* $C2CB/_cn.c:497-543 */
FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                   FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                   FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                   uint32_t numPoints, uint32_t numChannels)
{
   (void) numChannels;
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsY->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsW->dim_x != numPoints * 2u) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   return is_235_smooth(numPoints) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

/* $IC2CB/_cn.c:232-512 */
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                        FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                        FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                        uint32_t numPoints, uint32_t numChannels)
{
   (void) bufParamsX;
   (void) bufParamsW;
   (void) bufParamsY;
   c2c_core(pX, pW, pY, numPoints, numChannels, C2C_INVERSE_PRESCALE);
   return FFTLIB_SUCCESS;
}

/* $IC2CB/_cn.c:514-556: as the forward check, but a non-smooth N is
 * FFTLIB_ERR_NOT_IMPLEMENTED. */
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                    FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                    FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                    uint32_t numPoints, uint32_t numChannels)
{
   (void) numChannels;
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsY->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsW->dim_x != numPoints * 2u) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   return is_235_smooth(numPoints) ? FFTLIB_SUCCESS : FFTLIB_ERR_NOT_IMPLEMENTED;
}

/* $C2C1/_cn.c:203-489: one channel, N = dim_x / 2. */
FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                FFTLIB_bufParams1D_t *bufParamsY)
{
   (void) bufParamsW;
   (void) bufParamsY;
   c2c_core(pX, pW, pY, bufParamsX->dim_x >> 1, 1u, C2C_FORWARD);
   return FFTLIB_SUCCESS;
}

/* $C2C1/_cn.c:491-532: X, W and Y descriptors must all have the same dim_x. */
FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY)
{
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsW->dim_x || bufParamsX->dim_x != bufParamsY->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   return is_235_smooth(bufParamsX->dim_x / 2u) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

/* $IC2C1/_cn.c:216-503: one channel inverse, N = dim_x / 2, scaling by
 * division at the reorder store. */
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                 FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                 FFTLIB_bufParams1D_t *bufParamsY)
{
   (void) bufParamsW;
   (void) bufParamsY;
   c2c_core(pX, pW, pY, bufParamsX->dim_x >> 1, 1u, C2C_INVERSE_POSTDIV);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                             FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                             FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY)
{
   return FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams_cn(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY);
}

/* ------------------------------------------------------------------------ */
/* int16 fixed point ($I16B/_cn.c:11-338)                                   */
/* ------------------------------------------------------------------------ */

/* int -> int16 conversion modulo 2^16, the (implementation-defined) result of
 * the reference's int16_t assignments under gcc/x86 ($I16B/_cn.c:65-81). */
static int16_t wrap16(int32_t v) { return (int16_t) (uint16_t) ((uint32_t) v & 0xFFFFu); }

/* Real part of the saturating multiply by +j used by VCROT90H: -im, with
 * -(-32768) saturated to 32767 ($I16B/_cn.c:27-32). */
static int16_t rot90_re(int16_t im) { return (im == INT16_MIN) ? INT16_MAX : (int16_t) (-im); }

/* Q15 product rounding of VCMPYR1HH: floor((acc + 2^14) / 2^15) clamped to
 * int16 ($I16B/_cn.c:13-25). */
static int16_t q15_round_sat(int64_t acc)
{
   int64_t r = (acc + 16384) >> 15;
   if (r > 32767) r = 32767;
   if (r < -32768) r = -32768;
   return (int16_t) r;
}

/* Rounding right shift ($I16B/_cn.c:36-45): floor(v / 2^s) plus the last bit
 * shifted out. The reference forms the bit mask as int16(1u << (s - 1)); for
 * s == 0 that shift count is out of range in C and x86 takes it modulo 32
 * (mask 1<<31 -> int16 0), so s == 0 leaves v unchanged. Counts are taken
 * modulo 32 throughout, as x86 does. */
static int16_t shift_round(int16_t v, uint32_t s)
{
   int16_t mask = (int16_t) (uint16_t) ((1u << ((s - 1u) & 31u)) & 0xFFFFu);
   int32_t x    = v;
   int32_t sh   = x >> (s & 31u);
   return wrap16((x & mask) != 0 ? sh + 1 : sh);
}

/* Radix-4 int16 butterfly. With twiddles (w != NULL) outputs 1..3 are Q15
 * products; output 0 passes through (its Q15 round of x0 * 2^15 is exact).
 * Restates butterfly / butterflyNoTw ($I16B/_cn.c:47-168). */
static void bfly_i16(const int16_t *re, const int16_t *im, const int16_t *w, int16_t *oRe, int16_t *oIm)
{
   int16_t s1r = wrap16(re[0] + re[2]), s1i = wrap16(im[0] + im[2]);
   int16_t s2r = wrap16(re[1] + re[3]), s2i = wrap16(im[1] + im[3]);
   int16_t d1r = wrap16(re[0] - re[2]), d1i = wrap16(im[0] - im[2]);
   int16_t d2r = wrap16(re[1] - re[3]), d2i = wrap16(im[1] - im[3]);
   int16_t tr[4], ti[4];
   int     m;

   tr[0] = wrap16(s1r + s2r);
   ti[0] = wrap16(s1i + s2i);
   tr[1] = wrap16(d1r - rot90_re(d2i));
   ti[1] = wrap16(d1i - d2r);
   tr[2] = wrap16(s1r - s2r);
   ti[2] = wrap16(s1i - s2i);
   tr[3] = wrap16(d1r + rot90_re(d2i));
   ti[3] = wrap16(d1i + d2r);

   oRe[0] = tr[0];
   oIm[0] = ti[0];
   for (m = 1; m < 4; m++) {
      if (w == NULL) {
         oRe[m] = tr[m];
         oIm[m] = ti[m];
      }
      else {
         int64_t wr = w[2 * (m - 1)], wi = w[2 * (m - 1) + 1];
         oRe[m]     = q15_round_sat((int64_t) tr[m] * wr - (int64_t) ti[m] * wi);
         oIm[m]     = q15_round_sat((int64_t) tr[m] * wi + (int64_t) ti[m] * wr);
      }
   }
}

static void i16_core(int16_t *pX, const int16_t *pW, int16_t *pY, const uint32_t *pShift, uint32_t N, uint32_t B)
{
   uint32_t       M = N, stage = 0, shift = 0, k, l;
   const int16_t *pw = pW;

   while (M > 2u) {
      uint32_t q = M / 4u;
      shift      = pShift[stage];
      for (k = 0; k < N / 4u; k++) {
         uint32_t dft = k / q, point = k % q, base = dft * M + point;
         /* twiddle-free butterflies: the 4-point stage and point 0 of the
          * 8-point stage ($I16B/_cn.c:230) */
         int      noTw = (M == 4u) || (M == 8u && point == 0u);
         int16_t  w[6];
         int      m;
         for (m = 0; m < 3; m++) {
            w[2 * m]     = pw[2u * ((uint32_t) m * q + point)];
            w[2 * m + 1] = pw[2u * ((uint32_t) m * q + point) + 1u];
         }
         for (l = 0; l < B; l++) {
            int16_t *ch = pX + (size_t) l * N * 2u;
            int16_t  re[4], im[4], oRe[4], oIm[4];
            for (m = 0; m < 4; m++) {
               re[m] = ch[2u * (base + (uint32_t) m * q)];
               im[m] = ch[2u * (base + (uint32_t) m * q) + 1u];
            }
            bfly_i16(re, im, noTw ? NULL : w, oRe, oIm);
            /* the 8-point stage's shift is applied after the final radix-2
             * instead ($I16B/_cn.c:276-280) */
            if (M != 8u) {
               for (m = 0; m < 4; m++) {
                  oRe[m] = shift_round(oRe[m], shift);
                  oIm[m] = shift_round(oIm[m], shift);
               }
            }
            /* bit-swapped store of outputs 1 and 2 ($I16B/_cn.c:281-294) */
            ch[2u * base]                 = oRe[0];
            ch[2u * base + 1u]            = oIm[0];
            ch[2u * (base + q)]           = oRe[2];
            ch[2u * (base + q) + 1u]      = oIm[2];
            ch[2u * (base + 2u * q)]      = oRe[1];
            ch[2u * (base + 2u * q) + 1u] = oIm[1];
            ch[2u * (base + 3u * q)]      = oRe[3];
            ch[2u * (base + 3u * q) + 1u] = oIm[3];
         }
      }
      M /= 4u;
      pw += 2u * 3u * M;
      stage++;
   }
   if (M == 2u) {
      /* final radix-2 with the last radix-4 stage's shift ($I16B/_cn.c:302-323) */
      for (k = 0; k < N / 2u; k++) {
         for (l = 0; l < B; l++) {
            int16_t *a = pX + (size_t) l * N * 2u + 4u * k;
            int16_t  v0 = wrap16(a[0] + a[2]), v1 = wrap16(a[1] + a[3]);
            int16_t  v2 = wrap16(a[0] - a[2]), v3 = wrap16(a[1] - a[3]);
            a[0]        = shift_round(v0, shift);
            a[1]        = shift_round(v1, shift);
            a[2]        = shift_round(v2, shift);
            a[3]        = shift_round(v3, shift);
         }
      }
   }
   for (k = 0; k < N; k++) {
      uint32_t p = bit_reverse(k, N - 1u);
      for (l = 0; l < B; l++) {
         pY[(size_t) l * N * 2u + 2u * k]      = pX[(size_t) l * N * 2u + 2u * p];
         pY[(size_t) l * N * 2u + 2u * k + 1u] = pX[(size_t) l * N * 2u + 2u * p + 1u];
      }
   }
}

/* $I16B/_cn.c:191-338 */
FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                       FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                       FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                       FFTLIB_bufParams1D_t *bufParamsShift, uint32_t numPoints,
                                                       uint32_t numChannels)
{
   (void) bufParamsX;
   (void) bufParamsW;
   (void) bufParamsY;
   (void) bufParamsShift;
   i16_core(pX, pW, pY, pShift, numPoints, numChannels);
   return FFTLIB_SUCCESS;
}

/* $I16B/_cn.c:340-385. The last test accepts a twiddle table of 2*4^s or
 * 4*4^s elements for s = bufParamsShift->dim_x shift entries. */
FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_checkParams_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                   int16_t *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                   int16_t *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                   uint32_t *pShift,
                                                                   FFTLIB_bufParams1D_t *bufParamsShift,
                                                                   uint32_t numPoints, uint32_t numChannels)
{
   uint32_t pow4;
   if (pX == NULL || pW == NULL || pY == NULL || pShift == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsY->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->dim_x < numPoints * numChannels * 2u) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsW->dim_x != numPoints * 2u) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_INT16 || bufParamsW->data_type != FFTLIB_INT16 ||
       bufParamsY->data_type != FFTLIB_INT16)
      return FFTLIB_ERR_INVALID_TYPE;
   if (bufParamsShift->data_type != FFTLIB_UINT32) return FFTLIB_ERR_INVALID_TYPE;
   pow4 = 1u << ((bufParamsShift->dim_x * 2u) & 31u);
   if (pow4 * 2u != bufParamsW->dim_x && pow4 * 4u != bufParamsW->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   return FFTLIB_SUCCESS;
}

/* $I16/_cn.c: single channel, N = dim_x / 2 (products formed in int and
 * widened, which gives the same values as the batched kernel's int64). */
FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                FFTLIB_bufParams1D_t *bufParamsShift)
{
   (void) bufParamsW;
   (void) bufParamsY;
   (void) bufParamsShift;
   i16_core(pX, pW, pY, pShift, bufParamsX->dim_x >> 1, 1u);
   return FFTLIB_SUCCESS;
}

/* $I16/_cn.c checkParams: the batched checks with numChannels = 1 and
 * N = dim_x / 2. */
FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_checkParams_cn(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            int16_t *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            int16_t *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                            uint32_t *pShift, FFTLIB_bufParams1D_t *bufParamsShift)
{
   uint32_t pow4;
   if (pX == NULL || pW == NULL || pY == NULL || pShift == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsY->dim_x || bufParamsX->dim_x != bufParamsW->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_INT16 || bufParamsW->data_type != FFTLIB_INT16 ||
       bufParamsY->data_type != FFTLIB_INT16)
      return FFTLIB_ERR_INVALID_TYPE;
   if (bufParamsShift->data_type != FFTLIB_UINT32) return FFTLIB_ERR_INVALID_TYPE;
   pow4 = 1u << ((bufParamsShift->dim_x * 2u) & 31u);
   if (pow4 * 2u != bufParamsW->dim_x && pow4 * 4u != bufParamsW->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   return FFTLIB_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* fp32 real-to-complex / complex-to-real                                    */
/* ------------------------------------------------------------------------ */

/* Split post-processing of one channel ($R2CB/_cn.c:40-121). Z = pXFFT channel
 * (N/2 complex), A[i] = pSf[2i..], B[i] = pSf[2N-2i-2..]. */
static void r2c_split(const float *Z, const float *pSf, float *Y, uint32_t N, uint32_t nOut, int fullSpectrum)
{
   uint32_t i;
   for (i = 0; i < N / 2u; i++) {
      float v0r = Z[2u * i], v0i = Z[2u * i + 1u];
      float br = pSf[2u * N - 2u * i - 2u], bi = pSf[2u * N - 2u * i - 1u];
      float t0r = v0r * pSf[2u * i] - v0i * pSf[2u * i + 1u];
      float t0i = v0r * pSf[2u * i + 1u] + v0i * pSf[2u * i];
      float t1r, t1i;
      if (i == 0u) {
         /* i = 0 uses Z[0] with this sign pattern ($R2CB/_cn.c:73-78) */
         t1r = Z[0] * br + Z[1] * bi;
         t1i = Z[0] * br - Z[1] * bi;
      }
      else {
         float v1r = Z[N - 2u * i], v1i = Z[N - 2u * i + 1u];
         t1r       = v1r * br + v1i * bi;
         t1i       = v1r * bi - v1i * br;
      }
      Y[2u * i]      = t0r + t1r;
      Y[2u * i + 1u] = t0i + t1i;
      if (fullSpectrum && i > 0u) {
         /* Y[N-i] = conj(Y[i]) ($R2CB/_cn.c:102-105) */
         Y[2u * nOut - 2u * i]      = Y[2u * i];
         Y[2u * nOut - 2u * i + 1u] = Y[2u * i + 1u] * (-1.0f);
      }
   }
   /* Nyquist bin ($R2CB/_cn.c:113-118) */
   Y[N]      = Z[0] - Z[1];
   Y[N + 1u] = 0.0f;
}

/* $R2CB/_cn.c:10-125 */
FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                      FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                      FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                      FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                      FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                      uint32_t numPoints, uint32_t numChannels, uint8_t outmode)
{
   uint32_t nOut = (outmode == 1u) ? numPoints : (numPoints >> 1) + 1u;
   uint32_t ch;
   (void) bufParamsX;
   (void) bufParamsW;
   (void) bufParamsXFFT;
   (void) bufParamsSf;
   (void) bufParamsY;
   c2c_core(pX, pW, pXFFT, numPoints >> 1, numChannels, C2C_FORWARD);
   for (ch = 0; ch < numChannels; ch++) {
      r2c_split(pXFFT + (size_t) ch * numPoints, pSf, pY + (size_t) 2u * ch * nOut, numPoints, nOut, outmode == 1u);
   }
   return FFTLIB_SUCCESS;
}

/* $R2CB/_cn.c:127-177 */
FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                  FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                  FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                  uint32_t numPoints, uint32_t numChannels,
                                                                  uint8_t outmode)
{
   (void) numChannels;
   (void) outmode;
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsW->dim_x || bufParamsX->dim_x != bufParamsY->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   if (bufParamsX->dim_x & 1u) return FFTLIB_ERR_INVALID_DIMENSION;
   return is_235_smooth(numPoints / 2u) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

/* $R2C1/_cn.c:8-154: N = dim_x reals, full N-bin spectrum. The reference's
 * mirror step also stores Y[N] for i = 0, one element past a 2N-float output
 * buffer; that store lies outside the output and is not restated, and the
 * i = N/4 prelude it computes first is overwritten by its main loop. */
FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                               FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                               FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                               FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                               FFTLIB_bufParams1D_t *bufParamsY)
{
   uint32_t N = bufParamsX->dim_x;
   (void) bufParamsW;
   (void) bufParamsXFFT;
   (void) bufParamsSf;
   (void) bufParamsY;
   c2c_core(pX, pW, pXFFT, N >> 1, 1u, C2C_FORWARD);
   r2c_split(pXFFT, pSf, pY, N, N, 1);
   return FFTLIB_SUCCESS;
}

/* $R2C1/_cn.c:156-200 */
FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                           FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                           FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY)
{
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsW->dim_x || bufParamsX->dim_x != bufParamsY->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   if (bufParamsX->dim_x & 1u) return FFTLIB_ERR_INVALID_DIMENSION;
   return is_235_smooth(bufParamsX->dim_x >> 1) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

/* Split pre-processing of one channel ($C2RB/_cn.c:51-83): V[i] from
 * conj(X[N/2-i]) (inMode 0) or X[N/2+i] (inMode 1). */
static void c2r_split(const float *X, const float *pSf, float *Z, uint32_t N, int inMode)
{
   uint32_t i, h = N >> 1;
   for (i = 0; i < h; i++) {
      float v0r = X[2u * i], v0i = X[2u * i + 1u], v1r, v1i;
      float t0r, t0i, t1r, t1i;
      if (inMode == 0) {
         v1r = X[2u * (h - i)];
         v1i = -X[2u * (h - i) + 1u];
      }
      else {
         v1r = X[2u * i + N];
         v1i = X[2u * i + 1u + N];
      }
      t0r          = v0r * pSf[2u * i] - v0i * pSf[2u * i + 1u];
      t0i          = v0r * pSf[2u * i + 1u] + v0i * pSf[2u * i];
      t1r          = v1r * pSf[2u * i + N] - v1i * pSf[2u * i + 1u + N];
      t1i          = v1r * pSf[2u * i + 1u + N] + v1i * pSf[2u * i + N];
      Z[2u * i]      = t0r + t1r;
      Z[2u * i + 1u] = t0i + t1i;
   }
}

/* $C2RB/_cn.c:8-91 */
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                       FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                       FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                       FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                       FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                       uint32_t numPoints, uint32_t numChannels, uint8_t inMode)
{
   uint32_t stride = (inMode == 0u) ? 2u * ((numPoints >> 1) + 1u) : 2u * numPoints;
   uint32_t ch;
   (void) bufParamsX;
   (void) bufParamsW;
   (void) bufParamsXFFT;
   (void) bufParamsSf;
   (void) bufParamsY;
   for (ch = 0; ch < numChannels; ch++) {
      c2r_split(pX + (size_t) ch * stride, pSf, pXFFT + (size_t) ch * numPoints, numPoints, inMode);
   }
   c2c_core(pXFFT, pW, pY, numPoints / 2u, numChannels, C2C_INVERSE_PRESCALE);
   return FFTLIB_SUCCESS;
}

/* $C2RB/_cn.c:93-136 */
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                   FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                   FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                   uint32_t numPoints, uint32_t numChannels,
                                                                   uint8_t inMode)
{
   (void) numChannels;
   (void) inMode;
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsW->dim_x || bufParamsX->dim_x != bufParamsY->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   return is_235_smooth(numPoints / 2u) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

/* $C2R1/_cn.c:8-78: N = dim_x, full-spectrum input (inMode 1 form). */
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                                FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                                FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                                FFTLIB_bufParams1D_t *bufParamsY)
{
   uint32_t N = bufParamsX->dim_x;
   (void) bufParamsW;
   (void) bufParamsXFFT;
   (void) bufParamsSf;
   (void) bufParamsY;
   c2r_split(pX, pSf, pXFFT, N, 1);
   c2c_core(pXFFT, pW, pY, N >> 1, 1u, C2C_INVERSE_POSTDIV);
   return FFTLIB_SUCCESS;
}

/* $C2R1/_cn.c:80-118 */
FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_checkParams_cn(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY)
{
   if (pX == NULL || pW == NULL || pY == NULL) return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsW->dim_x || bufParamsX->dim_x != bufParamsY->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_FLOAT32 || bufParamsW->data_type != FFTLIB_FLOAT32 ||
       bufParamsY->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   return is_235_smooth(bufParamsX->dim_x >> 1) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

/* ------------------------------------------------------------------------ */
/* Table generators                                                         */
/* ------------------------------------------------------------------------ */

/* The reference's generators evaluate angles in double with PI = 3.141592654
 * (not M_PI) and round cos/-sin to float. */
static const double kPiRef = 3.141592654;

/* Mixed-radix C2C twiddles (test driver tw_gen, test/fft_c7x/
 * FFTLIB_fft1d_i32fc_c32fc_o32fc/..._d.c:31-97). For every stage of radix r
 * with sub-DFT length M (t = M/r), the block holds W_M^{m*i} for m = 1..r-1 at
 * complex offset (m-1)*t + i, i < t, where the angle of W_M^{m*i} is computed
 * as (2*m*PI*(i*j))/Ng with Ng the length of the radix group and j = Ng/M. */
void oracle_tw_gen_c2c(FFTLIB_F32 *pW, int32_t numPoints)
{
   int32_t k = 0, radIdx;
   static const int32_t radices[3] = {5, 3, 4};
   int32_t ng = numPoints;

   for (radIdx = 0; radIdx < 3; radIdx++) {
      int32_t r = radices[radIdx], stageLen, j, t;
      if (ng % r != 0) continue;
      stageLen = ng;
      t        = ng / r;
      for (j = 1; j <= ng / r && stageLen % r == 0; j *= r) {
         int32_t i;
         for (i = 0; i < ng / r; i += j) {
            int32_t m;
            for (m = 1; m < r; m++) {
               double ang            = 2.0 * (double) m * kPiRef * i / ng;
               pW[k + 2 * (m - 1) * t]     = (FFTLIB_F32) cos(ang);
               pW[k + 2 * (m - 1) * t + 1] = (FFTLIB_F32) -sin(ang);
            }
            k += 2;
         }
         k += t * 2 * (r - 2);
         t /= r;
         stageLen /= r;
      }
      if (r != 4) ng = stageLen;
   }
}

/* int16 radix-4 twiddles: floor(0.5 + 32767.5*v) saturated to int16
 * (test/fft_c7x/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/..._d.c:29-61). */
static int16_t to_q15(double v)
{
   double x = floor(0.5 + v);
   if (x >= 32767.0) return 32767;
   if (x <= -32768.0) return -32768;
   return (int16_t) x;
}

void oracle_tw_gen_i16sc(int16_t *pW, uint32_t numPoints)
{
   int32_t i, j, k = 0, t = (int32_t) (numPoints >> 2);
   for (j = 1; j <= (int32_t) (numPoints >> 2); j <<= 2) {
      for (i = 0; i < (int32_t) (numPoints >> 2); i += j) {
         int32_t m;
         for (m = 1; m < 4; m++) {
            double ang                  = 2.0 * (double) m * kPiRef * i / numPoints;
            pW[k + 2 * (m - 1) * t]     = to_q15(32767.5 * cos(ang));
            pW[k + 2 * (m - 1) * t + 1] = to_q15(32767.5 * (-sin(ang)));
         }
         k += 2;
      }
      k += t * 4;
      t >>= 2;
   }
}

/* N/2-point table for the real transforms ($R2CB/_tw_gen.cpp:9-78,
 * $C2RB/_tw_gen.cpp:9-79): the C2C table of length N/2. */
void oracle_tw_gen_r2c(FFTLIB_F32 *pW, int32_t numPoints) { oracle_tw_gen_c2c(pW, numPoints >> 1); }
void oracle_ifft_c2r_tw_gen(FFTLIB_F32 *pW, int32_t numPoints) { oracle_tw_gen_c2c(pW, numPoints >> 1); }

/* $R2CB/_tw_gen.cpp:80-107 */
void oracle_split_factor_gen_r2c(FFTLIB_F32 *pSf, int32_t numPoints)
{
   int32_t i;
   for (i = 0; i < numPoints >> 1; i++) {
      double ang                     = 2 * kPiRef * i / numPoints;
      pSf[2 * i]                     = (FFTLIB_F32) (0.5 * (1 - sin(ang)));
      pSf[2 * i + 1]                 = (FFTLIB_F32) (0.5 * (-cos(ang)));
      pSf[2 * numPoints - 2 * i - 2] = (FFTLIB_F32) (0.5 * (1 + sin(ang)));
      pSf[2 * numPoints - 2 * i - 1] = (FFTLIB_F32) (0.5 * (cos(ang)));
   }
}

/* $C2RB/_tw_gen.cpp:81-95 */
void oracle_ifft_c2r_split_factor_gen(FFTLIB_F32 *pSf, int32_t numPoints)
{
   int32_t i;
   for (i = 0; i < numPoints >> 1; i++) {
      double ang                 = 2 * kPiRef * i / numPoints;
      pSf[2 * i]                 = (FFTLIB_F32) (0.5 * (1 - sin(ang)));
      pSf[2 * i + 1]             = (FFTLIB_F32) (0.5 * (cos(ang)));
      pSf[2 * i + numPoints]     = (FFTLIB_F32) (0.5 * (1 + sin(ang)));
      pSf[2 * i + numPoints + 1] = (FFTLIB_F32) (0.5 * (-cos(ang)));
   }
}

/* ------------------------------------------------------------------------ */
/* All-cores CPU baseline: channel split over POSIX threads                 */
/* ------------------------------------------------------------------------ */

typedef struct {
   float         *x, *y;
   const float   *w;
   int16_t       *x16, *y16;
   const int16_t *w16;
   const uint32_t *shift;
   uint32_t       n, b, inverse;
} job_t;

static void *run_c2c_job(void *arg)
{
   job_t *j = (job_t *) arg;
   if (j->b) c2c_core(j->x, j->w, j->y, j->n, j->b, j->inverse ? C2C_INVERSE_PRESCALE : C2C_FORWARD);
   return NULL;
}

static void *run_i16_job(void *arg)
{
   job_t *j = (job_t *) arg;
   if (j->b) i16_core(j->x16, j->w16, j->y16, j->shift, j->n, j->b);
   return NULL;
}

static FFTLIB_STATUS run_split(job_t *proto, uint32_t numChannels, int nThreads, int isI16)
{
   pthread_t *th;
   job_t     *jobs;
   uint32_t   start = 0;
   int        t;
   if (nThreads < 1) nThreads = 1;
   th   = (pthread_t *) calloc((size_t) nThreads, sizeof(pthread_t));
   jobs = (job_t *) calloc((size_t) nThreads, sizeof(job_t));
   if (th == NULL || jobs == NULL) {
      free(th);
      free(jobs);
      return FFTLIB_ERR_FAILURE;
   }
   for (t = 0; t < nThreads; t++) {
      uint32_t cnt = numChannels / (uint32_t) nThreads + ((uint32_t) t < numChannels % (uint32_t) nThreads ? 1u : 0u);
      size_t   off = (size_t) start * proto->n * 2u;
      jobs[t]      = *proto;
      jobs[t].b    = cnt;
      if (isI16) {
         jobs[t].x16 = proto->x16 + off;
         jobs[t].y16 = proto->y16 + off;
         pthread_create(&th[t], NULL, run_i16_job, &jobs[t]);
      }
      else {
         jobs[t].x = proto->x + off;
         jobs[t].y = proto->y + off;
         pthread_create(&th[t], NULL, run_c2c_job, &jobs[t]);
      }
      start += cnt;
   }
   for (t = 0; t < nThreads; t++) pthread_join(th[t], NULL);
   free(th);
   free(jobs);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS oracle_fft1dBatched_c2c_threads(FFTLIB_F32 *pX, FFTLIB_F32 *pW, FFTLIB_F32 *pY, uint32_t numPoints,
                                              uint32_t numChannels, int inverse, int nThreads)
{
   job_t proto;
   memset(&proto, 0, sizeof(proto));
   proto.x       = pX;
   proto.y       = pY;
   proto.w       = pW;
   proto.n       = numPoints;
   proto.inverse = (uint32_t) (inverse != 0);
   return run_split(&proto, numChannels, nThreads, 0);
}

FFTLIB_STATUS oracle_fft1dBatched_i16_threads(int16_t *pX, int16_t *pW, int16_t *pY, uint32_t *pShift,
                                              uint32_t numShifts, uint32_t numPoints, uint32_t numChannels,
                                              int nThreads)
{
   job_t proto;
   (void) numShifts;
   memset(&proto, 0, sizeof(proto));
   proto.x16   = pX;
   proto.y16   = pY;
   proto.w16   = pW;
   proto.shift = pShift;
   proto.n     = numPoints;
   return run_split(&proto, numChannels, nThreads, 1);
}
