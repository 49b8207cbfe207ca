// Split-precision MFMA variants of the batched complex FFT: DFT-as-GEMM on
// the matrix cores for the lengths N = R1 * 16 * R3 (R1 <= 16, R3 in
// {1, 2, 4}) listed in FFTLIB_MFMA_LENGTHS below: 21 lengths from 32 to 1024.
//
// With C = 16 R3 columns, n = C a + b (a < R1, b < C), b = 16 b2 + b1
// (b1 < 16, b2 < R3) and k = c + R1 d2 + R1 R3 d1 (c < R1, d2 < R3, d1 < 16):
//
//   Y1[c][b]     = sum_a W_R1^(ac) x[C a + b]              radix-R1 stage (GEMM 1)
//   U[c][d2][b1] = sum_b2 W_R3^(b2 d2) W_N^(bc) Y1[c][b]    twiddle, radix-R3 stage (fp32 VALU)
//   X[k]         = sum_b1 W_16^(b1 d1) W_C^(b1 d2) U[c][d2][b1]
//                                                          twiddle, radix-16 stage (GEMM 2)
//
// (for N = 960: the radix-15 and radix-64 = 4 x 16 DFT stages). The outputs
// go through a per-wavefront LDS staging area so that every store
// instruction writes 64 consecutive points. Both DFT
// stages run in real form on v_mfma_f32_16x16x32_{f16,bf16} with FP32
// accumulation, and both are computed transposed: in GEMM 1 the data is the
// A operand (rows = the 16 columns b of a tile, K = the real and imaginary
// parts of x over a) and the DFT matrix the B operand; in GEMM 2 the DFT
// matrix is the A operand and the data the B operand (K = b1, columns = c).
// Tile mt of GEMM 1 holds the rows b = 16 mt + 4q + i (work item (q, r) =
// (lane / 16, lane % 16), i < 4) of column c = r, i.e. b2 = mt and b1 = 4q +
// i: the R3 inputs of every radix-R3 butterfly are in one work item, and after
// that stage tile d2 of the accumulators is exactly the B operand of GEMM 2's
// N-tile d2 with K index 8q + 2i + p = part p of b1 = 4q + i. The data
// therefore never leaves the registers between the stages, and each GEMM-2
// tile frees the accumulators it consumed. The input is read straight into
// the GEMM-1 operand layout (K index 8q + 2jj + p is part p of
// x[C (4q + jj) + b]: four 8-byte loads per tile, whole 128-byte lines per
// wavefront load). Besides the output staging, LDS holds the DFT-matrix
// fragments and the twiddles, shared by the workgroup.
//
// The data operands are split into sums of 16-bit parts so that the products
// recover FP32-level accuracy:
//   FP16x2: x = h + l (fp16), products h*h + h*l + l*h;
//   BF16x2: same with bf16 parts;
//   BF16x3: x = h + m + l (bf16; the data split by truncation, which is
//           exact, the matrices rounded), products hh + hm + mh + hl + lh
//           + mm (without mm the N = 960 error against FP64 is 2.6e-6 in a
//           CPU model of these products).
// FP16 parts are taken of the transform scaled by a power of two (exact) so
// that its largest input is in [0.5, 1): fp16 range then holds every operand.
// The twiddles and the radix-R3 stage are fp32.
//
// This is a different (but equivalent) algorithm from the reference's
// in-place mixed-radix DIF (src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/
// ..._cn.c:240-491): it does not read the caller's twiddle table; the DFT
// matrices and twiddles are generated here in double precision. It is
// selected with FFTLIB_GPU_setC2CVariant and its error against an FP64 DFT
// is a property of the split scheme.
//
// Work decomposition: 256-thread workgroups; each wavefront runs whole
// transforms (w, w + 4 gridDim.x, ...), loading tile mt of the next transform
// right after it has split tile mt of the current one.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <vector>

#include "fft_engine.h"
#include "fft_kernels.h"

namespace fftlib {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16   bf168 __attribute__((ext_vector_type(8)));
typedef float    f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int kWaves    = 4;
constexpr int kThreadsM = 64 * kWaves;

enum Split : int { SPLIT_FP16X2 = 0, SPLIT_BF16X2 = 1, SPLIT_BF16X3 = 2 };

template <int S>
struct SplitTraits {
   static constexpr int parts = S == SPLIT_BF16X3 ? 3 : 2;
};

// ---- operand splitting ------------------------------------------------------
__device__ __forceinline__ uint16_t f16_bits(float v)
{
   const _Float16 h = (_Float16) v;
   return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_value(uint16_t b) { return (float) __builtin_bit_cast(_Float16, b); }
__device__ __forceinline__ uint16_t bf16_bits(float v)
{
   const __bf16 h = (__bf16) v;
   return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float bf16_value(uint16_t b) { return (float) __builtin_bit_cast(__bf16, b); }

// Upper halves of two floats as one register (bf16 pair: a low, b high).
__device__ __forceinline__ uint32_t hi_halves(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x07060302u); }

// Splits 8 floats into the parts of scheme S (part 0 = high).
// BF16x3 splits by truncation, which is exact: h = x with the low 16 bits
// cleared, r = x - h (exact: the low 16 mantissa bits of x), m = r truncated
// the same way, l = r - m (exact, at most 8 significant bits, so a bf16
// value); each part is the upper half of its float. Two bit masks, two
// subtractions and 1.5 permutes per value, against 7.5 instructions for the
// rounding conversions the compiler emits for (__bf16) casts.
template <int S>
__device__ __forceinline__ void split8(const float (&v)[8], u16x8 (&p)[SplitTraits<S>::parts])
{
   if constexpr (S == SPLIT_BF16X3) {
      uint32_t ph[4], pm[4], pl[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
         uint32_t h[2], m[2], l[2];
#pragma unroll
         for (int k = 0; k < 2; ++k) {
            const float x  = v[2 * w + k];
            h[k]           = __float_as_uint(x) & 0xFFFF0000u;
            const float rr = x - __uint_as_float(h[k]);
            m[k]           = __float_as_uint(rr) & 0xFFFF0000u;
            l[k]           = __float_as_uint(rr - __uint_as_float(m[k]));
            h[k]           = __float_as_uint(x);
            m[k]           = __float_as_uint(rr);
         }
         ph[w] = hi_halves(h[0], h[1]);
         pm[w] = hi_halves(m[0], m[1]);
         pl[w] = hi_halves(l[0], l[1]);
      }
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      p[0] = __builtin_bit_cast(u16x8, (u32x4){ph[0], ph[1], ph[2], ph[3]});
      p[1] = __builtin_bit_cast(u16x8, (u32x4){pm[0], pm[1], pm[2], pm[3]});
      p[2] = __builtin_bit_cast(u16x8, (u32x4){pl[0], pl[1], pl[2], pl[3]});
      return;
   }
#pragma unroll
   for (int i = 0; i < 8; ++i) {
      if constexpr (S == SPLIT_FP16X2) {
         const uint16_t h = f16_bits(v[i]);
         p[0][i]          = h;
         p[1][i]          = f16_bits(v[i] - f16_value(h));
      }
      else {
         const uint16_t h = bf16_bits(v[i]);
         const float    r = v[i] - bf16_value(h);
         p[0][i]          = h;
         if constexpr (S == SPLIT_BF16X2) {
            p[1][i] = bf16_bits(r);
         }
         else {
            const uint16_t m = bf16_bits(r);
            p[1][i]          = m;
            p[2][i]          = bf16_bits(r - bf16_value(m));
         }
      }
   }
}

template <int S>
__device__ __forceinline__ f32x4 mma(u16x8 a, u16x8 b, f32x4 c)
{
   if constexpr (S == SPLIT_FP16X2)
      return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a), __builtin_bit_cast(half8, b), c, 0,
                                                    0, 0);
   else
      return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf168, a), __builtin_bit_cast(bf168, b), c,
                                                     0, 0, 0);
}

// Sum of the split products that the scheme keeps (smallest terms first).
template <int S>
__device__ __forceinline__ f32x4 split_mma(const u16x8 (&a)[SplitTraits<S>::parts],
                                           const u16x8 (&b)[SplitTraits<S>::parts], f32x4 c)
{
   if constexpr (S == SPLIT_BF16X3) {
      c = mma<S>(a[1], b[1], c);
      c = mma<S>(a[0], b[2], c);
      c = mma<S>(a[2], b[0], c);
   }
   c = mma<S>(a[0], b[1], c);
   c = mma<S>(a[1], b[0], c);
   return mma<S>(a[0], b[0], c);
}

// ---- the kernel ---------------------------------------------------------------
struct MfmaArgs {
   const float2   *x;
   float2         *y;
   const uint16_t *frag;      // GEMM-1 B fragments of every part, then GEMM-2 A fragments of every part
   const float2   *twN;       // twiddle rows: W_N^(r b) at r (C + 1) + b (r < 16, b < C), then W_C^(b1 d2) at 16 (C + 1) + 16 d2 + b1
   uint32_t        numChannels;
   int             mode;      // C2CMode
};

// Wave-wide maximum of v over the lanes.
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
   for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
   return v;
}

// Orders this wavefront's LDS stores before its other lanes' later LDS
// accesses: the LDS completes one wavefront's operations in order, and the
// fences and the wave barrier keep the compiler from moving or forwarding
// loads and stores across this point.
__device__ __forceinline__ void wave_lds_sync()
{
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2 cmul(float2 a, float2 w)
{
   return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// Halves per fragment table: 2 tiles x parts x 64 lanes x 8.
template <int S>
constexpr int frag_halves()
{
   return 2 * SplitTraits<S>::parts * 64 * 8;
}

// Twiddle entries in LDS: the rows W_N^(r b) (r < 16, b < C = 16 R3) with a
// stride of C + 1 entries, so that each work item reads its 4 R3 twiddles at
// one base plus immediate offsets and the 16 rows of a read fall in distinct
// banks (a stride of C, a multiple of the bank count, put them all in one);
// then W_C^(b1 d2) at 16 d2 + b1.
template <int R3>
constexpr int mfma_tw_row()
{
   return 16 * R3 + 1;
}
template <int R3>
constexpr int mfma_tw_entries()
{
   return 16 * mfma_tw_row<R3>() + 16 * R3;
}

// Outputs are staged in LDS per wavefront and stored 64 consecutive points
// per instruction: stored straight from registers (runs of R1 points, not
// line aligned for R1 = 15) N = 960 took 360-405 us instead of 191-198 us.

template <int S, int R1, int R3>
constexpr size_t mfma_lds_bytes()
{
   return (size_t) 2 * frag_halves<S>() * sizeof(uint16_t) +
          ((size_t) mfma_tw_entries<R3>() + (size_t) R1 * 16 * R3 * kWaves) * sizeof(float2);
}

// Wavefronts per SIMD the register allocation targets: 3 for the shapes
// with two or four tiles (with 4, the four-tile shapes spill 90-170 bytes per
// work item and ran 1.5x slower on MI355X), 4 for one tile.
template <int R3>
constexpr int mfma_waves_per_eu()
{
   return R3 == 1 ? 4 : 3;
}

template <int S, int R1, int R3>
__global__ void __launch_bounds__(kThreadsM) __attribute__((amdgpu_waves_per_eu(mfma_waves_per_eu<R3>())))
c2c_mfma_kernel(MfmaArgs a)
{
   constexpr int P = SplitTraits<S>::parts;
   constexpr int C = 16 * R3;
   constexpr int N   = R1 * C;
   constexpr int kTw = mfma_tw_entries<R3>();
   extern __shared__ __align__(16) unsigned char ldsm[];
   u16x8  *f1 = (u16x8 *) ldsm;   // [t][part][lane]
   u16x8  *f2 = f1 + 2 * P * 64;
   float2 *tw = (float2 *) (f2 + 2 * P * 64);
   for (uint32_t i = threadIdx.x; i < (uint32_t) (4 * P * 64); i += kThreadsM) f1[i] = ((const u16x8 *) a.frag)[i];
   for (uint32_t i = threadIdx.x; i < (uint32_t) kTw; i += kThreadsM) tw[i] = a.twN[i];
   __syncthreads();

   const uint32_t lane = threadIdx.x % 64;
   const uint32_t wave = threadIdx.x / 64;
   const uint32_t q    = lane / 16, r = lane % 16;
   const bool     live = r < (uint32_t) R1;   // column c = r exists (the others are padding)
   constexpr int  kRow = mfma_tw_row<R3>();
   const float2  *twr  = tw + (live ? r * kRow : 0u) + 4u * q;   // W_N^(r (16 mt + 4q + i)) at 16 mt + i
   const float2  *twc  = tw + 16 * kRow + 4u * q;                // W_C^((4q + i) d2) at 16 d2 + i

   const float sxIn  = a.mode == C2C_INVERSE_PRESCALE ? 1.0f / (float) N : 1.0f;
   const float syIn  = a.mode == C2C_FORWARD ? sxIn : -sxIn;
   const float syOut = a.mode == C2C_INVERSE_PRESCALE ? -1.0f : 1.0f;
   const bool  pdiv  = a.mode == C2C_INVERSE_POSTDIV;

   // x[C a + 16 mt + r] of tile mt, a = 4q + jj (zero for a >= R1)
   auto load_tile = [&](float2(&v)[4], const float2 *xf, int mt) __attribute__((always_inline)) {
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
         const uint32_t aa = 4u * q + (uint32_t) jj;
         v[jj] = aa < (uint32_t) R1 ? engine::load_stream(xf + C * aa + 16 * mt + r) : make_float2(0.0f, 0.0f);
      }
   };

   const uint32_t stride = gridDim.x * kWaves;
   uint32_t       f      = blockIdx.x * kWaves + wave;
   float2         v[R3][4];
   if (f < a.numChannels) {
#pragma unroll
      for (int mt = 0; mt < R3; ++mt) load_tile(v[mt], a.x + (size_t) f * N, mt);
   }
   for (; f < a.numChannels; f += stride) {
      const uint32_t fn   = f + stride;
      const bool     more = fn < a.numChannels;
      // ---- input scaling (inverse modes) and the fp16 range scale -----------------
      float scale = 1.0f, unscale = 1.0f;
      if (a.mode != C2C_FORWARD) {
#pragma unroll
         for (int mt = 0; mt < R3; ++mt)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) v[mt][jj] = make_float2(sxIn * v[mt][jj].x, syIn * v[mt][jj].y);
      }
      if constexpr (S == SPLIT_FP16X2) {
         float amax = 0.0f;
#pragma unroll
         for (int mt = 0; mt < R3; ++mt)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) amax = fmaxf(amax, fmaxf(fabsf(v[mt][jj].x), fabsf(v[mt][jj].y)));
         amax = wave_max(amax);
         if (amax > 0.0f && amax < 1e30f) {
            int e;
            (void) frexpf(amax, &e);
            scale   = ldexpf(1.0f, -e);
            unscale = ldexpf(1.0f, e);
         }
      }
      // ---- GEMM 1 (transposed): rows b = 16 mt + (0..15), K = parts of x over a ---
      f32x4 acc1[R3][2];
      {
         u16x8 F1[2][P];
#pragma unroll
         for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int p = 0; p < P; ++p) F1[t][p] = f1[(t * P + p) * 64 + lane];
#pragma unroll
         for (int mt = 0; mt < R3; ++mt) {
            float vv[8];
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) vv[2 * jj] = scale * v[mt][jj].x, vv[2 * jj + 1] = scale * v[mt][jj].y;
            u16x8 A[P];
            split8<S>(vv, A);
            if (more) load_tile(v[mt], a.x + (size_t) fn * N, mt);   // tile mt of the next transform
#pragma unroll
            for (int t = 0; t < 2; ++t) acc1[mt][t] = split_mma<S>(A, F1[t], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            __builtin_amdgcn_sched_barrier(0);   // one tile at a time: bounds the live registers
         }
      }
      // ---- twiddle W_N^(b c), radix-R3 stage over b2, twiddle W_C^(b1 d2) ----------
      // Element i of tile mt is row b = 16 mt + 4q + i of column c = r, so
      // b2 = mt and b1 = 4q + i: the R3 inputs of each radix-R3 butterfly sit in
      // one work item, and the stage runs in place in acc1.
#pragma unroll
      for (int i = 0; i < 4; ++i) {
         float2 t[R3];
#pragma unroll
         for (int mt = 0; mt < R3; ++mt)
            t[mt] = cmul(make_float2(acc1[mt][0][i], acc1[mt][1][i]), twr[16 * mt + i]);
         float2 z[R3];
         if constexpr (R3 == 1) {
            z[0] = t[0];
         }
         else if constexpr (R3 == 2) {
            z[0] = make_float2(t[0].x + t[1].x, t[0].y + t[1].y);
            z[1] = make_float2(t[0].x - t[1].x, t[0].y - t[1].y);
         }
         else {
            const float2 s02 = make_float2(t[0].x + t[2].x, t[0].y + t[2].y);
            const float2 d02 = make_float2(t[0].x - t[2].x, t[0].y - t[2].y);
            const float2 s13 = make_float2(t[1].x + t[3].x, t[1].y + t[3].y);
            const float2 d13 = make_float2(t[1].x - t[3].x, t[1].y - t[3].y);
            z[0]             = make_float2(s02.x + s13.x, s02.y + s13.y);
            z[1]             = make_float2(d02.x + d13.y, d02.y - d13.x);   // d02 - i d13
            z[2]             = make_float2(s02.x - s13.x, s02.y - s13.y);
            z[3]             = make_float2(d02.x - d13.y, d02.y + d13.x);   // d02 + i d13
         }
#pragma unroll
         for (int d2 = 0; d2 < R3; ++d2) {
            if (d2 > 0) z[d2] = cmul(z[d2], twc[16 * d2 + i]);
            acc1[d2][0][i] = z[d2].x;
            acc1[d2][1][i] = z[d2].y;
         }
      }
      __builtin_amdgcn_sched_barrier(0);
      // ---- GEMM 2 (transposed), N-tile d2: K index 8q + 2e + p = part p of element
      // e (b1 = 4q + e) of tile d2; tile d2 of acc1 is dead once split ---------------
      f32x4 acc2[R3][2];
      u16x8 F2[2][P];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
         for (int p = 0; p < P; ++p) F2[t][p] = f2[(t * P + p) * 64 + lane];
#pragma unroll
      for (int d2 = 0; d2 < R3; ++d2) {
         float vv[8];
#pragma unroll
         for (int e = 0; e < 4; ++e) vv[2 * e] = acc1[d2][0][e], vv[2 * e + 1] = acc1[d2][1][e];
         u16x8 B[P];
         split8<S>(vv, B);
#pragma unroll
         for (int t = 0; t < 2; ++t) acc2[d2][t] = split_mma<S>(F2[t], B, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
         __builtin_amdgcn_sched_barrier(0);   // one tile at a time: bounds the live registers
      }
      // ---- store X[r + R1 (d2 + R3 d1)] (d1 = 4q + i: row of the GEMM-2 output) ------
      // through this wavefront's LDS staging; the inverse modes conjugate, POSTDIV
      // divides by N as a product with the fp32 reciprocal.
      float2 *yo = a.y + (size_t) f * N;
      float2 *zs = tw + kTw + wave * N;
      const float ox = pdiv ? unscale * (1.0f / (float) N) : unscale;
      const float oy = pdiv ? -ox : syOut * unscale;
      if (live) {
#pragma unroll
         for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int d2 = 0; d2 < R3; ++d2)
               zs[r + R1 * (d2 + R3 * (4 * q + i))] = make_float2(ox * acc2[d2][0][i], oy * acc2[d2][1][i]);
      }
      wave_lds_sync();
#pragma unroll
      for (int j = 0; j < (N + 63) / 64; ++j)
         if (N % 64 == 0 || lane + 64u * j < (uint32_t) N) engine::store_stream(yo + lane + 64 * j, zs[lane + 64 * j]);
      wave_lds_sync();   // the next transform's staging stores come after these loads
   }
}

// ---- lengths ------------------------------------------------------------------
// Every supported length with its factors (R1, R3): the largest R1 <= 16
// (GEMM 1 uses 2 R1 of its 32 K slots; N = 512 as 8 x 16 x 4 ran at 1.5x the
// time of 16 x 16 x 2).
#define FFTLIB_MFMA_LENGTHS(X)                                                                                      \
   X(32, 2, 1) X(48, 3, 1) X(64, 4, 1) X(80, 5, 1) X(96, 6, 1) X(128, 8, 1) X(144, 9, 1) X(160, 10, 1)             \
   X(192, 12, 1) X(240, 15, 1) X(256, 16, 1) X(288, 9, 2) X(320, 10, 2) X(384, 12, 2) X(480, 15, 2) X(512, 16, 2) \
   X(576, 9, 4) X(640, 10, 4) X(768, 12, 4) X(960, 15, 4) X(1024, 16, 4)

bool mfma_factors(uint32_t n, int *r1, int *r3)
{
   switch (n) {
#define FFTLIB_MFMA_CASE(L, A, B)                                                                                   \
   case L: *r1 = A, *r3 = B; return true;
      FFTLIB_MFMA_LENGTHS(FFTLIB_MFMA_CASE)
#undef FFTLIB_MFMA_CASE
   default: return false;
   }
}

// ---- host side: fragment tables ------------------------------------------------
uint16_t host_f16(float v)
{
   const _Float16 h = (_Float16) v;
   return __builtin_bit_cast(uint16_t, h);
}
float host_f16v(uint16_t b) { return (float) __builtin_bit_cast(_Float16, b); }
uint16_t host_bf16(float v)
{
   // round to nearest even (values here are finite)
   uint32_t u = __builtin_bit_cast(uint32_t, v);
   u += 0x7FFFu + ((u >> 16) & 1u);
   return (uint16_t) (u >> 16);
}
float host_bf16v(uint16_t b) { return __builtin_bit_cast(float, (uint32_t) b << 16); }

void host_split(int S, double v, uint16_t *parts)
{
   const float f = (float) v;
   if (S == SPLIT_FP16X2) {
      parts[0] = host_f16(f);
      parts[1] = host_f16((float) (v - (double) host_f16v(parts[0])));
   }
   else {
      parts[0]       = host_bf16(f);
      const double r = v - (double) host_bf16v(parts[0]);
      parts[1]       = host_bf16((float) r);
      if (S == SPLIT_BF16X3) parts[2] = host_bf16((float) (r - (double) host_bf16v(parts[1])));
   }
}

// Real-form DFT coefficient: part t of sum_j W_m^(jk) z[j] contributed by
// part p of z[j] (jk = the exponent j k).
double real_form(int t, int p, int jk, int m)
{
   const double pi  = 3.14159265358979323846;
   const double ang = -2.0 * pi * (double) (jk % m) / (double) m;
   const double wr = cos(ang), wi = sin(ang);
   if (t == 0) return p == 0 ? wr : -wi;
   return p == 0 ? wi : wr;
}

// Fragment tables of one (split, R1, R3) in MFMA operand layout for lane
// l = 16 q + r, tile t and half j (e = j >> 1, p = j & 1):
//   GEMM-1 B operand: B[k = 8q + j][n = 16t + r] = coefficient of part p of
//     x[a = 4q + e] in part t of Y1[c = r] (0 when a or c is at least R1);
//   GEMM-2 A operand: A[m = 16t + r][k = 8q + j] = coefficient of part p of
//     Z[b1 = 4q + e] in part t of V[d1 = r].
// Layout: [table][t][part][lane][8 halves].
std::vector<uint16_t> build_frags(int S, int R1, int R3)
{
   const int             P = S == SPLIT_BF16X3 ? 3 : 2;
   std::vector<uint16_t> h((size_t) 2 * 2 * P * 64 * 8);
   uint16_t              parts[3];
   for (int tab = 0; tab < 2; ++tab)
      for (int t = 0; t < 2; ++t)
         for (int l = 0; l < 64; ++l)
            for (int j = 0; j < 8; ++j) {
               const int q = l / 16, r = l % 16, p = j & 1, e = j >> 1;
               double    v = 0.0;
               if (tab == 0) {
                  const int aa = 4 * q + e;
                  if (aa < R1 && r < R1) v = real_form(t, p, aa * r, R1);
               }
               else {
                  v = real_form(t, p, (4 * q + e) * r, 16);
               }
               host_split(S, v, parts);
               for (int pp = 0; pp < P; ++pp) h[((((size_t) tab * 2 + t) * P + pp) * 64 + l) * 8 + j] = parts[pp];
            }
   return h;
}

struct TableKey {
   int      dev, split;
   uint32_t n;
   bool     operator<(const TableKey &o) const
   {
      return dev != o.dev ? dev < o.dev : split != o.split ? split < o.split : n < o.n;
   }
};
struct DeviceTables {
   uint16_t *frag = nullptr;
   float2   *twN  = nullptr;
};
std::mutex                       g_tabMutex;
std::map<TableKey, DeviceTables> g_tables;   // device tables live until the process ends

hipError_t get_tables(int S, uint32_t n, int R1, int R3, const uint16_t **frag, const float2 **twN)
{
   int        dev = 0;
   hipError_t e   = hipGetDevice(&dev);
   if (e != hipSuccess) return e;
   std::lock_guard<std::mutex> lock(g_tabMutex);
   auto it = g_tables.find(TableKey{dev, S, n});
   if (it == g_tables.end()) {
      const std::vector<uint16_t> h = build_frags(S, R1, R3);
      const uint32_t              C = 16u * (uint32_t) R3, row = C + 1u, entries = 16u * row + 16u * (uint32_t) R3;
      std::vector<float2>         w(entries);
      const double                pi = 3.14159265358979323846;
      auto                        wn = [&](uint32_t m) {
         const double ang = -2.0 * pi * (double) (m % n) / (double) n;
         return make_float2((float) cos(ang), (float) sin(ang));
      };
      for (uint32_t r = 0; r < 16; ++r)
         for (uint32_t b = 0; b < C; ++b) w[r * row + b] = wn(r * b);
      for (uint32_t d2 = 0; d2 < (uint32_t) R3; ++d2)
         for (uint32_t b1 = 0; b1 < 16; ++b1) w[16 * row + 16 * d2 + b1] = wn((uint32_t) R1 * b1 * d2);
      DeviceTables t;
      if ((e = hipMalloc((void **) &t.frag, h.size() * sizeof(uint16_t))) != hipSuccess) return e;
      if ((e = hipMalloc((void **) &t.twN, entries * sizeof(float2))) != hipSuccess) {
         (void) hipFree(t.frag);
         return e;
      }
      if ((e = hipMemcpy(t.frag, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice)) != hipSuccess ||
          (e = hipMemcpy(t.twN, w.data(), entries * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) {
         (void) hipFree(t.frag);
         (void) hipFree(t.twN);
         return e;
      }
      it = g_tables.emplace(TableKey{dev, S, n}, t).first;
   }
   *frag = it->second.frag;
   *twN  = it->second.twN;
   return hipSuccess;
}

// Grid: `perWave` = transforms per wavefront the grid aims for (0 = the
// resident grid, every wavefront looping over numChannels / (resident
// wavefronts) transforms).
template <int S, int R1, int R3>
hipError_t launch_form(const MfmaArgs &a, uint32_t perWave, hipStream_t stream)
{
   auto             kern     = c2c_mfma_kernel<S, R1, R3>;
   constexpr size_t lds      = mfma_lds_bytes<S, R1, R3>();
   static int       resident = 0;
   if (resident == 0) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
         cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kThreadsM, lds) != hipSuccess || per <= 0) per = 1;
      resident = cus * per;
   }
   const uint32_t need = (a.numChannels + kWaves - 1) / kWaves;
   uint32_t       grid = need < (uint32_t) resident ? need : (uint32_t) resident;
   if (perWave != 0 && (need + perWave - 1u) / perWave > grid) grid = (need + perWave - 1u) / perWave;
   hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreadsM), lds, stream, a);
   return hipGetLastError();
}

template <int S>
hipError_t launch_split(uint32_t n, const MfmaArgs &a, uint32_t perWave, hipStream_t stream)
{
   switch (n) {
#define FFTLIB_MFMA_CASE(L, A, B)                                                                                   \
   case L: return launch_form<S, A, B>(a, perWave, stream);
      FFTLIB_MFMA_LENGTHS(FFTLIB_MFMA_CASE)
#undef FFTLIB_MFMA_CASE
   default: return hipErrorInvalidValue;
   }
}

// Transforms per wavefront the launches aim for (0 = the resident grid):
// about 4 per wavefront ran 2-7% faster than the resident grid at 240-960
// points (N = 960 BF16x3: 198.4 -> 191.0 us, FP16x2: 183.3 -> 172.9 us).
constexpr uint32_t kMfmaPerWave = 4;

}  // namespace

bool mfma_c2c_supported(uint32_t n)
{
   int r1, r3;
   return mfma_factors(n, &r1, &r3);
}

hipError_t launch_c2c_mfma_grid(const float *x, float *y, uint32_t n, uint32_t numChannels, int mode, int variant,
                                uint32_t perWave, hipStream_t stream)
{
   int R1 = 0, R3 = 0;
   if (!mfma_factors(n, &R1, &R3)) return hipErrorInvalidValue;
   if (numChannels == 0) return hipSuccess;
   int S;
   switch (variant) {
   case C2C_VARIANT_BF16X3: S = SPLIT_BF16X3; break;
   case C2C_VARIANT_FP16X2: S = SPLIT_FP16X2; break;
   case C2C_VARIANT_BF16X2: S = SPLIT_BF16X2; break;
   default: return hipErrorInvalidValue;
   }
   MfmaArgs   a{};
   hipError_t e = get_tables(S, n, R1, R3, &a.frag, &a.twN);
   if (e != hipSuccess) return e;
   a.x = (const float2 *) x, a.y = (float2 *) y, a.numChannels = numChannels, a.mode = mode;
   if (S == SPLIT_FP16X2) return launch_split<SPLIT_FP16X2>(n, a, perWave, stream);
   if (S == SPLIT_BF16X2) return launch_split<SPLIT_BF16X2>(n, a, perWave, stream);
   return launch_split<SPLIT_BF16X3>(n, a, perWave, stream);
}

hipError_t launch_c2c_mfma(const float *x, float *y, uint32_t n, uint32_t numChannels, int mode, int variant,
                           hipStream_t stream)
{
   return launch_c2c_mfma_grid(x, y, n, numChannels, mode, variant, kMfmaPerWave, stream);
}

}  // namespace fftlib
