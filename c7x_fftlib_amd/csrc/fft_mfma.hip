// Split-precision MFMA variant of the batched complex FFT, N = 960.
//
// The 960-point DFT is computed as two dense DFT stages on the matrix cores,
// with n = 64a + b (a < 15, b < 64) and k = c + 15d (c < 15, d < 64):
//
//   Y1[c][b] = sum_a W15^(ac) x[64a + b]            radix-15 stage (GEMM 1)
//   Y2[c][b] = W960^(bc) Y1[c][b]                    twiddle (fp32 VALU)
//   X[c + 15d] = sum_b Y2[c][b] W64^(bd)             radix-64 stage (GEMM 2)
//
// Both GEMMs run in real form on v_mfma_f32_16x16x32_{f16,bf16} with FP32
// accumulation. Operands are split into sums of 16-bit parts so the products
// recover FP32-level accuracy:
//   FP16x2: x = h + l (fp16), products h*h + h*l + l*h;
//   BF16x2: same with bf16 parts;
//   BF16x3: x = h + m + l (bf16), products hh + hm + mh + hl + lh + mm.
// FP16 parts are taken of the transform scaled by a power of two (exact) so
// that its largest input is near 1: fp16 range then holds every operand.
//
// This is a different (but equivalent) algorithm from the reference's
// in-place mixed-radix DIF (src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/
// ..._cn.c:240-491): it does not read the caller's twiddle table; the DFT
// matrices and W960 twiddles are generated here in double precision. It is
// selected with FFTLIB_GPU_setC2CVariant and its error against an FP64 DFT
// is a property of the split scheme.
//
// Work decomposition: 256-thread workgroups, persistent; each wavefront owns
// one 8.25 KB LDS region and runs whole transforms in it (transform
// w, w + 4*gridDim.x, ...), prefetching the next transform's input into
// registers while the current one goes through the two GEMMs. The DFT-matrix
// fragments (already in MFMA operand layout) and the twiddle table are
// copied to LDS once per workgroup.
#include <hip/hip_runtime.h>

#include <math.h>
#include <mutex>
#include <stdint.h>

#include "fft_kernels.h"

namespace fftlib {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16   bf168 __attribute__((ext_vector_type(8)));
typedef float    f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int      kN        = 960;
constexpr int      kWaves    = 4;
constexpr int      kThreadsM = 64 * kWaves;
constexpr uint32_t kXRow     = 66;                 // complex slots per staged input row (64 + pad)
constexpr uint32_t kY2Row    = 132;                // floats per Y2 row (128 + pad)
constexpr uint32_t kRegion   = 16 * kY2Row;        // floats per wave region (>= 15*kXRow*2 and 2*960)
constexpr int      kFragA1   = 2 * 64 * 8;         // halves of the two radix-15 A fragments, per part
constexpr int      kFragB2   = 2 * 2 * 4 * 64 * 8; // halves of the radix-64 B fragments (Fr, Fi), per part
static_assert(kRegion >= 15 * kXRow * 2 && kRegion >= 2 * kN, "wave region too small");

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

// Splits 8 floats into the parts of scheme S (part 0 = high).
template <int S>
__device__ __forceinline__ void split8(const float (&v)[8], u16x8 (&p)[SplitTraits<S>::parts])
{
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

// Sum of the split products that the scheme keeps.
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

__device__ __forceinline__ u16x8 neg8(u16x8 v) { return v ^ (uint16_t) 0x8000u; }

// ---- the kernel ---------------------------------------------------------------
struct MfmaArgs {
   const float2   *x;
   float2         *y;
   const uint16_t *frag;      // [part][A1 | B2] fragment tables
   const float2   *tw960;     // W960^m, m < 960
   uint32_t        numChannels;
   int             mode;      // C2CMode
};

// Wave-wide maximum of |v| over the lanes.
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
   for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
   return v;
}

template <int S>
__global__ void __launch_bounds__(kThreadsM) c2c960_mfma_kernel(MfmaArgs a)
{
   constexpr int P = SplitTraits<S>::parts;
   extern __shared__ float smf[];
   uint16_t *fragA1 = (uint16_t *) smf;                          // P * kFragA1 halves
   uint16_t *fragB2 = fragA1 + P * kFragA1;                       // P * kFragB2 halves
   float2   *tw     = (float2 *) (fragB2 + P * kFragB2);          // 960 complex
   float    *region = (float *) (tw + kN) + (threadIdx.x / 64) * kRegion;
   float2   *rc     = (float2 *) region;

   const uint32_t lane = threadIdx.x % 64;
   const uint32_t wave = threadIdx.x / 64;
   const uint32_t q    = lane / 16, r16 = lane % 16;

   // constants to LDS (once per workgroup)
   {
      const uint32_t *src   = (const uint32_t *) a.frag;
      uint32_t       *dst   = (uint32_t *) fragA1;
      const uint32_t  words = (uint32_t) P * (kFragA1 + kFragB2) / 2u;
      for (uint32_t i = threadIdx.x; i < words; i += kThreadsM) dst[i] = src[i];
      for (uint32_t i = threadIdx.x; i < (uint32_t) kN; i += kThreadsM) tw[i] = a.tw960[i];
      __syncthreads();
   }
   // radix-15 A fragments stay in registers
   u16x8 A1[2][P];
#pragma unroll
   for (int p = 0; p < P; ++p)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) A1[mt][p] = *(const u16x8 *) (fragA1 + p * kFragA1 + (mt * 64 + lane) * 8);

   const float sxIn  = a.mode == C2C_INVERSE_PRESCALE ? 1.0f / (float) kN : 1.0f;
   const float syIn  = a.mode == C2C_FORWARD ? sxIn : -sxIn;
   const float syOut = a.mode == C2C_INVERSE_PRESCALE ? -1.0f : 1.0f;
   const bool  pdiv  = a.mode == C2C_INVERSE_POSTDIV;

   const uint32_t stride = gridDim.x * kWaves;
   uint32_t       f      = blockIdx.x * kWaves + wave;
   float2         xin[15];
   if (f < a.numChannels) {
#pragma unroll
      for (int j = 0; j < 15; ++j) xin[j] = a.x[(size_t) f * kN + lane + 64 * j];
   }
   for (; f < a.numChannels; f += stride) {
      // ---- stage the input (x[64a + b] at row a, column b), scaled ----------
      float amax = 0.0f;
#pragma unroll
      for (int j = 0; j < 15; ++j) {
         const float2 v = make_float2(sxIn * xin[j].x, syIn * xin[j].y);
         amax           = fmaxf(amax, fmaxf(fabsf(v.x), fabsf(v.y)));
         rc[j * kXRow + lane] = v;
      }
      // Power-of-two scale bringing max|x| into [0.5, 1) for the fp16 parts
      // (1 when the transform is zero or for the bf16 schemes).
      float scale = 1.0f, unscale = 1.0f;
      if constexpr (S == SPLIT_FP16X2) {
         amax = wave_max(amax);
         if (amax > 0.0f && amax < 1e30f) {
            int e;
            (void) frexpf(amax, &e);
            scale   = ldexpf(1.0f, -e);
            unscale = ldexpf(1.0f, e);
         }
      }
      // ---- prefetch the next transform --------------------------------------
      const uint32_t fn = f + stride;
      if (fn < a.numChannels) {
#pragma unroll
         for (int j = 0; j < 15; ++j) xin[j] = a.x[(size_t) fn * kN + lane + 64 * j];
      }
      // ---- GEMM 1: C1 (32 x 64) = A1 (32 x 32) * B1 (32 x 64) ----------------
      // B1 row k: k < 15 -> Re x[a = k][b], 16 <= k < 31 -> Im x[a = k - 16][b].
      f32x4 acc1[2][4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
         const uint32_t b = 16 * nt + r16;
         float          v[8];
#pragma unroll
         for (int i = 0; i < 8; ++i) {
            const uint32_t aa = (q & 1u) * 8u + (uint32_t) i;
            const float2   z  = aa < 15u ? rc[aa * kXRow + b] : make_float2(0.0f, 0.0f);
            v[i]              = scale * ((q < 2u) ? z.x : z.y);
         }
         u16x8 B[P];
         split8<S>(v, B);
#pragma unroll
         for (int mt = 0; mt < 2; ++mt) acc1[mt][nt] = split_mma<S>(A1[mt], B, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      }
      // ---- twiddle, Y2 to LDS (row c: 64 real parts, then 64 imaginary) --------
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            const uint32_t c  = 4u * q + (uint32_t) i;   // row 15 is padding (zero)
            const uint32_t b  = 16u * nt + r16;
            const float2   w  = tw[(b * c) % (uint32_t) kN];
            const float    yr = acc1[0][nt][i], yi = acc1[1][nt][i];
            region[c * kY2Row + b]      = yr * w.x - yi * w.y;
            region[c * kY2Row + 64 + b] = yr * w.y + yi * w.x;
         }
      }
      // ---- GEMM 2: Z (16 x 128) = A2 (16 x 128) * B2 (128 x 128) -------------
      u16x8 A2[4][P];
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
         float          v[8];
         const float4  *src = (const float4 *) (region + r16 * kY2Row + 32 * ks + 8 * q);
         const float4   lo = src[0], hi = src[1];
         v[0] = lo.x, v[1] = lo.y, v[2] = lo.z, v[3] = lo.w, v[4] = hi.x, v[5] = hi.y, v[6] = hi.z, v[7] = hi.w;
         split8<S>(v, A2[ks]);
      }
      f32x4 acc2[8];
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
         f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
         for (int ks = 0; ks < 4; ++ks) {
            // B2 block: rows from Re Y2 (ks < 2) or Im Y2; columns to Re Z
            // (nt < 4) or Im Z: [[Fr, Fi], [-Fi, Fr]].
            const bool useFi = (ks < 2) != (nt < 4);
            const bool neg   = ks >= 2 && nt < 4;
            const int  mat   = useFi ? 1 : 0;
            u16x8      B[P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
               const u16x8 w = *(const u16x8 *) (fragB2 + p * kFragB2 +
                                                 (((mat * 2 + (ks & 1)) * 4 + (nt & 3)) * 64 + lane) * 8);
               B[p]          = neg ? neg8(w) : w;
            }
            c = split_mma<S>(A2[ks], B, c);
         }
         acc2[nt] = c;
      }
      // ---- Z to LDS in output order k = c + 15d, then coalesced stores --------
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            const uint32_t c = 4u * q + (uint32_t) i;
            const uint32_t d = 16u * nt + r16;
            if (c >= 15u) continue;
            const float zr = unscale * acc2[nt][i], zi = unscale * acc2[nt + 4][i];
            rc[c + 15u * d] = pdiv ? make_float2(zr / (float) kN, -(zi / (float) kN)) : make_float2(zr, syOut * zi);
         }
      }
      float2 *yo = a.y + (size_t) f * kN;
#pragma unroll
      for (int j = 0; j < 15; ++j) yo[lane + 64 * j] = rc[lane + 64 * j];
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

struct DeviceTables {
   uint16_t *frag[3]  = {nullptr, nullptr, nullptr};
   float2   *tw960    = nullptr;
};

std::mutex   g_tabMutex;
DeviceTables g_tables[16];

// A1 (32 x 32, real form of the 15-point DFT) and B2 (Fr, Fi of the 64-point
// DFT) in MFMA operand layout for lane l: A[m = l % 16][k = 8 (l / 16) + i],
// B[k = 8 (l / 16) + i][n = l % 16].
hipError_t build_tables(int S, uint16_t **out)
{
   const int P      = S == SPLIT_BF16X3 ? 3 : 2;
   const size_t cnt = (size_t) P * (kFragA1 + kFragB2);
   uint16_t    *h   = new uint16_t[cnt];
   const double pi  = 3.14159265358979323846;
   uint16_t     parts[3];
   auto A1 = [&](int m, int k) -> double {   // real form of F15 (rows/cols 15, 31 padding)
      const int c = m & 15, a = k & 15;
      if (c >= 15 || a >= 15) return 0.0;
      const double ang = -2.0 * pi * (double) ((a * c) % 15) / 15.0;
      const double fr = cos(ang), fi = sin(ang);
      if (m < 16) return k < 16 ? fr : -fi;
      return k < 16 ? fi : fr;
   };
   for (int mt = 0; mt < 2; ++mt)
      for (int l = 0; l < 64; ++l)
         for (int i = 0; i < 8; ++i) {
            host_split(S, A1(16 * mt + l % 16, 8 * (l / 16) + i), parts);
            for (int p = 0; p < P; ++p) h[(size_t) p * kFragA1 + (mt * 64 + l) * 8 + i] = parts[p];
         }
   uint16_t *hb = h + (size_t) P * kFragA1;
   for (int mat = 0; mat < 2; ++mat)
      for (int kh = 0; kh < 2; ++kh)
         for (int nt = 0; nt < 4; ++nt)
            for (int l = 0; l < 64; ++l)
               for (int i = 0; i < 8; ++i) {
                  const int    b   = 32 * kh + 8 * (l / 16) + i;
                  const int    d   = 16 * nt + l % 16;
                  const double ang = -2.0 * pi * (double) ((b * d) % 64) / 64.0;
                  host_split(S, mat == 0 ? cos(ang) : sin(ang), parts);
                  for (int p = 0; p < P; ++p)
                     hb[(size_t) p * kFragB2 + (((mat * 2 + kh) * 4 + nt) * 64 + l) * 8 + i] = parts[p];
               }
   // layout: all A1 parts, then all B2 parts (copied to LDS as one block)
   uint16_t  *d = nullptr;
   hipError_t e = hipMalloc((void **) &d, cnt * sizeof(uint16_t));
   if (e == hipSuccess) e = hipMemcpy(d, h, cnt * sizeof(uint16_t), hipMemcpyHostToDevice);
   delete[] h;
   if (e != hipSuccess) {
      if (d) (void) hipFree(d);
      return e;
   }
   *out = d;
   return hipSuccess;
}

hipError_t get_tables(int S, const uint16_t **frag, const float2 **tw)
{
   int dev = 0;
   hipError_t e = hipGetDevice(&dev);
   if (e != hipSuccess) return e;
   if (dev < 0 || dev >= 16) return hipErrorInvalidDevice;
   std::lock_guard<std::mutex> lock(g_tabMutex);
   DeviceTables &t = g_tables[dev];
   if (t.tw960 == nullptr) {
      float2      *h  = new float2[kN];
      const double pi = 3.14159265358979323846;
      for (int m = 0; m < kN; ++m) {
         const double ang = -2.0 * pi * (double) m / (double) kN;
         h[m]             = make_float2((float) cos(ang), (float) sin(ang));
      }
      float2 *d = nullptr;
      e         = hipMalloc((void **) &d, kN * sizeof(float2));
      if (e == hipSuccess) e = hipMemcpy(d, h, kN * sizeof(float2), hipMemcpyHostToDevice);
      delete[] h;
      if (e != hipSuccess) {
         if (d) (void) hipFree(d);
         return e;
      }
      t.tw960 = d;
   }
   if (t.frag[S] == nullptr) {
      e = build_tables(S, &t.frag[S]);
      if (e != hipSuccess) return e;
   }
   *frag = t.frag[S];
   *tw   = t.tw960;
   return hipSuccess;
}

template <int S>
hipError_t launch_s(const MfmaArgs &a, hipStream_t stream)
{
   constexpr int P   = SplitTraits<S>::parts;
   const size_t  lds = (size_t) P * (kFragA1 + kFragB2) * sizeof(uint16_t) + kN * sizeof(float2) +
                      (size_t) kWaves * kRegion * sizeof(float);
   static int resident = 0;
   if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute((const void *) c2c960_mfma_kernel<S>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      if (e != hipSuccess) return e;
   }
   if (resident == 0) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
         cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, c2c960_mfma_kernel<S>, kThreadsM, lds) != hipSuccess ||
          per <= 0)
         per = 1;
      resident = cus * per;
   }
   const uint32_t need = (a.numChannels + kWaves - 1) / kWaves;
   const uint32_t grid = need < (uint32_t) resident ? need : (uint32_t) resident;
   hipLaunchKernelGGL(c2c960_mfma_kernel<S>, dim3(grid), dim3(kThreadsM), lds, stream, a);
   return hipGetLastError();
}

}  // namespace

bool mfma_c2c_supported(uint32_t n) { return n == (uint32_t) kN; }

hipError_t launch_c2c_mfma(const float *x, float *y, uint32_t n, uint32_t numChannels, int mode, int variant,
                           hipStream_t stream)
{
   if (n != (uint32_t) kN) return hipErrorInvalidValue;
   if (numChannels == 0) return hipSuccess;
   int S;
   switch (variant) {
   case C2C_VARIANT_BF16X3: S = SPLIT_BF16X3; break;
   case C2C_VARIANT_FP16X2: S = SPLIT_FP16X2; break;
   case C2C_VARIANT_BF16X2: S = SPLIT_BF16X2; break;
   default: return hipErrorInvalidValue;
   }
   MfmaArgs   a{};
   hipError_t e = get_tables(S, &a.frag, &a.tw960);
   if (e != hipSuccess) return e;
   a.x = (const float2 *) x, a.y = (float2 *) y, a.numChannels = numChannels, a.mode = mode;
   if (S == SPLIT_FP16X2) return launch_s<SPLIT_FP16X2>(a, stream);
   if (S == SPLIT_BF16X2) return launch_s<SPLIT_BF16X2>(a, stream);
   return launch_s<SPLIT_BF16X3>(a, stream);
}

}  // namespace fftlib
