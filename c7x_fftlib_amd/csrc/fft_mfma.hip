// Split-precision MFMA variant of the batched complex FFT, N = 960.
//
// The 960-point DFT is computed with n = 64a + b (a < 15, b < 64),
// b = 4 b1 + b2 (b1 < 16, b2 < 4) and k = c + 15 d1 + 240 d2
// (c < 15, d1 < 16, d2 < 4):
//
//   Y1[c][b]     = sum_a W15^(ac) x[64a + b]             radix-15 stage (GEMM 1)
//   Y2[c][b]     = W960^(bc) Y1[c][b]                     twiddle
//   V[c][b2][d1] = sum_b1 W16^(b1 d1) Y2[c][4 b1 + b2]    radix-16 stage (GEMM 2)
//   X[c + 15 d1 + 240 d2] = sum_b2 W4^(b2 d2) W64^(b2 d1) V[c][b2][d1]
//                                                        radix-4 stage (fp32 VALU)
//
// The radix-15 and radix-16 DFT stages run in real form on
// v_mfma_f32_16x16x32_{f16,bf16} with FP32 accumulation; their DFT matrices
// (32 x 32 real) stay in registers in MFMA operand layout. The data operands
// are split into sums of 16-bit parts so that the products recover
// FP32-level accuracy:
//   FP16x2: x = h + l (fp16), products h*h + h*l + l*h;
//   BF16x2: same with bf16 parts;
//   BF16x3: x = h + m + l (bf16), products hh + hm + mh + hl + lh + mm.
// FP16 parts are taken of the transform scaled by a power of two (exact) so
// that its largest input is in [0.5, 1): fp16 range then holds every operand.
// The twiddles and the radix-4 stage are fp32.
//
// This is a different (but equivalent) algorithm from the reference's
// in-place mixed-radix DIF (src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/
// ..._cn.c:240-491): it does not read the caller's twiddle table; the DFT
// matrices and twiddles are generated here in double precision. It is
// selected with FFTLIB_GPU_setC2CVariant and its error against an FP64 DFT
// is a property of the split scheme.
//
// Work decomposition: 256-thread workgroups, each looping over transforms;
// each wavefront owns one 9 KB LDS region and runs whole transforms in it
// (transform w, w + 4*gridDim.x, ...), prefetching the next transform's input
// into registers while the current one is transformed. The W960 table is
// copied to LDS once per workgroup.
#include <hip/hip_runtime.h>

#include <math.h>
#include <mutex>
#include <stdint.h>

#include "fft_engine.h"
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
constexpr uint32_t kXRow     = 66;          // complex slots per staged input row (64 + pad)
constexpr uint32_t kVRow     = 36;          // floats per GEMM-2 operand row (re[16] im[16] + pad)
constexpr uint32_t kRegion   = 64 * kVRow;  // floats per wave region
constexpr int      kFragA1   = 2 * 64 * 8;  // halves of the two radix-15 A fragments, per part
constexpr int      kFragB16  = 2 * 64 * 8;  // halves of the two radix-16 B fragments, per part
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
   const uint16_t *frag;      // A1 fragments of every part, then B16 fragments of every part
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
__global__ void __launch_bounds__(kThreadsM) __attribute__((amdgpu_waves_per_eu(3, 3))) c2c960_mfma_kernel(MfmaArgs a)
{
   constexpr int P = SplitTraits<S>::parts;
   extern __shared__ float smf[];
   float2 *tw     = (float2 *) smf;                                       // W960^m, m < 960
   float  *region = smf + 2 * kN + (threadIdx.x / 64) * kRegion;
   float2 *rc     = (float2 *) region;

   const uint32_t lane = threadIdx.x % 64;
   const uint32_t wave = threadIdx.x / 64;
   const uint32_t q    = lane / 16, r16 = lane % 16;

   for (uint32_t i = threadIdx.x; i < (uint32_t) kN; i += kThreadsM) tw[i] = a.tw960[i];
   // DFT-matrix fragments in registers
   u16x8 A1[2][P], B16[2][P];
#pragma unroll
   for (int p = 0; p < P; ++p)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
         A1[t][p]  = *(const u16x8 *) (a.frag + p * kFragA1 + (t * 64 + lane) * 8);
         B16[t][p] = *(const u16x8 *) (a.frag + P * kFragA1 + p * kFragB16 + (t * 64 + lane) * 8);
      }
   __syncthreads();
   // W64^(b2 d1) for d1 = lane % 16, b2 = 1..3 (= W960^(15 b2 d1))
   float2 w64[3];
#pragma unroll
   for (int b2 = 1; b2 < 4; ++b2) w64[b2 - 1] = tw[(15u * (uint32_t) b2 * r16) % (uint32_t) kN];

   const float sxIn  = a.mode == C2C_INVERSE_PRESCALE ? 1.0f / (float) kN : 1.0f;
   const float syIn  = a.mode == C2C_FORWARD ? sxIn : -sxIn;
   const float syOut = a.mode == C2C_INVERSE_PRESCALE ? -1.0f : 1.0f;
   const bool  pdiv  = a.mode == C2C_INVERSE_POSTDIV;

   const uint32_t stride = gridDim.x * kWaves;
   uint32_t       f      = blockIdx.x * kWaves + wave;
   float2         xin[15];
   if (f < a.numChannels) {
#pragma unroll
      for (int j = 0; j < 15; ++j) xin[j] = engine::load_stream(a.x + (size_t) f * kN + lane + 64 * j);
   }
   for (; f < a.numChannels; f += stride) {
      // ---- stage the input (x[64a + b] at row a, column b) --------------------
      float amax = 0.0f;
#pragma unroll
      for (int j = 0; j < 15; ++j) {
         const float2 v = make_float2(sxIn * xin[j].x, syIn * xin[j].y);
         amax           = fmaxf(amax, fmaxf(fabsf(v.x), fabsf(v.y)));
         rc[j * kXRow + lane] = v;
      }
      // Power-of-two scale bringing max|x| into [0.5, 1) for the fp16 parts
      // (1 when the transform is zero or not finite, and for bf16).
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
      // ---- prefetch the next transform ------------------------------------------
      const uint32_t fn = f + stride;
      if (fn < a.numChannels) {
#pragma unroll
         for (int j = 0; j < 15; ++j) xin[j] = engine::load_stream(a.x + (size_t) fn * kN + lane + 64 * j);
      }
      // ---- GEMM 1: C1 (32 x 64) = A1 (32 x 32) * B1 (32 x 64) -------------------
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
      // ---- twiddle W960^(bc); Y2 to LDS as rows (c, b2): re[b1], im[b1] ---------
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            const uint32_t c   = 4u * q + (uint32_t) i;   // row 15 is padding (zero)
            const uint32_t b   = 16u * nt + r16;
            const float2   w   = tw[(b * c) % (uint32_t) kN];
            const float    yr  = acc1[0][nt][i], yi = acc1[1][nt][i];
            float         *row = region + (4u * c + (b & 3u)) * kVRow + (b >> 2);
            row[0]             = yr * w.x - yi * w.y;
            row[16]            = yr * w.y + yi * w.x;
         }
      }
      // ---- GEMM 2 (radix 16), twiddle W64, radix 4, Z to LDS in output order ---
      // Row m = 16 mt + lane % 16 of the operand is (c = 4 mt + lane % 16 / 4,
      // b2 = lane % 4); output rows 16 mt + 4 q + i are (c = 4 mt + q, b2 = i).
      f32x4 vr[4], vi[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
         const float4 *src = (const float4 *) (region + (16u * mt + r16) * kVRow + 8u * q);
         const float4  lo = src[0], hi = src[1];
         float         v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};   // already scaled
         u16x8         A[P];
         split8<S>(v, A);
         vr[mt] = split_mma<S>(A, B16[0], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
         vi[mt] = split_mma<S>(A, B16[1], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
         const uint32_t c = 4u * mt + q;
         if (c >= 15u) continue;
         // t[b2] = W64^(b2 d1) V[c][b2][d1], d1 = lane % 16
         float2 t[4];
         t[0] = make_float2(vr[mt][0], vi[mt][0]);
#pragma unroll
         for (int b2 = 1; b2 < 4; ++b2) {
            const float2 w = w64[b2 - 1];
            t[b2]          = make_float2(vr[mt][b2] * w.x - vi[mt][b2] * w.y, vr[mt][b2] * w.y + vi[mt][b2] * w.x);
         }
         const float2 s02 = make_float2(t[0].x + t[2].x, t[0].y + t[2].y);
         const float2 d02 = make_float2(t[0].x - t[2].x, t[0].y - t[2].y);
         const float2 s13 = make_float2(t[1].x + t[3].x, t[1].y + t[3].y);
         const float2 d13 = make_float2(t[1].x - t[3].x, t[1].y - t[3].y);
         float2       z[4];
         z[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
         z[1] = make_float2(d02.x + d13.y, d02.y - d13.x);   // d02 - i d13
         z[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
         z[3] = make_float2(d02.x - d13.y, d02.y + d13.x);   // d02 + i d13
#pragma unroll
         for (int d2 = 0; d2 < 4; ++d2) {
            const float zr = unscale * z[d2].x, zi = unscale * z[d2].y;
            rc[c + 15u * r16 + 240u * (uint32_t) d2] =
               pdiv ? make_float2(zr / (float) kN, -(zi / (float) kN)) : make_float2(zr, syOut * zi);
         }
      }
      float2 *yo = a.y + (size_t) f * kN;
#pragma unroll
      for (int j = 0; j < 15; ++j) engine::store_stream(yo + lane + 64 * j, rc[lane + 64 * j]);
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

// A1 (32 x 32, real form of the 15-point DFT; rows/columns 15 and 31 are
// padding) and B16 (32 x 32, real form of the 16-point DFT) in MFMA operand
// layout for lane l: A[m = 16 t + l % 16][k = 8 (l / 16) + i],
// B[k = 8 (l / 16) + i][n = 16 t + l % 16]. Layout: A1 of every part, then
// B16 of every part.
hipError_t build_tables(int S, uint16_t **out)
{
   const int    P   = S == SPLIT_BF16X3 ? 3 : 2;
   const size_t cnt = (size_t) P * (kFragA1 + kFragB16);
   uint16_t    *h   = new uint16_t[cnt];
   const double pi  = 3.14159265358979323846;
   uint16_t     parts[3];
   auto A1 = [&](int m, int k) -> double {
      const int c = m & 15, a = k & 15;
      if (c >= 15 || a >= 15) return 0.0;
      const double ang = -2.0 * pi * (double) ((a * c) % 15) / 15.0;
      const double fr = cos(ang), fi = sin(ang);
      if (m < 16) return k < 16 ? fr : -fi;   // Re out = fr Re in - fi Im in
      return k < 16 ? fi : fr;                // Im out = fi Re in + fr Im in
   };
   auto B16 = [&](int k, int n) -> double {
      const int    b1 = k & 15, d1 = n & 15;
      const double ang = -2.0 * pi * (double) ((b1 * d1) % 16) / 16.0;
      const double fr = cos(ang), fi = sin(ang);
      if (n < 16) return k < 16 ? fr : -fi;
      return k < 16 ? fi : fr;
   };
   for (int t = 0; t < 2; ++t)
      for (int l = 0; l < 64; ++l)
         for (int i = 0; i < 8; ++i) {
            host_split(S, A1(16 * t + l % 16, 8 * (l / 16) + i), parts);
            for (int p = 0; p < P; ++p) h[(size_t) p * kFragA1 + (t * 64 + l) * 8 + i] = parts[p];
            host_split(S, B16(8 * (l / 16) + i, 16 * t + l % 16), parts);
            for (int p = 0; p < P; ++p)
               h[(size_t) P * kFragA1 + (size_t) p * kFragB16 + (t * 64 + l) * 8 + i] = parts[p];
         }
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
   const size_t  lds = kN * sizeof(float2) + (size_t) kWaves * kRegion * sizeof(float);
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
   // Grid: fp16x2 and bf16x2 run about 7% faster with about 4 transforms per
   // wavefront (65536 transforms: 4096 workgroups, fp16x2 187.6 -> 173.6 us,
   // bf16x2 186.2 -> 173.0 us on MI355X) than with the resident grid (768
   // workgroups, 21 per wavefront); bf16x3, with twice the matrix products,
   // was fastest with the resident grid.
   const uint32_t need = (a.numChannels + kWaves - 1) / kWaves;
   uint32_t       grid = need < (uint32_t) resident ? need : (uint32_t) resident;
   if (S != SPLIT_BF16X3 && (need + 3u) / 4u > grid) grid = (need + 3u) / 4u;
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
