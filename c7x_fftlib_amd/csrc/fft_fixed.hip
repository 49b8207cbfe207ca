// Length-specialised gfx950 kernels for the batched fp32 complex FFT.
//
// The generic kernels (fft_kernels.hip) read the stage sequence from a
// runtime StageTable, which costs registers (dynamic indexing, a dispatch
// over pass shapes) and occupancy. For the common lengths the same in-place
// decimation-in-frequency schedule is instantiated at compile time: radices,
// sub-lengths, twiddle block offsets, the register passes and the output
// digit reversal are constants, so every index is a constant plus a multiple
// of the work-item number.
//
// Work decomposition: T work items per transform (T = a power of two close to
// N/16, at most 256) and 256/T transforms per workgroup. A register pass of
// radices (r1, r2) is executed by each work item for ceil((N/(r1 r2))/T)
// groups of r1*r2 points; between passes the data stays in LDS (one padding
// slot per 16 points). The first pass reads HBM directly when its point
// groups are at least 32 apart (coalesced rows), otherwise the transforms are
// first copied to LDS with flat coalesced loads. The result is read from LDS
// in digit-reversed order and written in natural order with flat coalesced
// stores, so HBM sees exactly one read and one write per point.
//
// Arithmetic and twiddle use are those of the generic kernels (reference
// src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/..._cn.c:240-491): the
// same butterflies, slot order and twiddle-table entries.
#include <hip/hip_runtime.h>

#include "fft_engine.h"
#include "fft_kernels.h"

namespace fftlib {
namespace {

using namespace engine;

constexpr int kFixedThreads = 256;
constexpr int kMaxFixedSt   = 24;

// Compile-time stage schedule of a 2/3/5-smooth length (same stage order and
// twiddle layout as build_stage_table).
struct FixedPlan {
   int  n;
   int  nStages;
   int  radix[kMaxFixedSt];
   int  M[kMaxFixedSt];
   int  twOff[kMaxFixedSt];
   int  nPasses;
   int  passFirst[kMaxFixedSt];
   int  passLen[kMaxFixedSt];
   int  nDigits;
   int  digit[kMaxFixedSt];
   bool pow2;
   int  log2n;
   int  T;     // work items per transform
   int  FPB;   // transforms per workgroup
};

constexpr FixedPlan make_fixed_plan(int n)
{
   FixedPlan p{};
   p.n    = n;
   int t  = n, M = n, off = 0;
   bool odd = false;
   while (t % 5 == 0) {
      p.radix[p.nStages] = 5, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      off += 4 * (M / 5), M /= 5, t /= 5, odd = true;
      p.digit[p.nDigits++] = 5;
   }
   while (t % 3 == 0) {
      p.radix[p.nStages] = 3, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      off += 2 * (M / 3), M /= 3, t /= 3, odd = true;
      p.digit[p.nDigits++] = 3;
   }
   while (t % 4 == 0) {
      p.radix[p.nStages] = 4, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      off += 3 * (M / 4), M /= 4, t /= 4;
      p.digit[p.nDigits++] = 2;
      p.digit[p.nDigits++] = 2;
   }
   if (t == 2) {
      p.radix[p.nStages] = 2, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      M /= 2, t = 1;
      p.digit[p.nDigits++] = 2;
   }
   p.pow2  = !odd;
   p.log2n = 0;
   while ((1 << p.log2n) < n) ++p.log2n;
   for (int s = 0; s < p.nStages;) {
      const bool pair         = s + 1 < p.nStages && p.radix[s] * p.radix[s + 1] <= 25;
      p.passFirst[p.nPasses]  = s;
      p.passLen[p.nPasses]    = pair ? 2 : 1;
      ++p.nPasses;
      s += pair ? 2 : 1;
   }
   int T = 1;
   while (T < 256 && T * 16 < n) T *= 2;
   p.T   = T;
   p.FPB = kFixedThreads / T;
   return p;
}

template <int N>
struct Fixed {
   static constexpr FixedPlan P = make_fixed_plan(N);
};

// Position of frequency k in the in-place DIF result.
template <int N>
__device__ __forceinline__ uint32_t fixed_position(uint32_t k)
{
   constexpr FixedPlan P = Fixed<N>::P;
   if constexpr (P.pow2) {
      if constexpr (P.log2n == 0) return 0u;
      else return __brev(k) >> (32 - P.log2n);
   }
   else {
      uint32_t pos = 0;
#pragma unroll
      for (int i = 0; i < P.nDigits; ++i) {
         const uint32_t r  = (uint32_t) P.digit[i];
         const uint32_t qd = k / r;
         pos               = pos * r + (k - qd * r);
         k                 = qd;
      }
      return pos;
   }
}

template <int N, int PI, bool DIRECT>
__device__ __forceinline__ void fixed_pass(float2 *data, const float2 *xin, bool active, uint32_t t,
                                           const float2 *__restrict__ tw, int mode, float scale)
{
   constexpr FixedPlan P     = Fixed<N>::P;
   constexpr int       s     = P.passFirst[PI];
   constexpr int       R1    = P.radix[s];
   constexpr int       R2    = P.passLen[PI] == 2 ? P.radix[s + 1] : 1;
   constexpr int       R     = R1 * R2;
   constexpr int       M     = P.M[s];
   constexpr int       q1    = M / R1;
   constexpr int       q2    = q1 / R2;
   constexpr int       items = N / R;
   constexpr int       T     = P.T;
   constexpr int       iters = (items + T - 1) / T;
   const float2       *w1    = tw + P.twOff[s];
   const float2       *w2    = tw + P.twOff[R2 > 1 ? s + 1 : s];
#pragma unroll
   for (int it = 0; it < iters; ++it) {
      const uint32_t u = t + (uint32_t) (it * T);
      if (items % T != 0 && u >= (uint32_t) items) break;
      const uint32_t d   = u / (uint32_t) q2;
      const uint32_t p2  = u - d * (uint32_t) q2;
      const uint32_t loc = d * (uint32_t) M + p2;
      float2         v[R];
      if constexpr (DIRECT) {
#pragma unroll
         for (int j = 0; j < R; ++j) {
            float2 a = active ? xin[loc + (uint32_t) (q2 * j)] : make_float2(0.0f, 0.0f);
            if (mode != C2C_FORWARD) a.y = -a.y;
            if (mode == C2C_INVERSE_PRESCALE) {
               a.x = scale * a.x;
               a.y = scale * a.y;
            }
            v[j] = a;
         }
      }
      else {
#pragma unroll
         for (int j = 0; j < R; ++j) v[j] = data[lds_idx(loc + (uint32_t) (q2 * j))];
      }
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) butterfly<R1>(v, j2, R2, w1, q1, p2 + (uint32_t) (q2 * j2));
      if constexpr (R2 > 1) {
#pragma unroll
         for (int m = 0; m < R1; ++m) butterfly<R2>(v, R2 * m, 1, w2, q2, p2);
      }
#pragma unroll
      for (int j = 0; j < R; ++j) data[lds_idx(loc + (uint32_t) (q2 * j))] = v[j];
   }
}

template <int N, int PI>
__device__ __forceinline__ void fixed_passes_from(float2 *data, uint32_t t, const float2 *__restrict__ tw)
{
   constexpr FixedPlan P = Fixed<N>::P;
   if constexpr (PI < P.nPasses) {
      fixed_pass<N, PI, false>(data, nullptr, false, t, tw, C2C_FORWARD, 1.0f);
      __syncthreads();
      fixed_passes_from<N, PI + 1>(data, t, tw);
   }
}

// LDS slots of one transform: points padded with one slot per 16 (lds_idx).
template <int N>
constexpr uint32_t fixed_stride()
{
   return (uint32_t) (N + (N - 1) / 16 + 1);
}

template <int N>
__global__ void __launch_bounds__(kFixedThreads) c2c_fixed_kernel(const float2 *__restrict__ x,
                                                                  float2 *__restrict__ y,
                                                                  const float2 *__restrict__ tw,
                                                                  uint32_t numChannels, int mode)
{
   constexpr FixedPlan P     = Fixed<N>::P;
   constexpr int       T     = P.T;
   constexpr int       FPB   = P.FPB;
   constexpr int       s0    = P.passFirst[0];
   constexpr int       R0    = P.radix[s0] * (P.passLen[0] == 2 ? P.radix[s0 + 1] : 1);
   constexpr bool      DIRECT = (N / R0) >= 32;
   constexpr uint32_t  STRIDE = fixed_stride<N>();
   extern __shared__ float2 smem[];
   const uint32_t first = blockIdx.x * FPB;
   const uint32_t nf    = min((uint32_t) FPB, numChannels - first);
   const uint32_t fl    = threadIdx.x / T;
   const uint32_t t     = threadIdx.x % T;
   const float    scale = 1.0f / (float) N;
   const float2  *xb    = x + (size_t) first * N;
   float2        *data  = smem + fl * STRIDE;

   if constexpr (DIRECT) {
      fixed_pass<N, 0, true>(data, xb + (size_t) fl * N, fl < nf, t, tw, mode, scale);
      __syncthreads();
      fixed_passes_from<N, 1>(data, t, tw);
   }
   else {
      for (uint32_t i = threadIdx.x; i < nf * N; i += kFixedThreads) {
         const uint32_t f = i / (uint32_t) N;
         float2         a = xb[i];
         if (mode != C2C_FORWARD) a.y = -a.y;
         if (mode == C2C_INVERSE_PRESCALE) {
            a.x = scale * a.x;
            a.y = scale * a.y;
         }
         smem[f * STRIDE + lds_idx(i - f * (uint32_t) N)] = a;
      }
      __syncthreads();
      fixed_passes_from<N, 0>(data, t, tw);
   }
   float2 *yb = y + (size_t) first * N;
   for (uint32_t i = threadIdx.x; i < nf * N; i += kFixedThreads) {
      const uint32_t f = i / (uint32_t) N;
      const uint32_t k = i - f * (uint32_t) N;
      float2         v = smem[f * STRIDE + lds_idx(fixed_position<N>(k))];
      if (mode == C2C_INVERSE_PRESCALE) v.y = -v.y;
      if (mode == C2C_INVERSE_POSTDIV) {
         v.x = v.x / (float) N;
         v.y = -(v.y / (float) N);
      }
      yb[i] = v;
   }
}

template <int N>
hipError_t launch_fixed(const float2 *x, float2 *y, const float2 *tw, uint32_t numChannels, int mode,
                        hipStream_t stream)
{
   constexpr FixedPlan P    = Fixed<N>::P;
   const size_t        lds  = (size_t) P.FPB * fixed_stride<N>() * sizeof(float2);
   if (lds > 65536) {
      const hipError_t e = hipFuncSetAttribute((const void *) c2c_fixed_kernel<N>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      if (e != hipSuccess) return e;
   }
   const uint32_t blocks = (numChannels + P.FPB - 1) / P.FPB;
   hipLaunchKernelGGL(c2c_fixed_kernel<N>, dim3(blocks), dim3(kFixedThreads), lds, stream, x, y, tw, numChannels,
                      mode);
   return hipGetLastError();
}

}  // namespace

bool fixed_c2c_supported(uint32_t n)
{
   switch (n) {
   case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048: case 4096: case 8192:
   case 16384: case 48: case 96: case 192: case 240: case 480: case 960: case 1920: case 3840: case 1200:
   case 1536: case 1000: case 2000: case 4000: case 3200: case 6400:
      return true;
   default: return false;
   }
}

hipError_t launch_c2c_fixed(const float *x, float *y, const float *tw, uint32_t n, uint32_t numChannels, int mode,
                            hipStream_t stream)
{
   const float2 *xi = (const float2 *) x, *wi = (const float2 *) tw;
   float2       *yo = (float2 *) y;
   switch (n) {
#define FFTLIB_FIXED(L)                                                                                             \
   case L: return launch_fixed<L>(xi, yo, wi, numChannels, mode, stream);
      FFTLIB_FIXED(16)
      FFTLIB_FIXED(32)
      FFTLIB_FIXED(64)
      FFTLIB_FIXED(128)
      FFTLIB_FIXED(256)
      FFTLIB_FIXED(512)
      FFTLIB_FIXED(1024)
      FFTLIB_FIXED(2048)
      FFTLIB_FIXED(4096)
      FFTLIB_FIXED(8192)
      FFTLIB_FIXED(16384)
      FFTLIB_FIXED(48)
      FFTLIB_FIXED(96)
      FFTLIB_FIXED(192)
      FFTLIB_FIXED(240)
      FFTLIB_FIXED(480)
      FFTLIB_FIXED(960)
      FFTLIB_FIXED(1920)
      FFTLIB_FIXED(3840)
      FFTLIB_FIXED(1200)
      FFTLIB_FIXED(1536)
      FFTLIB_FIXED(1000)
      FFTLIB_FIXED(2000)
      FFTLIB_FIXED(4000)
      FFTLIB_FIXED(3200)
      FFTLIB_FIXED(6400)
#undef FFTLIB_FIXED
   default: return hipErrorInvalidValue;
   }
}

}  // namespace fftlib
