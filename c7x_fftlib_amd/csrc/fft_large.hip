// Complex FFTs longer than one workgroup's LDS on the generic (runtime stage
// table) kernels. The specialised path runs these lengths in two passes on
// the class kernels (fft_fixed.hip); this path serves the generic kernels and
// lengths beyond the two-pass range.
//
// The first register passes run over HBM (outer_pass_kernel: each work item
// loads its R = r1*r2 points at stride N/R or M/R, applies one or two DIF
// stages in registers and stores them back to a scratch copy) until the
// remaining sub-DFTs fit in LDS; inner_block_kernel then finishes each
// contiguous sub-block in LDS and writes its outputs in natural order. HBM
// traffic: one read + one write of the data per outer pass, plus the final
// read + write (2 outer + inner = 4 x 8 bytes per point for one outer pass).
//
// The outputs of one sub-block are interleaved with those of the other
// sub-blocks of the same transform (stride N/M2). Blocks of one transform are
// mapped to consecutive dispatch slots of the same XCD, so their partial
// cache lines are completed in that XCD's L2 before write-back.
#include <hip/hip_runtime.h>

#include "fft_engine.h"
#include "fft_kernels.h"

namespace fftlib {
namespace {

using namespace engine;

constexpr int      kThreadsL    = 256;
constexpr uint32_t kInnerPoints = 8192;  // largest sub-block finished in LDS
constexpr uint32_t kXcds        = 8;     // MI355X: 8 XCDs, round-robin dispatch

// Frequency index whose in-place DIF position is P (inverse of dif_position).
__device__ __forceinline__ uint32_t dif_inverse(uint32_t P, const StageTable &st)
{
   if (st.pow2) return st.log2n ? (__brev(P) >> (32u - st.log2n)) : 0u;
   uint32_t k = 0;
   for (int i = (int) st.nDigits - 1; i >= 0; --i) {
      const uint32_t r  = st.digit[i];
      const uint32_t qd = div_digit(P, r);
      k                 = k * r + (P - qd * r);
      P                 = qd;
   }
   return k;
}

// One pass (stages s, s+1) over global memory; src -> dst may alias.
template <int R1, int R2, int MODE, bool FIRST>
__global__ void __launch_bounds__(kThreadsL) outer_pass_kernel(const float2 *src, float2 *dst,
                                                               const float2 *__restrict__ tw, StageTable st,
                                                               uint32_t s, uint32_t numChannels)
{
   constexpr int  R     = R1 * R2;
   const uint32_t N     = st.n;
   const uint32_t M     = st.stageM[s];
   const uint32_t q1    = M / R1;
   const uint32_t q2    = q1 / R2;
   const uint32_t items = N / R;
   const uint32_t u     = blockIdx.x * kThreadsL + threadIdx.x;
   if (u >= items) return;
   const FastDiv  dQ2   = make_div(q2);
   const uint32_t d     = fdiv(u, dQ2);
   const uint32_t p2    = u - d * q2;
   const uint32_t loc   = d * M + p2;
   const float2  *w1    = tw + st.twOff[s];
   const float2  *w2    = tw + st.twOff[R2 > 1 ? s + 1 : s];
   const float    scale = 1.0f / (float) N;
   for (uint32_t f = blockIdx.y; f < numChannels; f += gridDim.y) {
      const size_t base  = (size_t) f * N + loc;
      float2       v[R];
#pragma unroll
      for (int j = 0; j < R; ++j) {
         float2 a = src[base + (size_t) q2 * j];
         if (FIRST && MODE != C2C_FORWARD) a.y = -a.y;
         if (FIRST && MODE == C2C_INVERSE_PRESCALE) {
            a.x = scale * a.x;
            a.y = scale * a.y;
         }
         v[j] = a;
      }
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) butterfly<R1>(v, j2, R2, w1, q1, p2 + q2 * j2);
      if constexpr (R2 > 1) {
#pragma unroll
         for (int m = 0; m < R1; ++m) butterfly<R2>(v, R2 * m, 1, w2, q2, p2);
      }
#pragma unroll
      for (int j = 0; j < R; ++j) dst[base + (size_t) q2 * j] = v[j];
   }
}

// Finishes sub-block b (length M2, passes [p0, nPasses)) of transform f and
// writes its outputs y[f*N + c + (N/M2)*k'] in natural order.
template <int MODE>
__global__ void __launch_bounds__(kThreadsL) inner_block_kernel(const float2 *__restrict__ t,
                                                                float2 *__restrict__ y,
                                                                const float2 *__restrict__ tw, StageTable st,
                                                                uint32_t p0, uint32_t M2, uint32_t numChannels)
{
   extern __shared__ float2 smem[];
   const uint32_t N       = st.n;
   const uint32_t nBlocks = N / M2;
   const uint32_t total   = numChannels * nBlocks;
   const uint32_t per     = (total + kXcds - 1) / kXcds;
   const uint32_t logical = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
   if (logical >= total) return;
   const uint32_t f   = logical / nBlocks;
   const uint32_t b   = logical - f * nBlocks;
   const float2  *src = t + (size_t) f * N + (size_t) b * M2;
   for (uint32_t i = threadIdx.x; i < M2; i += blockDim.x) smem[lds_idx(i)] = src[i];
   __syncthreads();
   dif_passes_lds(smem, 1, tw, st, p0, M2);
   const uint32_t stride = nBlocks;
   const uint32_t c      = dif_inverse(b * M2, st);
   const uint32_t lo     = b * M2;
   for (uint32_t kk = threadIdx.x; kk < M2; kk += blockDim.x) {
      const uint32_t k = c + stride * kk;
      const uint32_t P = dif_position(k, st);
      if (P - lo >= M2) continue;  // cannot happen for a consistent plan
      float2 v = smem[lds_idx(P - lo)];
      if (MODE == C2C_INVERSE_PRESCALE) v.y = -v.y;
      if (MODE == C2C_INVERSE_POSTDIV) {
         v.x = v.x / (float) N;
         v.y = -(v.y / (float) N);
      }
      y[(size_t) f * N + k] = v;
   }
}

template <int MODE, bool FIRST>
hipError_t launch_outer(const float2 *src, float2 *dst, const float2 *tw, const StageTable &st, uint32_t pass,
                        uint32_t numChannels, hipStream_t stream)
{
   const uint32_t s  = st.passFirst[pass];
   const uint32_t r1 = st.radix[s];
   const uint32_t r2 = st.passLen[pass] == 2 ? st.radix[s + 1] : 1u;
   const uint32_t R  = r1 * r2;
   const dim3     grid((st.n / R + kThreadsL - 1) / kThreadsL, numChannels < 65535u ? numChannels : 65535u);
   const dim3     block(kThreadsL);
#define FFTLIB_OUTER(A, B)                                                                                          \
   case A * 8 + B:                                                                                                   \
      hipLaunchKernelGGL((outer_pass_kernel<A, B, MODE, FIRST>), grid, block, 0, stream, src, dst, tw, st, s,       \
                         numChannels);                                                                               \
      break;
   switch (r1 * 8 + r2) {
      FFTLIB_OUTER(5, 5)
      FFTLIB_OUTER(5, 3)
      FFTLIB_OUTER(5, 4)
      FFTLIB_OUTER(5, 2)
      FFTLIB_OUTER(5, 1)
      FFTLIB_OUTER(3, 3)
      FFTLIB_OUTER(3, 4)
      FFTLIB_OUTER(3, 2)
      FFTLIB_OUTER(3, 1)
      FFTLIB_OUTER(4, 4)
      FFTLIB_OUTER(4, 2)
      FFTLIB_OUTER(4, 1)
   default: return hipErrorInvalidValue;
   }
#undef FFTLIB_OUTER
   return hipGetLastError();
}

template <int MODE>
hipError_t run_large(const float2 *x, float2 *y, const float2 *tw, const StageTable &st, uint32_t numChannels,
                     float2 *scratch, uint32_t p0, hipStream_t stream)
{
   hipError_t e;
   for (uint32_t p = 0; p < p0; ++p) {
      e = (p == 0) ? launch_outer<MODE, true>(x, scratch, tw, st, p, numChannels, stream)
                   : launch_outer<MODE, false>(scratch, scratch, tw, st, p, numChannels, stream);
      if (e != hipSuccess) return e;
   }
   const uint32_t M2      = st.stageM[st.passFirst[p0]];
   const size_t   lds     = (size_t) lds_points(1, M2) * sizeof(float2);
   const uint32_t total   = numChannels * (st.n / M2);
   const uint32_t per     = (total + kXcds - 1) / kXcds;
   static bool ldsSet = false;   // largest sub-block's LDS, set once (graph capture safe)
   if (!ldsSet) {
      e = hipFuncSetAttribute((const void *) inner_block_kernel<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int) (lds_points(1, kInnerPoints) * sizeof(float2)));
      if (e != hipSuccess) return e;
      ldsSet = true;
   }
   hipLaunchKernelGGL(inner_block_kernel<MODE>, dim3(per * kXcds), dim3(kThreadsL), lds, stream,
                      (const float2 *) scratch, y, tw, st, p0, M2, numChannels);
   return hipGetLastError();
}

uint32_t large_first_inner_pass(const StageTable &st)
{
   for (uint32_t p = 1; p < st.nPasses; ++p)
      if (st.stageM[st.passFirst[p]] <= kInnerPoints) return p;
   return 0;
}

}  // namespace

hipError_t launch_c2c_large(const float *x, float *y, const float *tw, const StageTable &st, uint32_t numChannels,
                            int mode, float *scratch, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   const uint32_t p0 = large_first_inner_pass(st);
   if (p0 == 0) return hipErrorInvalidValue;
   const float2 *xi = (const float2 *) x, *wi = (const float2 *) tw;
   float2       *yo = (float2 *) y, *t = (float2 *) scratch;
   if (mode == C2C_FORWARD) return run_large<C2C_FORWARD>(xi, yo, wi, st, numChannels, t, p0, stream);
   if (mode == C2C_INVERSE_PRESCALE)
      return run_large<C2C_INVERSE_PRESCALE>(xi, yo, wi, st, numChannels, t, p0, stream);
   return run_large<C2C_INVERSE_POSTDIV>(xi, yo, wi, st, numChannels, t, p0, stream);
}

}  // namespace fftlib
