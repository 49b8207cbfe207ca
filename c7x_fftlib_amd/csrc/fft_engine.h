// Device-side FFT engine shared by the gfx950 kernels: complex helpers,
// radix-2/3/4/5 butterflies with the reference's in-place slot order, the
// register-pass scheduler over an LDS-resident batch and the output
// permutation (digit reversal) of the reference's in-place DIF.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_kernels.h"

namespace fftlib {
namespace engine {

// Streaming (non-temporal) global accesses: transform inputs are read once
// and outputs written once, so they bypass the caches' retention policy.
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 load_stream(const float2 *p)
{
   const f32x2_t t = __builtin_nontemporal_load((const f32x2_t *) p);
   return make_float2(t.x, t.y);
}
__device__ __forceinline__ void store_stream(float2 *p, float2 v)
{
   const f32x2_t t = {v.x, v.y};
   __builtin_nontemporal_store(t, (f32x2_t *) p);
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 w)
{
   return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}
// -i * z
__device__ __forceinline__ float2 mul_mi(float2 z) { return make_float2(z.y, -z.x); }

// k / r and k % r for the digit radices 2, 3, 5 (constant divisors).
__device__ __forceinline__ uint32_t div_digit(uint32_t k, uint32_t r)
{
   switch (r) {
   case 2: return k >> 1;
   case 3: return k / 3u;
   case 5: return k / 5u;
   default: return k / r;
   }
}

// Position, in the in-place DIF result, of frequency k.
__device__ __forceinline__ uint32_t dif_position(uint32_t k, const StageTable &st)
{
   if (st.pow2) {
      return st.log2n ? (__brev(k) >> (32u - st.log2n)) : 0u;
   }
   uint32_t pos = 0;
   for (uint32_t i = 0; i < st.nDigits; ++i) {
      const uint32_t r  = st.digit[i];
      const uint32_t qd = div_digit(k, r);
      pos               = pos * r + (k - qd * r);
      k                 = qd;
   }
   return pos;
}

// cos/sin of 2*pi/5, 4*pi/5 and 2*pi/3.
constexpr float kC51 = 0.30901699437494745f;
constexpr float kS51 = 0.95105651629515353f;
constexpr float kC52 = -0.80901699437494734f;
constexpr float kS52 = 0.58778525229247325f;
constexpr float kS3  = 0.86602540378443871f;

// ---------------------------------------------------------------------------
// fp32 DIF engine: register passes over an LDS-resident batch
// ---------------------------------------------------------------------------
// LDS slot of point i of the block's data: one pad slot per 16 points breaks
// the power-of-two strides of the late passes across the 64 LDS banks.
__device__ __forceinline__ uint32_t lds_idx(uint32_t i) { return i + (i >> 4); }

// n / d for n < 2^22 via a float reciprocal and one correction step.
struct FastDiv {
   uint32_t d;
   float    rcp;
};
__device__ __forceinline__ FastDiv make_div(uint32_t d) { return FastDiv{d, 1.0f / (float) d}; }
__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f)
{
   uint32_t q = (uint32_t) ((float) n * f.rcp);
   const int r = (int) (n - q * f.d);
   if (r < 0) --q;
   else if (r >= (int) f.d) ++q;
   return q;
}

// Radix-R butterfly on v[o + S*m], m < R: the R-point DFT, outputs m >= 1
// multiplied by their stage twiddles w[(m-1)*wq + p], results stored in the
// reference's in-place slot order (radix-4 outputs 1 and 2 swapped,
// ..._cn.c:368-381). TW = false is the last stage of a transform (wq == 1,
// p == 0), whose twiddles are the table's W^0 = (1, -0) entries: the
// products are skipped (they change at most the sign of a zero).
// Entry i of a table: the base plus a 32-bit byte offset, so that loads from
// global memory take the table's (uniform) address from scalar registers
// and one 32-bit offset register each; with the offset scaled in 64 bits
// every twiddle load held its own 64-bit address (N = 16384: spills).
__device__ __forceinline__ float2 table_at(const float2 *w, uint32_t i)
{
   return *(const float2 *) ((const char *) w + i * (uint32_t) sizeof(float2));
}

template <int R, bool TW = true>
__device__ __forceinline__ void butterfly(float2 *v, const int o, const int S, const float2 *w, uint32_t wq,
                                          uint32_t p)
{
   auto tw = [&](float2 a, uint32_t i) { return TW ? cmul(a, table_at(w, i)) : a; };
   if constexpr (R == 2) {
      const float2 a = v[o], b = v[o + S];
      v[o]           = add2(a, b);
      v[o + S]       = sub2(a, b);
   }
   else if constexpr (R == 3) {
      const float2 x0 = v[o], x1 = v[o + S], x2 = v[o + 2 * S];
      const float2 a = add2(x1, x2), b = sub2(x1, x2);
      const float2 c = make_float2(x0.x - 0.5f * a.x, x0.y - 0.5f * a.y);
      const float2 s = mul_mi(make_float2(kS3 * b.x, kS3 * b.y));
      v[o]         = add2(x0, a);
      v[o + S]     = tw(add2(c, s), p);
      v[o + 2 * S] = tw(sub2(c, s), wq + p);
   }
   else if constexpr (R == 4) {
      const float2 x0 = v[o], x1 = v[o + S], x2 = v[o + 2 * S], x3 = v[o + 3 * S];
      const float2 s02 = add2(x0, x2), d02 = sub2(x0, x2), s13 = add2(x1, x3), d13 = sub2(x1, x3);
      v[o]         = add2(s02, s13);
      v[o + S]     = tw(sub2(s02, s13), wq + p);             // X2 -> slot 1
      v[o + 2 * S] = tw(add2(d02, mul_mi(d13)), p);          // X1 -> slot 2
      v[o + 3 * S] = tw(sub2(d02, mul_mi(d13)), 2 * wq + p); // X3
   }
   else {
      static_assert(R == 5, "radix 2, 3, 4 or 5");
      const float2 x0 = v[o], x1 = v[o + S], x2 = v[o + 2 * S], x3 = v[o + 3 * S], x4 = v[o + 4 * S];
      const float2 a1 = add2(x1, x4), b1 = sub2(x1, x4), a2 = add2(x2, x3), b2 = sub2(x2, x3);
      const float2 c1 = make_float2(x0.x + kC51 * a1.x + kC52 * a2.x, x0.y + kC51 * a1.y + kC52 * a2.y);
      const float2 c2 = make_float2(x0.x + kC52 * a1.x + kC51 * a2.x, x0.y + kC52 * a1.y + kC51 * a2.y);
      const float2 s1 = make_float2(kS51 * b1.x + kS52 * b2.x, kS51 * b1.y + kS52 * b2.y);
      const float2 s2 = make_float2(kS52 * b1.x - kS51 * b2.x, kS52 * b1.y - kS51 * b2.y);
      v[o]         = add2(add2(x0, a1), a2);
      v[o + S]     = tw(add2(c1, mul_mi(s1)), p);
      v[o + 2 * S] = tw(add2(c2, mul_mi(s2)), wq + p);
      v[o + 3 * S] = tw(sub2(c2, mul_mi(s2)), 2 * wq + p);
      v[o + 4 * S] = tw(sub2(c1, mul_mi(s1)), 3 * wq + p);
   }
}

// Input of the first pass when it reads global memory directly: point i of
// transform f, with the inverse transforms' conj / pre-scale applied.
struct GlobalIn {
   const float2 *x;      // first transform of this workgroup
   uint32_t      n;
   int           mode;   // C2CMode
   float         scale;
   __device__ __forceinline__ float2 load(uint32_t f, uint32_t i) const
   {
      float2 v = x[(size_t) f * n + i];
      if (mode != C2C_FORWARD) v.y = -v.y;
      if (mode == C2C_INVERSE_PRESCALE) {
         v.x = scale * v.x;
         v.y = scale * v.y;
      }
      return v;
   }
};

// One register pass: stage s (radix R1, sub-length M) and, when R2 > 1, stage
// s+1 (radix R2). Work item (transform f, sub-DFT d, point p2) holds the
// R1*R2 points d*M + p2 + q2*j, j < R1*R2, q2 = M/(R1*R2): stage s combines
// j = j2 + R2*m over m for each j2, stage s+1 combines j = R2*m + m2 over m2.
// N is the length of each contiguous transform (or sub-block) in LDS.
//
// Each pass is a separate (non-inlined) function reading its stage
// parameters from the kernel-argument StageTable. Two other builds of the
// same passes gave wrong results on gfx950 for every multi-pass length:
// all pass shapes inlined into one kernel-wide loop, and calls that took the
// stage parameters as scalar arguments. Only this form is verified.
template <int R1, int R2, bool FROM_GLOBAL>
__device__ __attribute__((noinline)) void run_pass(float2 *data, uint32_t nf, uint32_t N, const float2 *twl,
                                         const StageTable &st, uint32_t s, const GlobalIn &gin)
{
   constexpr int  R     = R1 * R2;
   const uint32_t M     = st.stageM[s];
   const uint32_t q1    = M / R1;
   const uint32_t q2    = q1 / R2;
   const uint32_t items = N / R;
   const FastDiv  dItems = make_div(items), dQ2 = make_div(q2);
   const float2  *w1    = twl + st.twOff[s];
   const float2  *w2    = twl + st.twOff[R2 > 1 ? s + 1 : s];
   for (uint32_t idx = threadIdx.x; idx < items * nf; idx += blockDim.x) {
      const uint32_t f    = fdiv(idx, dItems);
      const uint32_t u    = idx - f * items;
      const uint32_t d    = fdiv(u, dQ2);
      const uint32_t p2   = u - d * q2;
      const uint32_t loc  = d * M + p2;
      const uint32_t base = f * N + loc;
      float2         v[R];
#pragma unroll
      for (int j = 0; j < R; ++j) v[j] = FROM_GLOBAL ? gin.load(f, loc + q2 * j) : data[lds_idx(base + q2 * j)];
#pragma unroll
      for (int j2 = 0; j2 < R2; ++j2) butterfly<R1>(v, j2, R2, w1, q1, p2 + q2 * j2);
      if constexpr (R2 > 1) {
#pragma unroll
         for (int m = 0; m < R1; ++m) butterfly<R2>(v, R2 * m, 1, w2, q2, p2);
      }
#pragma unroll
      for (int j = 0; j < R; ++j) data[lds_idx(base + q2 * j)] = v[j];
   }
}

template <bool FROM_GLOBAL>
__device__ __forceinline__ void dispatch_pass(float2 *data, uint32_t nf, uint32_t len, const float2 *twl,
                                              const StageTable &st, uint32_t pass, const GlobalIn &gin)
{
   const uint32_t s  = st.passFirst[pass];
   const uint32_t r1 = st.radix[s];
   const uint32_t r2 = st.passLen[pass] == 2 ? st.radix[s + 1] : 1u;
   switch (r1 * 8u + r2) {
   case 5 * 8 + 5: run_pass<5, 5, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 5 * 8 + 3: run_pass<5, 3, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 5 * 8 + 4: run_pass<5, 4, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 5 * 8 + 2: run_pass<5, 2, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 5 * 8 + 1: run_pass<5, 1, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 3 * 8 + 3: run_pass<3, 3, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 3 * 8 + 4: run_pass<3, 4, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 3 * 8 + 2: run_pass<3, 2, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 3 * 8 + 1: run_pass<3, 1, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 4 * 8 + 4: run_pass<4, 4, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 4 * 8 + 2: run_pass<4, 2, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   case 4 * 8 + 1: run_pass<4, 1, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   default: run_pass<2, 1, FROM_GLOBAL>(data, nf, len, twl, st, s, gin); break;
   }
}

// Passes [p0, nPasses) on LDS-resident data (already written by the caller);
// len is the length of each contiguous transform (or sub-block) in LDS.
__device__ inline void dif_passes_lds(float2 *data, uint32_t nf, const float2 *twl, const StageTable &st,
                               uint32_t p0 = 0, uint32_t len = 0)
{
   const GlobalIn none{nullptr, 0, 0, 0.0f};
   if (len == 0) len = st.n;
   for (uint32_t p = p0; p < st.nPasses; ++p) {
      dispatch_pass<false>(data, nf, len, twl, st, p, none);
      __syncthreads();
   }
}

// All passes, the first one reading the transforms straight from global memory.
__device__ inline void dif_passes_global_in(float2 *data, uint32_t nf, const float2 *twl, const StageTable &st,
                                     const GlobalIn &gin)
{
   for (uint32_t p = 0; p < st.nPasses; ++p) {
      if (p == 0)
         dispatch_pass<true>(data, nf, st.n, twl, st, p, gin);
      else
         dispatch_pass<false>(data, nf, st.n, twl, st, p, gin);
      __syncthreads();
   }
}

__device__ __forceinline__ void load_twiddles(float2 *twl, const float2 *tw, uint32_t count)
{
   for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) twl[i] = tw[i];
}

// LDS slots used by F transforms of n points (with padding).
__host__ __device__ __forceinline__ uint32_t lds_points(uint32_t F, uint32_t n) { return F * n + (F * n) / 16u + 1u; }


}  // namespace engine
}  // namespace fftlib
