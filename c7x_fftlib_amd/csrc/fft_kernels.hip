// gfx950 kernels of the FFTLIB fft_c7x hot path: batched mixed-radix fp32
// complex FFT (forward / inverse), fused real-to-complex and complex-to-real
// transforms, and the bit-exact int16 fixed-point FFT.
//
// Design (single pass, LDS resident): one workgroup owns F whole transforms.
// It reads them from HBM once with coalesced loads, stages the caller's
// twiddle table into LDS, runs every decimation-in-frequency stage in LDS
// (one butterfly per work item, a barrier between stages) and writes the
// result once, reading LDS in the digit-reversed order so that HBM writes stay
// coalesced and in natural order. HBM traffic is therefore one read of the
// input and one write of the output per transform -- the algorithmic minimum --
// instead of the reference's one streaming pass per radix stage.
//
// The stage sequence, twiddle-table layout, radix-4 bit-swapped stores and
// output permutation follow the reference natural-C kernels
// (src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/..._cn.c:240-491); the
// int16 arithmetic follows src/fft_c7x/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/
// ..._cn.c:11-338 bit for bit.
#include <hip/hip_runtime.h>

#include "fft_kernels.h"
#include "fft_engine.h"

namespace fftlib {
namespace {

using namespace engine;

constexpr int      kThreads         = 256;
constexpr uint32_t kTargetLdsPoints = 4096;  // complex data points resident per workgroup

// ---------------------------------------------------------------------------
// fp32 kernels
// ---------------------------------------------------------------------------
template <int MODE, bool TWLDS>
__global__ void __launch_bounds__(kThreads) c2c_lds_kernel(const float2 *__restrict__ x, float2 *__restrict__ y,
                                                           const float2 *__restrict__ tw, StageTable st,
                                                           uint32_t numChannels, uint32_t F)
{
   extern __shared__ float2 smem[];
   const uint32_t N     = st.n;
   const uint32_t first = blockIdx.x * F;
   const uint32_t nf    = min(F, numChannels - first);
   float2        *data  = smem;
   const float2  *twl   = TWLDS ? smem + lds_points(F, N) : tw;
   const size_t   base  = (size_t) first * N;

   if (TWLDS) {
      load_twiddles(smem + lds_points(F, N), tw, st.twCount);
      __syncthreads();
   }
   if (st.nPasses > 0) {
      const GlobalIn gin{x + base, N, MODE, 1.0f / (float) N};
      dif_passes_global_in(data, nf, twl, st, gin);
   }
   else {
      for (uint32_t i = threadIdx.x; i < nf * N; i += blockDim.x) {
         const GlobalIn gin{x + base, N, MODE, 1.0f / (float) N};
         data[lds_idx(i)] = gin.load(0, i);
      }
      __syncthreads();
   }
   for (uint32_t i = threadIdx.x; i < nf * N; i += blockDim.x) {
      const uint32_t f = i / N;
      const uint32_t k = i - f * N;
      float2         v = data[lds_idx(f * N + dif_position(k, st))];
      if (MODE == C2C_INVERSE_PRESCALE) v.y = -v.y;
      if (MODE == C2C_INVERSE_POSTDIV) {
         v.x = v.x / (float) N;
         v.y = -(v.y / (float) N);
      }
      y[base + i] = v;
   }
}

// Real-to-complex: N = 2*Nc reals per channel viewed as Nc complex points,
// Nc-point FFT in LDS, then the split with factors A[i] = sf[2i..],
// B[i] = sf[2N-2i-2..] (reference src/fft_c7x/FFTLIB_fft1dBatched_i32f_c32fc_o32fc/
// ..._cn.c:40-121).
template <bool TWLDS>
__global__ void __launch_bounds__(kThreads) r2c_lds_kernel(const float2 *__restrict__ x, float2 *__restrict__ y,
                                                           const float2 *__restrict__ tw,
                                                           const float *__restrict__ sf, StageTable st,
                                                           uint32_t numChannels, uint32_t F, uint32_t outmode)
{
   extern __shared__ float2 smem[];
   const uint32_t Nc    = st.n;
   const uint32_t N     = 2u * Nc;
   const uint32_t nOut  = outmode ? N : Nc + 1u;
   const uint32_t first = blockIdx.x * F;
   const uint32_t nf    = min(F, numChannels - first);
   float2        *data  = smem;
   const float2  *twl   = TWLDS ? smem + lds_points(F, Nc) : tw;

   if (TWLDS) load_twiddles(smem + lds_points(F, Nc), tw, st.twCount);
   for (uint32_t i = threadIdx.x; i < nf * Nc; i += blockDim.x) data[lds_idx(i)] = x[(size_t) first * Nc + i];
   __syncthreads();
   dif_passes_lds(data, nf, twl, st);
   for (uint32_t idx = threadIdx.x; idx < nf * Nc; idx += blockDim.x) {
      const uint32_t f   = idx / Nc;
      const uint32_t i   = idx - f * Nc;
      const uint32_t c   = f * Nc;
      float2        *out = y + (size_t) (first + f) * nOut;
      const float2   zi  = data[lds_idx(c + dif_position(i, st))];
      const float    ar = sf[2u * i], ai = sf[2u * i + 1u];
      const float    br = sf[2u * N - 2u * i - 2u], bi = sf[2u * N - 2u * i - 1u];
      float2         t1;
      if (i == 0u) {
         t1 = make_float2(zi.x * br + zi.y * bi, zi.x * br - zi.y * bi);
      }
      else {
         const float2 zm = data[lds_idx(c + dif_position(Nc - i, st))];
         t1              = make_float2(zm.x * br + zm.y * bi, zm.x * bi - zm.y * br);
      }
      const float2 v = make_float2(zi.x * ar - zi.y * ai + t1.x, zi.x * ai + zi.y * ar + t1.y);
      out[i]         = v;
      if (outmode && i > 0u) out[N - i] = make_float2(v.x, -v.y);
      if (i == 0u) out[Nc] = make_float2(zi.x - zi.y, 0.0f);
   }
}

// Complex-to-real: Z[i] = X[i] A'[i] + V[i] B'[i] with V from the Hermitian
// mirror (inMode 0) or the upper half (inMode 1), then the Nc-point inverse
// (reference src/fft_c7x/FFTLIB_ifft1dBatched_i32fc_c32fc_o32f/..._cn.c:35-87).
template <bool POSTDIV, bool TWLDS>
__global__ void __launch_bounds__(kThreads) c2r_lds_kernel(const float2 *__restrict__ x, float2 *__restrict__ y,
                                                           const float2 *__restrict__ tw,
                                                           const float *__restrict__ sf, StageTable st,
                                                           uint32_t numChannels, uint32_t F, uint32_t inMode)
{
   extern __shared__ float2 smem[];
   const uint32_t Nc       = st.n;
   const uint32_t N        = 2u * Nc;
   const uint32_t inStride = inMode ? N : Nc + 1u;
   const uint32_t first    = blockIdx.x * F;
   const uint32_t nf       = min(F, numChannels - first);
   float2        *data     = smem;
   const float2  *twl      = TWLDS ? smem + lds_points(F, Nc) : tw;
   const float    scale    = 1.0f / (float) Nc;

   if (TWLDS) load_twiddles(smem + lds_points(F, Nc), tw, st.twCount);
   for (uint32_t idx = threadIdx.x; idx < nf * Nc; idx += blockDim.x) {
      const uint32_t f  = idx / Nc;
      const uint32_t i  = idx - f * Nc;
      const float2  *in = x + (size_t) (first + f) * inStride;
      const float2   v0 = in[i];
      float2         v1;
      if (inMode) {
         v1 = in[Nc + i];
      }
      else {
         v1   = in[Nc - i];
         v1.y = -v1.y;
      }
      const float ar = sf[2u * i], ai = sf[2u * i + 1u];
      const float br = sf[N + 2u * i], bi = sf[N + 2u * i + 1u];
      float2      z  = make_float2(v0.x * ar - v0.y * ai + (v1.x * br - v1.y * bi),
                                   v0.x * ai + v0.y * ar + (v1.x * bi + v1.y * br));
      z.y            = -z.y;
      if (!POSTDIV) {
         z.x = scale * z.x;
         z.y = scale * z.y;
      }
      data[lds_idx(idx)] = z;
   }
   __syncthreads();
   dif_passes_lds(data, nf, twl, st);
   for (uint32_t idx = threadIdx.x; idx < nf * Nc; idx += blockDim.x) {
      const uint32_t f = idx / Nc;
      const uint32_t n = idx - f * Nc;
      float2         v = data[lds_idx(f * Nc + dif_position(n, st))];
      if (POSTDIV) {
         v.x = v.x / (float) Nc;
         v.y = -(v.y / (float) Nc);
      }
      else {
         v.y = -v.y;
      }
      y[(size_t) first * Nc + idx] = v;
   }
}

// ---------------------------------------------------------------------------
// int16 fixed point
// ---------------------------------------------------------------------------
// Values are kept as sign-extended int in registers and packed (re in the low
// half-word, im in the high half-word -- the memory order of interleaved
// int16 pairs) in LDS.
__device__ __forceinline__ int lo16(uint32_t v) { return (int) (short) (v & 0xFFFFu); }
__device__ __forceinline__ int hi16(uint32_t v) { return (int) (short) (v >> 16); }
__device__ __forceinline__ uint32_t pack16(int re, int im)
{
   return ((uint32_t) re & 0xFFFFu) | (((uint32_t) im & 0xFFFFu) << 16);
}
// int -> int16 modulo 2^16 (the reference's int16_t assignment under gcc)
__device__ __forceinline__ int wrap16(int v) { return (int) (short) v; }
// real part of the saturating rotation by +j (VCROT90H): -im, -(-32768) -> 32767
__device__ __forceinline__ int rot90_re(int im) { return im == -32768 ? 32767 : -im; }
// Q15 product rounding: floor((acc + 2^14) / 2^15) clamped to int16. acc is a
// sum of two int16*int16 products, |acc| <= 2^31 - 2^15, exact in int32.
__device__ __forceinline__ int q15_round_sat(int acc)
{
   const int r = (acc + 16384) >> 15;
   return min(max(r, -32768), 32767);
}
// Rounding right shift: floor(v / 2^s) plus the last bit shifted out; the
// bit mask is int16(1 << (s-1)) with shift counts taken modulo 32, so s == 0
// leaves v unchanged (reference ..._cn.c:36-45 as executed on x86).
__device__ __forceinline__ int shift_round(int v, uint32_t s)
{
   const int mask = (int) (short) (unsigned short) ((1u << ((s - 1u) & 31u)) & 0xFFFFu);
   const int sh   = v >> (s & 31u);
   return wrap16((v & mask) != 0 ? sh + 1 : sh);
}

// One radix-4 DIF butterfly of the int16 kernels ($I16B/_cn.c butterfly /
// butterflyNoTw): on the packed points at offsets 0, q, 2q, 3q of its block,
// 16-bit wrapping sums, the saturating rotation, Q15 twiddle products
// w[p], w[q+p], w[2q+p] unless noTw, the rounding shift unless the stage is
// the 8-point one (whose shift the final radix-2 stage applies), and the
// outputs in the reference's storage order (o0, o2, o1, o3).
__device__ __forceinline__ void i16_bfly4(uint32_t &v0, uint32_t &v1, uint32_t &v2, uint32_t &v3, const uint32_t *w,
                                          uint32_t p, uint32_t q, bool noTw, bool doShift, uint32_t sh)
{
   const int s1r = wrap16(lo16(v0) + lo16(v2)), s1i = wrap16(hi16(v0) + hi16(v2));
   const int s2r = wrap16(lo16(v1) + lo16(v3)), s2i = wrap16(hi16(v1) + hi16(v3));
   const int d1r = wrap16(lo16(v0) - lo16(v2)), d1i = wrap16(hi16(v0) - hi16(v2));
   const int d2r = wrap16(lo16(v1) - lo16(v3)), d2i = wrap16(hi16(v1) - hi16(v3));
   int       o0r = wrap16(s1r + s2r), o0i = wrap16(s1i + s2i);
   int       o1r = wrap16(d1r - rot90_re(d2i)), o1i = wrap16(d1i - d2r);
   int       o2r = wrap16(s1r - s2r), o2i = wrap16(s1i - s2i);
   int       o3r = wrap16(d1r + rot90_re(d2i)), o3i = wrap16(d1i + d2r);
   if (!noTw) {
      const uint32_t w1 = w[p], w2 = w[q + p], w3 = w[2 * q + p];
      int            t;
      t   = q15_round_sat(o1r * lo16(w1) - o1i * hi16(w1));
      o1i = q15_round_sat(o1r * hi16(w1) + o1i * lo16(w1));
      o1r = t;
      t   = q15_round_sat(o2r * lo16(w2) - o2i * hi16(w2));
      o2i = q15_round_sat(o2r * hi16(w2) + o2i * lo16(w2));
      o2r = t;
      t   = q15_round_sat(o3r * lo16(w3) - o3i * hi16(w3));
      o3i = q15_round_sat(o3r * hi16(w3) + o3i * lo16(w3));
      o3r = t;
   }
   if (doShift) {
      o0r = shift_round(o0r, sh);
      o0i = shift_round(o0i, sh);
      o1r = shift_round(o1r, sh);
      o1i = shift_round(o1i, sh);
      o2r = shift_round(o2r, sh);
      o2i = shift_round(o2i, sh);
      o3r = shift_round(o3r, sh);
      o3i = shift_round(o3i, sh);
   }
   v0 = pack16(o0r, o0i);
   v1 = pack16(o2r, o2i);
   v2 = pack16(o1r, o1i);
   v3 = pack16(o3r, o3i);
}

__global__ void __launch_bounds__(kThreads) i16_lds_kernel(const uint32_t *__restrict__ x, uint32_t *__restrict__ y,
                                                           const uint32_t *__restrict__ tw,
                                                           const uint32_t *__restrict__ shift, StageTable st,
                                                           uint32_t numChannels, uint32_t F)
{
   extern __shared__ uint32_t smem16[];
   const uint32_t N     = st.n;
   const uint32_t first = blockIdx.x * F;
   const uint32_t nf    = min(F, numChannels - first);
   uint32_t      *data  = smem16;
   uint32_t      *twl   = smem16 + F * N;
   const size_t   base  = (size_t) first * N;

   for (uint32_t i = threadIdx.x; i < st.twCount; i += blockDim.x) twl[i] = tw[i];
   for (uint32_t i = threadIdx.x; i < nf * N; i += blockDim.x) data[i] = x[base + i];
   __syncthreads();

   uint32_t M = N, stage = 0, sh = 0, twOff = 0;
   while (M > 2u) {
      sh                  = shift[stage];
      const uint32_t q    = M >> 2;
      const uint32_t lq   = 31u - __clz(q);
      const uint32_t nb   = N >> 2;
      const uint32_t lnb  = 31u - __clz(nb);
      const uint32_t *w   = twl + twOff;
      for (uint32_t idx = threadIdx.x; idx < nb * nf; idx += blockDim.x) {
         const uint32_t f   = idx >> lnb;
         const uint32_t k   = idx & (nb - 1u);
         const uint32_t dft = k >> lq;
         const uint32_t p   = k & (q - 1u);
         uint32_t      *b0  = data + f * N + dft * M + p;
         uint32_t       v0 = b0[0], v1 = b0[q], v2 = b0[2 * q], v3 = b0[3 * q];
         i16_bfly4(v0, v1, v2, v3, w, p, q, (M == 4u) || (M == 8u && p == 0u), M != 8u, sh);
         b0[0]     = v0;
         b0[q]     = v1;
         b0[2 * q] = v2;
         b0[3 * q] = v3;
      }
      __syncthreads();
      twOff += 3u * q;
      M = q;
      ++stage;
   }
   if (M == 2u) {
      const uint32_t nb = N >> 1;
      for (uint32_t idx = threadIdx.x; idx < nb * nf; idx += blockDim.x) {
         const uint32_t f  = idx / nb;
         const uint32_t k  = idx - f * nb;
         uint32_t      *a  = data + f * N + 2u * k;
         const uint32_t u0 = a[0], u1 = a[1];
         a[0] = pack16(shift_round(wrap16(lo16(u0) + lo16(u1)), sh), shift_round(wrap16(hi16(u0) + hi16(u1)), sh));
         a[1] = pack16(shift_round(wrap16(lo16(u0) - lo16(u1)), sh), shift_round(wrap16(hi16(u0) - hi16(u1)), sh));
      }
      __syncthreads();
   }
   for (uint32_t i = threadIdx.x; i < nf * N; i += blockDim.x) {
      const uint32_t f = i / N;
      const uint32_t k = i - f * N;
      y[base + i]      = data[f * N + dif_position(k, st)];
   }
}

// int16 FFT of N = 4^K points (16 <= N <= 4096), specialised per length:
// T = N/16 work items per transform and 256/T transforms per workgroup; the
// rows are staged into LDS with 16-byte coalesced loads (one pad word per 16
// points), the radix-4 stages run two at a time with 16 points per work item
// in registers (the last stage alone when K is odd), and the bit-reversed
// result leaves with 16-byte stores. The butterfly arithmetic is
// i16_bfly4's, as in i16_lds_kernel, so the results are identical.
constexpr int i16_log4(int n) { return n <= 1 ? 0 : 1 + i16_log4(n / 4); }

__device__ __forceinline__ uint32_t i16_slot(uint32_t i) { return i + (i >> 4); }

// twiddle offset of radix-4 stage t (stage t has q = N / 4^(t+1) and uses 3q entries)
template <int N>
constexpr int i16_tw_off(int t)
{
   int off = 0, q = N / 4;
   for (int r = 0; r < t; ++r, q /= 4) off += 3 * q;
   return off;
}

// Stages S and S+1 (block sizes M = N / 4^S and M/4) on the 16 points
// base + q2*j of work item u, base = (u / q2) * M + u % q2, q2 = M / 16.
template <int N, int S>
__device__ __forceinline__ void i16_pass2(uint32_t *data, uint32_t u, const uint32_t *twl, uint32_t sh1, uint32_t sh2)
{
   constexpr uint32_t M = (uint32_t) N >> (2 * S), q1 = M / 4u, q2 = M / 16u;
   const uint32_t     d = u / q2, p2 = u % q2;
   const uint32_t     base = d * M + p2;
   uint32_t           v[16];
#pragma unroll
   for (int j = 0; j < 16; ++j) v[j] = data[i16_slot(base + q2 * (uint32_t) j)];
#pragma unroll
   for (int m = 0; m < 4; ++m)
      i16_bfly4(v[m], v[m + 4], v[m + 8], v[m + 12], twl + i16_tw_off<N>(S), p2 + q2 * (uint32_t) m, q1, false, true,
                sh1);
#pragma unroll
   for (int b = 0; b < 4; ++b)
      i16_bfly4(v[4 * b], v[4 * b + 1], v[4 * b + 2], v[4 * b + 3], twl + i16_tw_off<N>(S + 1), p2, q2, q2 == 1u,
                true, sh2);
#pragma unroll
   for (int j = 0; j < 16; ++j) data[i16_slot(base + q2 * (uint32_t) j)] = v[j];
}

// The last stage alone (block size 4, no twiddles) on points 16u .. 16u+15.
__device__ __forceinline__ void i16_pass_last(uint32_t *data, uint32_t u, uint32_t sh)
{
   uint32_t v[16];
#pragma unroll
   for (int j = 0; j < 16; ++j) v[j] = data[i16_slot(16u * u + (uint32_t) j)];
#pragma unroll
   for (int b = 0; b < 4; ++b) i16_bfly4(v[4 * b], v[4 * b + 1], v[4 * b + 2], v[4 * b + 3], nullptr, 0u, 1u, true, true, sh);
#pragma unroll
   for (int j = 0; j < 16; ++j) data[i16_slot(16u * u + (uint32_t) j)] = v[j];
}

template <int N, int S>
__device__ __forceinline__ void i16_passes_from(uint32_t *data, uint32_t u, bool live, const uint32_t *twl,
                                                const uint32_t (&sh)[i16_log4(N)])
{
   constexpr int K = i16_log4(N);
   if constexpr (S + 1 < K) {
      if (live) i16_pass2<N, S>(data, u, twl, sh[S], sh[S + 1]);
      __syncthreads();
      i16_passes_from<N, S + 2>(data, u, live, twl, sh);
   }
   else if constexpr (S < K) {
      if (live) i16_pass_last(data, u, sh[S]);
      __syncthreads();
   }
}

template <int N>
constexpr size_t i16_pow4_lds()
{
   return ((size_t) (kThreads / (N / 16)) * (N + N / 16) + (size_t) (N - 1)) * sizeof(uint32_t);
}

template <int N>
__global__ void __launch_bounds__(kThreads) i16_pow4_kernel(const uint4 *__restrict__ x, uint4 *__restrict__ y,
                                                            const uint32_t *__restrict__ tw,
                                                            const uint32_t *__restrict__ shift, uint32_t numChannels)
{
   constexpr int      K      = i16_log4(N);
   constexpr int      T      = N / 16;
   constexpr int      F      = kThreads / T;
   constexpr uint32_t STRIDE = (uint32_t) (N + N / 16);
   constexpr int      LOGN   = 2 * K;
   extern __shared__ uint32_t smem16[];
   uint32_t          *twl   = smem16 + F * STRIDE;
   const uint32_t     first = blockIdx.x * (uint32_t) F;
   const uint32_t     nf    = min((uint32_t) F, numChannels - first);
   uint32_t           sh[K];
#pragma unroll
   for (int t = 0; t < K; ++t) sh[t] = shift[t];
   for (uint32_t i = threadIdx.x; i < (uint32_t) (N - 1); i += kThreads) twl[i] = tw[i];
   const uint4 *src = x + (size_t) first * (N / 4);
   for (uint32_t i = threadIdx.x; i < nf * (uint32_t) (N / 4); i += kThreads) {
      const uint4    v = src[i];
      const uint32_t f = (4u * i) / (uint32_t) N;
      const uint32_t k = 4u * i - f * (uint32_t) N;   // multiple of 4: slots k..k+3 are adjacent
      uint32_t      *d = smem16 + f * STRIDE + i16_slot(k);
      d[0] = v.x, d[1] = v.y, d[2] = v.z, d[3] = v.w;
   }
   __syncthreads();
   const uint32_t fl = threadIdx.x / (uint32_t) T;
   i16_passes_from<N, 0>(smem16 + fl * STRIDE, threadIdx.x % (uint32_t) T, fl < nf, twl, sh);
   uint4 *dst = y + (size_t) first * (N / 4);
   for (uint32_t i = threadIdx.x; i < nf * (uint32_t) (N / 4); i += kThreads) {
      const uint32_t  f   = (4u * i) / (uint32_t) N;
      const uint32_t  k   = 4u * i - f * (uint32_t) N;
      const uint32_t *row = smem16 + f * STRIDE;
      uint4           v;
      v.x    = row[i16_slot(__brev(k) >> (32 - LOGN))];
      v.y    = row[i16_slot(__brev(k + 1u) >> (32 - LOGN))];
      v.z    = row[i16_slot(__brev(k + 2u) >> (32 - LOGN))];
      v.w    = row[i16_slot(__brev(k + 3u) >> (32 - LOGN))];
      dst[i] = v;
   }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
uint32_t g_maxLdsBytes = 0;

uint32_t max_lds_bytes()
{
   if (g_maxLdsBytes == 0) {
      int dev = 0, v = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
         g_maxLdsBytes = (uint32_t) v;
      else
         g_maxLdsBytes = 65536u;
   }
   return g_maxLdsBytes;
}

// Launch geometry: transforms per workgroup (F) and LDS bytes. The twiddle
// table is staged in LDS when it fits next to the data, otherwise the stages
// read it from global memory (L1/L2 resident).
// Split pass of the long real-to-complex path: z holds the nc-point complex
// FFT of every channel; one work item per output bin i < nc (the arithmetic
// of r2c_lds_kernel's store loop).
__global__ void __launch_bounds__(kThreads) r2c_split_kernel(const float2 *__restrict__ z, float2 *__restrict__ y,
                                                             const float *__restrict__ sf, uint32_t Nc,
                                                             uint32_t numChannels, uint32_t outmode)
{
   const uint32_t N     = 2u * Nc;
   const uint32_t nOut  = outmode ? N : Nc + 1u;
   const size_t   total = (size_t) Nc * numChannels;
   for (size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x; idx < total;
        idx += (size_t) gridDim.x * blockDim.x) {
      const uint32_t f   = (uint32_t) (idx / Nc);
      const uint32_t i   = (uint32_t) (idx - (size_t) f * Nc);
      const float2  *zc  = z + (size_t) f * Nc;
      float2        *out = y + (size_t) f * nOut;
      const float2   zi  = zc[i];
      const float    ar = sf[2u * i], ai = sf[2u * i + 1u];
      const float    br = sf[2u * N - 2u * i - 2u], bi = sf[2u * N - 2u * i - 1u];
      float2         t1;
      if (i == 0u) {
         t1 = make_float2(zi.x * br + zi.y * bi, zi.x * br - zi.y * bi);
      }
      else {
         const float2 zm = zc[Nc - i];
         t1              = make_float2(zm.x * br + zm.y * bi, zm.x * bi - zm.y * br);
      }
      const float2 v = make_float2(zi.x * ar - zi.y * ai + t1.x, zi.x * ai + zi.y * ar + t1.y);
      out[i]         = v;
      if (outmode && i > 0u) out[N - i] = make_float2(v.x, -v.y);
      if (i == 0u) out[Nc] = make_float2(zi.x - zi.y, 0.0f);
   }
}

// Pre-split pass of the long complex-to-real path: Z[i] = X[i] A'[i] +
// V[i] B'[i] with V = conj(X[nc-i]) (inMode 0) or X[nc+i] (inMode 1), the
// arithmetic of c2r_lds_kernel's load loop before its conjugation and scale.
__global__ void __launch_bounds__(kThreads) c2r_presplit_kernel(const float2 *__restrict__ x, float2 *__restrict__ z,
                                                                const float *__restrict__ sf, uint32_t Nc,
                                                                uint32_t numChannels, uint32_t inMode)
{
   const uint32_t N        = 2u * Nc;
   const uint32_t inStride = inMode ? N : Nc + 1u;
   const size_t   total    = (size_t) Nc * numChannels;
   for (size_t idx = (size_t) blockIdx.x * blockDim.x + threadIdx.x; idx < total;
        idx += (size_t) gridDim.x * blockDim.x) {
      const uint32_t f  = (uint32_t) (idx / Nc);
      const uint32_t i  = (uint32_t) (idx - (size_t) f * Nc);
      const float2  *in = x + (size_t) f * inStride;
      const float2   v0 = in[i];
      float2         v1;
      if (inMode) {
         v1 = in[Nc + i];
      }
      else {
         v1   = in[Nc - i];
         v1.y = -v1.y;
      }
      const float ar = sf[2u * i], ai = sf[2u * i + 1u];
      const float br = sf[N + 2u * i], bi = sf[N + 2u * i + 1u];
      z[idx] = make_float2(v0.x * ar - v0.y * ai + (v1.x * br - v1.y * bi),
                           v0.x * ai + v0.y * ar + (v1.x * bi + v1.y * br));
   }
}

uint32_t elementwise_blocks(size_t items)
{
   const size_t b = (items + kThreads - 1) / kThreads;
   return (uint32_t) (b < 65536 ? (b > 0 ? b : 1) : 65536);
}

struct Geometry {
   uint32_t F;
   size_t   ldsBytes;
   bool     twInLds;
};

bool pick_geometry(uint32_t n, uint32_t elemBytes, uint32_t twCount, uint32_t numChannels, Geometry *g,
                   bool padded = true)
{
   auto slots = [padded](uint32_t f, uint32_t nn) { return padded ? lds_points(f, nn) : f * nn; };
   const size_t limit = max_lds_bytes();
   uint32_t     F     = n >= kTargetLdsPoints ? 1u : kTargetLdsPoints / n;
   if (F > numChannels) F = numChannels;
   if (F == 0) F = 1;
   for (int withTw = 1; withTw >= 0; --withTw) {
      uint32_t f = F;
      while (f > 0 && ((size_t) slots(f, n) + (withTw ? twCount : 0u)) * elemBytes > limit) f >>= 1;
      if (f > 0) {
         g->F        = f;
         g->twInLds  = withTw != 0;
         g->ldsBytes = ((size_t) slots(f, n) + (withTw ? twCount : 0u)) * elemBytes;
         return true;
      }
   }
   return false;
}

template <typename K>
hipError_t prepare_lds(K kernel, size_t ldsBytes)
{
   if (ldsBytes > 65536) {
      return hipFuncSetAttribute((const void *) kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int) ldsBytes);
   }
   return hipSuccess;
}

template <typename K, typename... Args>
hipError_t launch(K kernel, const Geometry &g, uint32_t numChannels, hipStream_t stream, Args... args)
{
   hipError_t err = prepare_lds(kernel, g.ldsBytes);
   if (err != hipSuccess) return err;
   hipLaunchKernelGGL(kernel, dim3((numChannels + g.F - 1) / g.F), dim3(kThreads), g.ldsBytes, stream, args...);
   return hipGetLastError();
}

}  // namespace

uint32_t max_lds_points() { return max_lds_bytes() / 16u; }

bool fits_lds(const StageTable &st)
{
   Geometry g;
   return pick_geometry(st.n, 8u, st.twCount, 1u, &g);
}

hipError_t launch_c2c(const float *x, float *y, const float *tw, const StageTable &st, uint32_t numChannels,
                      int mode, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   Geometry g;
   if (!pick_geometry(st.n, 8u, st.twCount, numChannels, &g)) return hipErrorInvalidValue;
   const float2 *xi = (const float2 *) x, *wi = (const float2 *) tw;
   float2       *yo = (float2 *) y;
   if (mode == C2C_FORWARD)
      return g.twInLds ? launch(c2c_lds_kernel<C2C_FORWARD, true>, g, numChannels, stream, xi, yo, wi, st, numChannels, g.F)
                       : launch(c2c_lds_kernel<C2C_FORWARD, false>, g, numChannels, stream, xi, yo, wi, st, numChannels, g.F);
   if (mode == C2C_INVERSE_PRESCALE)
      return g.twInLds ? launch(c2c_lds_kernel<C2C_INVERSE_PRESCALE, true>, g, numChannels, stream, xi, yo, wi, st, numChannels, g.F)
                       : launch(c2c_lds_kernel<C2C_INVERSE_PRESCALE, false>, g, numChannels, stream, xi, yo, wi, st, numChannels, g.F);
   return g.twInLds ? launch(c2c_lds_kernel<C2C_INVERSE_POSTDIV, true>, g, numChannels, stream, xi, yo, wi, st, numChannels, g.F)
                    : launch(c2c_lds_kernel<C2C_INVERSE_POSTDIV, false>, g, numChannels, stream, xi, yo, wi, st, numChannels, g.F);
}

hipError_t launch_r2c(const float *x, float *y, const float *tw, const float *sf, const StageTable &st,
                      uint32_t numChannels, uint32_t outmode, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   Geometry g;
   if (!pick_geometry(st.n, 8u, st.twCount, numChannels, &g)) return hipErrorInvalidValue;
   const float2 *xi = (const float2 *) x, *wi = (const float2 *) tw;
   float2       *yo = (float2 *) y;
   return g.twInLds ? launch(r2c_lds_kernel<true>, g, numChannels, stream, xi, yo, wi, sf, st, numChannels, g.F, outmode)
                    : launch(r2c_lds_kernel<false>, g, numChannels, stream, xi, yo, wi, sf, st, numChannels, g.F, outmode);
}

hipError_t launch_c2r(const float *x, float *y, const float *tw, const float *sf, const StageTable &st,
                      uint32_t numChannels, uint32_t inMode, int postdiv, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   Geometry g;
   if (!pick_geometry(st.n, 8u, st.twCount, numChannels, &g)) return hipErrorInvalidValue;
   const float2 *xi = (const float2 *) x, *wi = (const float2 *) tw;
   float2       *yo = (float2 *) y;
   if (postdiv)
      return g.twInLds ? launch(c2r_lds_kernel<true, true>, g, numChannels, stream, xi, yo, wi, sf, st, numChannels, g.F, inMode)
                       : launch(c2r_lds_kernel<true, false>, g, numChannels, stream, xi, yo, wi, sf, st, numChannels, g.F, inMode);
   return g.twInLds ? launch(c2r_lds_kernel<false, true>, g, numChannels, stream, xi, yo, wi, sf, st, numChannels, g.F, inMode)
                    : launch(c2r_lds_kernel<false, false>, g, numChannels, stream, xi, yo, wi, sf, st, numChannels, g.F, inMode);
}

hipError_t launch_r2c_split(const float *z, float *y, const float *sf, uint32_t nc, uint32_t numChannels,
                            uint32_t outmode, hipStream_t stream)
{
   if (numChannels == 0 || nc == 0) return hipSuccess;
   hipLaunchKernelGGL(r2c_split_kernel, dim3(elementwise_blocks((size_t) nc * numChannels)), dim3(kThreads), 0, stream,
                      (const float2 *) z, (float2 *) y, sf, nc, numChannels, outmode);
   return hipGetLastError();
}

hipError_t launch_c2r_presplit(const float *x, float *z, const float *sf, uint32_t nc, uint32_t numChannels,
                               uint32_t inMode, hipStream_t stream)
{
   if (numChannels == 0 || nc == 0) return hipSuccess;
   hipLaunchKernelGGL(c2r_presplit_kernel, dim3(elementwise_blocks((size_t) nc * numChannels)), dim3(kThreads), 0,
                      stream, (const float2 *) x, (float2 *) z, sf, nc, numChannels, inMode);
   return hipGetLastError();
}

template <int N>
hipError_t launch_i16_pow4(const int16_t *x, int16_t *y, const int16_t *tw, const uint32_t *shift,
                           uint32_t numChannels, hipStream_t stream)
{
   constexpr uint32_t F = kThreads / (N / 16);
   hipLaunchKernelGGL(i16_pow4_kernel<N>, dim3((numChannels + F - 1) / F), dim3(kThreads), i16_pow4_lds<N>(), stream,
                      (const uint4 *) x, (uint4 *) y, (const uint32_t *) tw, shift, numChannels);
   return hipGetLastError();
}

hipError_t launch_i16(const int16_t *x, int16_t *y, const int16_t *tw, const uint32_t *shift,
                      const StageTable &st, uint32_t numChannels, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   // specialised lengths 4^2 .. 4^6 (rows of 16-byte multiples: 4 B per point)
   if (((uintptr_t) x & 15u) == 0 && ((uintptr_t) y & 15u) == 0) {
      switch (st.n) {
      case 16: return launch_i16_pow4<16>(x, y, tw, shift, numChannels, stream);
      case 64: return launch_i16_pow4<64>(x, y, tw, shift, numChannels, stream);
      case 256: return launch_i16_pow4<256>(x, y, tw, shift, numChannels, stream);
      case 1024: return launch_i16_pow4<1024>(x, y, tw, shift, numChannels, stream);
      case 4096: return launch_i16_pow4<4096>(x, y, tw, shift, numChannels, stream);
      default: break;
      }
   }
   Geometry g;
   if (!pick_geometry(st.n, 4u, st.twCount, numChannels, &g, false) || !g.twInLds) return hipErrorInvalidValue;
   return launch(i16_lds_kernel, g, numChannels, stream, (const uint32_t *) x, (uint32_t *) y,
                 (const uint32_t *) tw, shift, st, numChannels, g.F);
}

}  // namespace fftlib
