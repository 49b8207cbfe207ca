// Length-specialised gfx950 kernels (templates): see fft_fixed.hip for the
// design. Included by fft_fixed.hip (dedicated kernels, two-pass launches)
// and by the translation units that instantiate the multi-length class
// kernels (fft_fixed_class0..3.hip), so the instantiations compile in
// parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "fft_engine.h"
#include "fft_kernels.h"

namespace fftlib {
namespace fixedk {

using namespace engine;

inline constexpr int kFixedThreads = 256;    // work items per workgroup up to N = 4096
inline constexpr int kMaxFixedThreads = 1024; // longer lengths: one transform of N / 16 work items
inline constexpr int kMaxFixedSt   = 24;

enum FixedKind : int { KIND_C2C = 0, KIND_R2C = 1, KIND_C2R = 2 };

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
   int  nOdd;              // radix-5 and radix-3 digits (lsd first), then log2 digits of 2
   int  oddDigit[kMaxFixedSt];
   int  rOdd;              // product of the odd digits
   int  log2Pow;           // n = rOdd * 2^log2Pow
   int  twCount;           // complex twiddles the stages read
   int  T;                 // work items per transform
   int  FPB;               // transforms per workgroup
   int  threads;           // work items per workgroup: T * FPB
};

constexpr inline FixedPlan make_fixed_plan(int n)
{
   FixedPlan p{};
   p.n      = n;
   int t    = n, M = n, off = 0;
   p.rOdd   = 1;
   while (t % 5 == 0) {
      p.radix[p.nStages] = 5, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      off += 4 * (M / 5), M /= 5, t /= 5;
      p.oddDigit[p.nOdd++] = 5, p.rOdd *= 5;
   }
   while (t % 3 == 0) {
      p.radix[p.nStages] = 3, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      off += 2 * (M / 3), M /= 3, t /= 3;
      p.oddDigit[p.nOdd++] = 3, p.rOdd *= 3;
   }
   while (t % 4 == 0) {
      p.radix[p.nStages] = 4, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      off += 3 * (M / 4), M /= 4, t /= 4;
      p.log2Pow += 2;
   }
   if (t == 2) {
      p.radix[p.nStages] = 2, p.M[p.nStages] = M, p.twOff[p.nStages] = off, ++p.nStages;
      M /= 2, t = 1;
      p.log2Pow += 1;
   }
   p.twCount = off;
   for (int s = 0; s < p.nStages;) {
      const bool pair        = s + 1 < p.nStages && p.radix[s] * p.radix[s + 1] <= 25;
      p.passFirst[p.nPasses] = s;
      p.passLen[p.nPasses]   = pair ? 2 : 1;
      ++p.nPasses;
      s += pair ? 2 : 1;
   }
   // at most 16 points per work item: 256-item workgroups of FPB transforms
   // up to N = 4096, then one transform per workgroup of N / 16 (512 or
   // 1024) work items, so that the LDS passes of the longest lengths (one
   // workgroup per CU) still run 8 or 16 waves per CU
   int T = 1;
   while (T < kMaxFixedThreads && T * 16 < n) T *= 2;
   p.T       = T;
   p.threads = T > kFixedThreads ? T : kFixedThreads;
   p.FPB     = p.threads / T;
   return p;
}

template <int N>
struct Fixed {
   static constexpr FixedPlan P = make_fixed_plan(N);
};

// LDS slot of point i of a transform, and slots per transform.
__device__ __forceinline__ uint32_t fslot(uint32_t i) { return i + (i >> 5); }
template <int N>
constexpr uint32_t fixed_stride()
{
   return (uint32_t) (N + (N - 1) / 32 + 1);
}

// Position of frequency k in the in-place DIF result: the digit reversal over
// [5.., 3.., 2..] (lsd first) splits into the reversal of k mod rOdd over the
// odd digits, placed above the bit reversal of k / rOdd.
template <int N>
__device__ __forceinline__ uint32_t fixed_position(uint32_t k)
{
   constexpr FixedPlan P = Fixed<N>::P;
   constexpr int       A = P.log2Pow;
   uint32_t            m = k, ro = 0;
   if constexpr (P.rOdd > 1) {
      m          = k / (uint32_t) P.rOdd;
      uint32_t r = k - m * (uint32_t) P.rOdd;
#pragma unroll
      for (int i = 0; i < P.nOdd; ++i) {
         const uint32_t d  = (uint32_t) P.oddDigit[i];
         const uint32_t qd = r / d;
         ro                = ro * d + (r - qd * d);
         r                 = qd;
      }
   }
   if constexpr (A == 0) return ro;
   else return (ro << A) | (__brev(m) >> (32 - A));
}

// One work item of register pass PI: the R1*R2 points d*M + p2 + q2*j of
// item u (d = u / q2, p2 = u % q2), both stages in registers.
template <int N, int PI>
__device__ __forceinline__ void fixed_item(float2 *data, uint32_t u, const float2 *tw)
{
   constexpr FixedPlan P   = Fixed<N>::P;
   constexpr int       s   = P.passFirst[PI];
   constexpr int       R1  = P.radix[s];
   constexpr int       R2  = P.passLen[PI] == 2 ? P.radix[s + 1] : 1;
   constexpr int       R   = R1 * R2;
   constexpr int       M   = P.M[s];
   constexpr int       q1  = M / R1;
   constexpr int       q2  = q1 / R2;
   const float2       *w1  = tw + P.twOff[s];
   const float2       *w2  = tw + P.twOff[R2 > 1 ? s + 1 : s];
   const uint32_t      d   = u / (uint32_t) q2;
   const uint32_t      p2  = u - d * (uint32_t) q2;
   const uint32_t      loc = d * (uint32_t) M + p2;
   float2              v[R];
#pragma unroll
   for (int j = 0; j < R; ++j) v[j] = data[fslot(loc + (uint32_t) (q2 * j))];
#pragma unroll
   for (int j2 = 0; j2 < R2; ++j2) butterfly<R1, (q1 > 1)>(v, j2, R2, w1, q1, p2 + (uint32_t) (q2 * j2));
   if constexpr (R2 > 1) {
#pragma unroll
      for (int m = 0; m < R1; ++m) butterfly<R2, (q2 > 1)>(v, R2 * m, 1, w2, q2, p2);
   }
#pragma unroll
   for (int j = 0; j < R; ++j) data[fslot(loc + (uint32_t) (q2 * j))] = v[j];
}

// Register pass PI over one transform in LDS: work item t takes items
// t, t + T, ... (unrolled for up to 4 items per work item).
template <int N, int PI>
__device__ __forceinline__ void fixed_pass(float2 *data, uint32_t t, const float2 *tw)
{
   constexpr FixedPlan P     = Fixed<N>::P;
   constexpr int       s     = P.passFirst[PI];
   constexpr int       R     = P.radix[s] * (P.passLen[PI] == 2 ? P.radix[s + 1] : 1);
   constexpr int       items = N / R;
   constexpr int       T     = P.T;
   constexpr int       iters = (items + T - 1) / T;
   if constexpr (iters <= 4) {
#pragma unroll
      for (int it = 0; it < iters; ++it) {
         const uint32_t u = t + (uint32_t) (it * T);
         if (items % T != 0 && u >= (uint32_t) items) break;
         fixed_item<N, PI>(data, u, tw);
      }
   }
   else {
#pragma unroll 1
      for (uint32_t u = t; u < (uint32_t) items; u += T) fixed_item<N, PI>(data, u, tw);
   }
}

template <int N, int PI>
__device__ __forceinline__ void fixed_passes_from(float2 *data, uint32_t t, const float2 *tw)
{
   constexpr FixedPlan P = Fixed<N>::P;
   if constexpr (PI < P.nPasses) {
      fixed_pass<N, PI>(data, t, tw);
      __syncthreads();
      fixed_passes_from<N, PI + 1>(data, t, tw);
   }
}

// Scaling of the inverse transforms' input as two factors: forward (1, 1);
// batched inverse (1/N, -1/N) = conj then scale; single-channel inverse
// (1, -1) = conj. Multiplying by +-1 or by the scale is exact or identical
// to the separate conj and scale steps.
struct InScale {
   float sx, sy;
};
__device__ __forceinline__ InScale in_scale(int mode, float scale)
{
   if (mode == C2C_INVERSE_PRESCALE) return InScale{scale, -scale};
   if (mode == C2C_INVERSE_POSTDIV) return InScale{1.0f, -1.0f};
   return InScale{1.0f, 1.0f};
}

// Everything a kernel needs besides the LDS data: transform kind and mode,
// the caller's buffers and the per-channel strides (in complex points).
struct FixedArgs {
   const float2 *x;
   float2       *y;
   const float2 *tw;
   const float  *sf;          // split factors (real transforms)
   const float2 *sf2;         // the same table when 8-byte aligned (one load per factor), else null
   uint32_t      numChannels;
   uint32_t      numGroups;
   int           mode;        // C2CMode
   uint32_t      aux;         // R2C outmode / C2R inMode
   uint32_t      inStride;    // complex points per input channel
   uint32_t      outStride;   // complex points per output channel
   const uint64_t *inOff;     // ragged batches: complex offset of each channel in x (else null)
   const uint64_t *outOff;    // ... and in y (of each full transform in the two-pass layouts)
   // Two-pass layouts of a transform longer than one workgroup, fullN =
   // N1 * M2, input n = p + M2*i, output k = k1 + N1*k2:
   //   LAYOUT_COLUMNS (this kernel's N = N1, span = M2): channel f*M2 + p is
   //     column p of transform f, points x[in(f) + p + M2*i]; its outputs
   //     Z[k1] = W_fullN^(p*k1) * Y[k1] go to y[f*fullN + k1*M2 + p];
   //   LAYOUT_ROWS (N = M2, span = N1): channel f*N1 + k1 is row k1 of the
   //     column results (contiguous, x[(f*N1 + k1)*M2 + p]); its outputs go
   //     to y[out(f) + k1 + N1*k2].
   // A workgroup's FPB channels are adjacent columns (rows), so every access
   // to HBM is a run of FPB consecutive points.
   uint32_t      layout;
   uint32_t      span;
   uint32_t      fullN;
   const float2 *colTw;       // LAYOUT_COLUMNS: W_fullN^e for e < 2^colBits ...
   const float2 *colTwHi;     // ... and W_fullN^(h 2^colBits): W^e = colTwHi[e >> colBits] colTw[e & mask]
   uint32_t      colBits;
   uint32_t      scrStride;   // two-pass: points per row of the intermediate copy (span of the columns rounded up to 16)
   // Two-pass with rowAlign = A (16 or 8 points, else 0): the rows of
   // transform f are rotated in the intermediate copy (row r holds k1 =
   // (r + rot_f) mod N1, rot_f from the address of its output in yOut, see
   // row_rotation) so that the row groups' output runs start on 8 A-byte
   // boundaries.
   uint32_t      rowAlign;
   const float2 *yOut;        // the final output (also in the column pass, whose y is the intermediate copy)
};

enum : uint32_t { LAYOUT_PLAIN = 0u, LAYOUT_COLUMNS = 1u, LAYOUT_ROWS = 2u };

// Row rotation of a two-pass transform of N1 = n1 rows whose output starts
// at `out`, for A = rowAlign (N1 and the row groups' FPB multiples of A): 0
// when out is on an 8 A-byte boundary, else N1 - A + s with s the points
// from out to the next one. Each block of A rows A c .. A c + A - 1 (within
// one row group) then holds k1 from A c - A + s on, so that its outputs
// k1 + N1 k2 fill whole 8 A-byte pieces (block 0's rows k1 >= N1 - A + s at
// k2 and k1 < s at k2 + 1 share pieces too). A = 16 fills 128-byte lines: a
// line written in parts by several workgroups, on different dies, is
// written back in parts, and the partial writes cost HBM read-modify-write
// cycles (2^17 x 65536-point transforms: row pass 400 us with the output on
// a line, 520 us 64 bytes off it, 620 us 8 bytes off it).
__device__ __forceinline__ uint32_t row_rotation(const FixedArgs &a, const float2 *out, uint32_t n1)
{
   if (!a.rowAlign) return 0u;
   const uint32_t s = (uint32_t) ((0u - (uintptr_t) out) >> 3) & (a.rowAlign - 1u);
   return s ? n1 - a.rowAlign + s : 0u;
}

// First complex element of channel ch in the input and in the output.
__device__ __forceinline__ size_t in_base(const FixedArgs &a, uint32_t ch)
{
   return a.inOff ? (size_t) a.inOff[ch] : (size_t) ch * a.inStride;
}
__device__ __forceinline__ size_t out_base(const FixedArgs &a, uint32_t ch)
{
   return a.outOff ? (size_t) a.outOff[ch] : (size_t) ch * a.outStride;
}

// Complex entry j (floats 2j, 2j+1) of the split-factor table.
__device__ __forceinline__ float2 split_factor(const FixedArgs &a, uint32_t j)
{
   if (a.sf2) return a.sf2[j];
   return make_float2(a.sf[2u * j], a.sf[2u * j + 1u]);
}

// Point i (< N) of the complex sequence the FFT of channel `in` transforms:
// the input itself (C2C, R2C), or the C2R pre-split
// Z[i] = X[i] A'[i] + V[i] B'[i], V = conj(X[N-i]) (inMode 0) or X[N+i]
// (inMode 1), A' = sf[2i..], B' = sf[2N+2i..] (2N reals per channel).
template <int N, int KIND>
__device__ __forceinline__ float2 fixed_input(const float2 *in, uint32_t i, const FixedArgs &a)
{
   if constexpr (KIND != KIND_C2R) {
      return load_stream(in + i);
   }
   else {
      const float2 v0 = in[i];
      float2       v1;
      if (a.aux) {
         v1 = in[N + i];
      }
      else {
         v1   = in[N - i];
         v1.y = -v1.y;
      }
      const float2 A = split_factor(a, i);
      const float2 B = split_factor(a, (uint32_t) N + i);
      return make_float2(v0.x * A.x - v0.y * A.y + (v1.x * B.x - v1.y * B.y),
                         v0.x * A.y + v0.y * A.x + (v1.x * B.y + v1.y * B.x));
   }
}

// First pass split in two halves so that the HBM reads of the next group of
// transforms can be issued before the current group's LDS passes: the loads
// of pass 0 (work item u reads points u + q2*j of its transform; pass 0 has
// d = 0) and the butterflies + LDS stores of pass 0.
template <int N>
struct Pass0 {
   static constexpr FixedPlan P     = Fixed<N>::P;
   static constexpr int       s     = P.passFirst[0];
   static constexpr int       R1    = P.radix[s];
   static constexpr int       R2    = P.passLen[0] == 2 ? P.radix[s + 1] : 1;
   static constexpr int       R     = R1 * R2;
   static constexpr int       q2    = N / R;
   static constexpr int       items = N / R;
   static constexpr int       iters = (items + P.T - 1) / P.T;
};

// Loads are unconditional so that they issue back to back: work items of a
// transform beyond the batch re-read channel 0 of the group (their results
// are never stored); items beyond the pass's item count re-read the last item.
template <int N, int KIND>
__device__ __forceinline__ void pass0_load_one(float2 (&v)[Pass0<N>::R], uint32_t first, uint32_t fl, uint32_t nf,
                                               uint32_t u, const FixedArgs &a)
{
   using Q           = Pass0<N>;
   const float2 *in  = a.x + in_base(a, first + (fl < nf ? fl : 0u));
   if (Q::items % Q::P.T != 0) u = min(u, (uint32_t) Q::items - 1u);
#pragma unroll
   for (int j = 0; j < Q::R; ++j) v[j] = fixed_input<N, KIND>(in, u + (uint32_t) (Q::q2 * j), a);
}

template <int N>
__device__ __forceinline__ void pass0_compute_one(float2 (&w)[Pass0<N>::R], float2 *data, uint32_t u,
                                                  const float2 *tw, InScale sc)
{
   using Q          = Pass0<N>;
   constexpr int q1 = N / Q::R1;
   if (Q::items % Q::P.T != 0 && u >= (uint32_t) Q::items) return;
   const float2 *w1 = tw + Q::P.twOff[Q::s];
   const float2 *w2 = tw + Q::P.twOff[Q::R2 > 1 ? Q::s + 1 : Q::s];
#pragma unroll
   for (int j = 0; j < Q::R; ++j) {
      w[j].x = sc.sx * w[j].x;
      w[j].y = sc.sy * w[j].y;
   }
#pragma unroll
   for (int j2 = 0; j2 < Q::R2; ++j2) butterfly<Q::R1, (q1 > 1)>(w, j2, Q::R2, w1, q1, u + (uint32_t) (Q::q2 * j2));
   if constexpr (Q::R2 > 1) {
#pragma unroll
      for (int m = 0; m < Q::R1; ++m) butterfly<Q::R2, (Q::q2 > 1)>(w, Q::R2 * m, 1, w2, Q::q2, u);
   }
#pragma unroll
   for (int j = 0; j < Q::R; ++j) data[fslot(u + (uint32_t) (Q::q2 * j))] = w[j];
}

template <int N, int KIND>
__device__ __forceinline__ void pass0_load(float2 (&v)[Pass0<N>::iters][Pass0<N>::R], uint32_t first, uint32_t fl,
                                           uint32_t nf, uint32_t t, const FixedArgs &a)
{
#pragma unroll
   for (int it = 0; it < Pass0<N>::iters; ++it)
      pass0_load_one<N, KIND>(v[it], first, fl, nf, t + (uint32_t) (it * Pass0<N>::P.T), a);
}

template <int N>
__device__ __forceinline__ void pass0_compute(float2 (&v)[Pass0<N>::iters][Pass0<N>::R], float2 *data, uint32_t t,
                                              const float2 *tw, InScale sc)
{
#pragma unroll
   for (int it = 0; it < Pass0<N>::iters; ++it)
      pass0_compute_one<N>(v[it], data, t + (uint32_t) (it * Pass0<N>::P.T), tw, sc);
}

// Output k of the complex FFT in LDS after the single-channel inverse's
// division (the other modes are a sign factor applied by the caller).
__device__ __forceinline__ float2 out_postdiv(float2 v, float n) { return make_float2(v.x / n, -(v.y / n)); }

// LDS slots the store stage of work item t reads: outputs k = t + T*it of
// its transform (and, for R2C, the mirrored N - k). They are the same for
// every group of a persistent workgroup, so they are computed once when they
// fit in a few registers (at most 16 outputs per work item). R2C computes
// them per output instead: its store runs in chunks of 8 with the split
// factors in flight, and indexing the cached arrays by the chunk would put
// them in scratch memory, while unrolling every chunk needs more than 200
// VGPRs.
template <int N, int KIND, bool CACHE = (KIND != KIND_R2C)>
struct StoreMap {
   static constexpr FixedPlan P      = Fixed<N>::P;
   static constexpr int       iters  = (N + P.T - 1) / P.T;
   static constexpr bool      cached = CACHE && iters <= 16;
   static constexpr int       slots  = cached ? iters : 1;
   uint32_t                   off[slots];
   uint32_t                   mir[KIND == KIND_R2C ? slots : 1];

   __device__ __forceinline__ void init(uint32_t t)
   {
      if constexpr (cached) {
#pragma unroll
         for (int it = 0; it < iters; ++it) {
            const uint32_t k = min(t + (uint32_t) (it * P.T), (uint32_t) N - 1u);
            off[it]          = fslot(fixed_position<N>(k));
            if constexpr (KIND == KIND_R2C) mir[it] = fslot(fixed_position<N>(k == 0u ? 0u : (uint32_t) N - k));
         }
      }
   }
   __device__ __forceinline__ uint32_t at(int it, uint32_t t) const
   {
      if constexpr (cached) return off[it];
      else return fslot(fixed_position<N>(min(t + (uint32_t) (it * P.T), (uint32_t) N - 1u)));
   }
   __device__ __forceinline__ uint32_t mirror_at(int it, uint32_t t) const
   {
      if constexpr (cached && KIND == KIND_R2C) return mir[it];
      else {
         const uint32_t k = min(t + (uint32_t) (it * P.T), (uint32_t) N - 1u);
         return fslot(fixed_position<N>(k == 0u ? 0u : (uint32_t) N - k));
      }
   }
};

// C2R inMode 0 pre-split of a whole group when a workgroup holds FPB > 1
// transforms: work item w takes the mirror pairs (i, N - i), i = w + 256c <
// N/2, of every transform of the group, loading the four split factors of a
// pair once (pair 0 takes those of N/2 for the self-mirrored Z[N/2], its
// partner being the Nyquist bin). Same formula per point as the per-transform
// pre-split in fixed_kernel. Used where every work item has pairs (N/2 >= 256)
// and a group holds 2-4 transforms: C2R 2048 (503 MB per call) 244.0 ->
// 220.4 us, 960 248.6 -> 217.3 us, 1024 239.9 -> 232.4 us; at 512 (8 per
// group) it was slower (226.7 -> 249.1 us).
template <int N>
constexpr bool c2r_group_presplit()
{
   return Fixed<N>::P.FPB > 1 && Fixed<N>::P.FPB <= 4 && N / 2 >= kFixedThreads;
}

// Writes transform `ch` from its LDS slots `src`: work item t handles outputs
// t, t + T, ... in chunks of up to 16 (8 for R2C; all LDS reads of a chunk
// are issued before its stores). C2C and C2R write the N complex values in natural
// order; R2C applies the split Y[i] = Z[i] A[i] + (Z[N-i] mirrored) B[i] for
// i < N, the Nyquist bin Y[N] = Z0re - Z0im, and for outmode 1 the conjugate
// upper half Y[2N-i] = conj(Y[i]).
template <int N, int KIND, bool C>
__device__ __forceinline__ void fixed_store(const float2 *src, uint32_t ch, uint32_t t, const FixedArgs &a,
                                            const StoreMap<N, KIND, C> &map)
{
   constexpr FixedPlan P     = Fixed<N>::P;
   constexpr int       T     = P.T;
   constexpr int       iters = StoreMap<N, KIND, C>::iters;
   constexpr int       MAXC  = KIND == KIND_R2C ? 8 : 16;
   constexpr int       CH    = iters < MAXC ? iters : MAXC;
   constexpr bool      EXACT = N % T == 0 && iters % CH == 0;
   float2             *out   = a.y + out_base(a, ch);
   const float         sy    = a.mode == C2C_INVERSE_PRESCALE ? -1.0f : 1.0f;
   const bool          pdiv  = a.mode == C2C_INVERSE_POSTDIV;
   // one chunk of CH outputs starting at output index t + T*c0; with a cached
   // map the chunk loop is unrolled so that the map's arrays are indexed by
   // constants (a runtime index would put them in scratch memory)
   auto chunk = [&](int c0) __attribute__((always_inline)) {
      float2 v[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) v[c] = src[map.at(c0 + c, t)];
      if constexpr (KIND == KIND_R2C) {
         float2 vm[CH];
#pragma unroll
         for (int c = 0; c < CH; ++c) vm[c] = src[map.mirror_at(c0 + c, t)];
#pragma unroll
         for (int c = 0; c < CH; ++c) {
            const uint32_t i = t + (uint32_t) ((c0 + c) * T);
            if (!EXACT && (c0 + c >= iters || i >= (uint32_t) N)) break;
            const float2 zi = v[c], zm = vm[c];
            const float2 A  = split_factor(a, i);
            const float2 B  = split_factor(a, 2u * N - 1u - i);
            float2       t1;
            if (i == 0u) t1 = make_float2(zi.x * B.x + zi.y * B.y, zi.x * B.x - zi.y * B.y);
            else t1 = make_float2(zm.x * B.x + zm.y * B.y, zm.x * B.y - zm.y * B.x);
            const float2 y = make_float2(zi.x * A.x - zi.y * A.y + t1.x, zi.x * A.y + zi.y * A.x + t1.y);
            out[i]         = y;
            if (a.aux && i > 0u) out[2u * N - i] = make_float2(y.x, -y.y);
            if (i == 0u) out[N] = make_float2(zi.x - zi.y, 0.0f);
         }
      }
      else if (pdiv) {
#pragma unroll
         for (int c = 0; c < CH; ++c) {
            const uint32_t k = t + (uint32_t) ((c0 + c) * T);
            if (!EXACT && (c0 + c >= iters || k >= (uint32_t) N)) break;
            store_stream(out + k, out_postdiv(v[c], (float) N));
         }
      }
      else {
#pragma unroll
         for (int c = 0; c < CH; ++c) {
            const uint32_t k = t + (uint32_t) ((c0 + c) * T);
            if (!EXACT && (c0 + c >= iters || k >= (uint32_t) N)) break;
            store_stream(out + k, make_float2(v[c].x, sy * v[c].y));
         }
      }
   };
   if constexpr (StoreMap<N, KIND, C>::cached) {
#pragma unroll
      for (int c0 = 0; c0 < iters; c0 += CH) chunk(c0);
   }
   else {
#pragma unroll 1
      for (int c0 = 0; c0 < iters; c0 += CH) chunk(c0);
   }
}

constexpr inline int largest_divisor_upto(int n, int cap)
{
   int c = n < cap ? n : cap;
   while (n % c != 0) --c;
   return c;
}

// R2C store of a whole group when a workgroup holds FPB > 1 transforms: work
// item w handles outputs i = w, w + 256, ... of every transform of the group,
// so the split factors A[i], B[i] and the digit-reversed slots of i and N - i
// are loaded and computed once per output index instead of once per
// transform (R2C 2048, 503 MB per call: 233.7 -> 205.3 us; 960: 231.6 ->
// 202.2 us). Same formula per output as fixed_store; the compiler may contract its
// products into FMAs differently, so the two can differ in the last bit. Used from
// N = 256 on, where every work item has outputs; shorter lengths would leave
// most work items idle while the others loop over many transforms.
template <int N>
constexpr bool r2c_group_store()
{
   return Fixed<N>::P.FPB > 1 && N >= kFixedThreads;
}

template <int N>
__device__ __forceinline__ void r2c_store_group(const float2 *smem, uint32_t first, uint32_t nf, const FixedArgs &a)
{
   static_assert(Fixed<N>::P.threads == kFixedThreads, "groups of several transforms: 256 work items");
   constexpr int      FPB    = Fixed<N>::P.FPB;
   constexpr uint32_t STRIDE = fixed_stride<N>();
   constexpr int      iters  = (N + kFixedThreads - 1) / kFixedThreads;
   constexpr int      CH     = largest_divisor_upto(iters, 4);
#pragma unroll 1
   for (int c0 = 0; c0 < iters; c0 += CH) {
      float2   A[CH], B[CH];
      uint32_t s[CH], sm[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t i = min(threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads), (uint32_t) N - 1u);
         A[c]             = split_factor(a, i);
         B[c]             = split_factor(a, 2u * N - 1u - i);
         s[c]             = fslot(fixed_position<N>(i));
         sm[c]            = fslot(fixed_position<N>(i == 0u ? 0u : (uint32_t) N - i));
      }
#pragma unroll
      for (int f = 0; f < FPB; ++f) {
         if ((uint32_t) f >= nf) break;
         const float2 *src = smem + f * STRIDE;
         float2       *out = a.y + out_base(a, first + (uint32_t) f);
         float2        zi[CH], zm[CH];
#pragma unroll
         for (int c = 0; c < CH; ++c) zi[c] = src[s[c]], zm[c] = src[sm[c]];
#pragma unroll
         for (int c = 0; c < CH; ++c) {
            const uint32_t i = threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads);
            if (N % kFixedThreads != 0 && i >= (uint32_t) N) break;
            float2 t1;
            if (i == 0u) t1 = make_float2(zi[c].x * B[c].x + zi[c].y * B[c].y, zi[c].x * B[c].x - zi[c].y * B[c].y);
            else t1 = make_float2(zm[c].x * B[c].x + zm[c].y * B[c].y, zm[c].x * B[c].y - zm[c].y * B[c].x);
            const float2 y = make_float2(zi[c].x * A[c].x - zi[c].y * A[c].y + t1.x,
                                         zi[c].x * A[c].y + zi[c].y * A[c].x + t1.y);
            out[i]         = y;
            if (a.aux && i > 0u) out[2u * N - i] = make_float2(y.x, -y.y);
            if (i == 0u) out[N] = make_float2(zi[c].x - zi[c].y, 0.0f);
         }
      }
   }
}

// The split factors A[i], B[2N-1-i] of the outputs i = w + 256c a work item
// stores in r2c_store_group, loaded once per workgroup by the prefetching
// kernels: loaded per group they were the youngest vector-memory loads of
// the store phase, and waiting for them (the counter completes in order)
// also waited for the next group's prefetched input. The LDS slots are
// recomputed per group (held too, N = 2048 took 177 VGPRs, one wave per
// SIMD below what its LDS allows).
template <int N>
struct R2CSplitCache {
   static constexpr int iters = (N + kFixedThreads - 1) / kFixedThreads;
   float2               A[iters], B[iters];

   __device__ __forceinline__ void init(const FixedArgs &a)
   {
#pragma unroll
      for (int c = 0; c < iters; ++c) {
         const uint32_t i = min(threadIdx.x + (uint32_t) (c * kFixedThreads), (uint32_t) N - 1u);
         A[c]             = split_factor(a, i);
         B[c]             = split_factor(a, 2u * N - 1u - i);
      }
   }
};

// r2c_store_group with the factors and slots of an R2CSplitCache; strided
// outputs (the dedicated kernels have no offset tables).
template <int N>
__device__ __forceinline__ void r2c_store_group_cached(const float2 *smem, uint32_t first, uint32_t nf,
                                                       const FixedArgs &a, const R2CSplitCache<N> &rc)
{
   static_assert(Fixed<N>::P.threads == kFixedThreads, "groups of several transforms: 256 work items");
   constexpr int      FPB    = Fixed<N>::P.FPB;
   constexpr uint32_t STRIDE = fixed_stride<N>();
   constexpr int      iters  = R2CSplitCache<N>::iters;
   uint32_t           w      = threadIdx.x;
   asm volatile("" : "+v"(w));   // slots recomputed per group, not hoisted into registers
#pragma unroll
   for (int f = 0; f < FPB; ++f) {
      if ((uint32_t) f >= nf) break;
      const float2 *src = smem + f * STRIDE;
      float2       *out = a.y + (size_t) (first + (uint32_t) f) * a.outStride;
      float2        zi[iters], zm[iters];
#pragma unroll
      for (int c = 0; c < iters; ++c) {
         const uint32_t i = min(w + (uint32_t) (c * kFixedThreads), (uint32_t) N - 1u);
         zi[c]            = src[fslot(fixed_position<N>(i))];
         zm[c]            = src[fslot(fixed_position<N>(i == 0u ? 0u : (uint32_t) N - i))];
      }
#pragma unroll
      for (int c = 0; c < iters; ++c) {
         const uint32_t i = threadIdx.x + (uint32_t) (c * kFixedThreads);
         if (N % kFixedThreads != 0 && i >= (uint32_t) N) break;
         const float2 A = rc.A[c], B = rc.B[c];
         float2       t1;
         if (i == 0u) t1 = make_float2(zi[c].x * B.x + zi[c].y * B.y, zi[c].x * B.x - zi[c].y * B.y);
         else t1 = make_float2(zm[c].x * B.x + zm[c].y * B.y, zm[c].x * B.y - zm[c].y * B.x);
         const float2 y = make_float2(zi[c].x * A.x - zi[c].y * A.y + t1.x, zi[c].x * A.y + zi[c].y * A.x + t1.y);
         out[i]         = y;
         if (a.aux && i > 0u) out[2u * N - i] = make_float2(y.x, -y.y);
         if (i == 0u) out[N] = make_float2(zi[c].x - zi[c].y, 0.0f);
      }
   }
}

// C2C / C2R store of a group whose outputs are contiguous (no output offsets,
// stride N complex points) for lengths with at most 8 work items per
// transform (the C2R output is the inner complex FFT's): work item w
// writes points o = w + 256c of the group's FPB * N outputs, so every store
// instruction covers consecutive addresses (per transform, T work items would
// otherwise write runs of T points, up to 64 lines per store instruction).
// 503 MB per call: N = 16 1243 -> 222 us, 32 1113 -> 224 us, 64 315 -> 226
// us, 128 254 -> 229 us, 48 326 -> 247 us, 96 274 -> 250 us; C2R 32 1030 ->
// 318 us, 64 398 -> 321 us, 128 353 -> 316 us.
template <int N>
constexpr bool c2c_linear_store()
{
   return Fixed<N>::P.T <= 8 && (Fixed<N>::P.FPB * N) % kFixedThreads == 0;
}

template <int N>
__device__ __forceinline__ void c2c_store_linear(const float2 *smem, uint32_t first, uint32_t nf, const FixedArgs &a)
{
   static_assert(Fixed<N>::P.threads == kFixedThreads, "groups of several transforms: 256 work items");
   constexpr int      FPB    = Fixed<N>::P.FPB;
   constexpr uint32_t STRIDE = fixed_stride<N>();
   constexpr int      IT     = FPB * N / kFixedThreads;
   // chunks divide IT, so every LDS read below stays inside the group's
   // FPB transforms (N = 48 and 96 have IT = 12: chunks of 6)
   constexpr int      CH     = largest_divisor_upto(IT, 8);
   static_assert(IT % CH == 0 && FPB * N % kFixedThreads == 0, "linear store chunks must tile the group");
   float2            *out    = a.y + (size_t) first * N;
   const float        sy     = a.mode == C2C_INVERSE_PRESCALE ? -1.0f : 1.0f;
   const bool         pdiv   = a.mode == C2C_INVERSE_POSTDIV;
#pragma unroll 1
   for (int c0 = 0; c0 < IT; c0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t o = threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads);
         const uint32_t f = o / (uint32_t) N, k = o - f * (uint32_t) N;
         v[c]             = smem[f * STRIDE + fslot(fixed_position<N>(k))];
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t o = threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads);
         if (o / (uint32_t) N >= nf) break;
         store_stream(out + o, pdiv ? out_postdiv(v[c], (float) N) : make_float2(v[c].x, sy * v[c].y));
      }
   }
}

// R2C (outmode 0) store of a group whose output rows of N + 1 bins are
// contiguous, for lengths with at most 2 work items per transform (N = 32):
// work item w writes bins o = w + 256c of the group's FPB (N + 1) outputs
// (transform o / (N + 1), bin o % (N + 1)), so every store instruction
// covers consecutive addresses. Same formula per bin as fixed_store,
// including bin 0 and the Nyquist bin N. 503 MB per call: N = 32 447.5 ->
// 354.0 us; at N = 64 and 128 it raised register use past 3 workgroups per
// CU and was 4-8% slower, so they keep the per-transform store.
template <int N>
constexpr bool r2c_linear_store()
{
   return Fixed<N>::P.T <= 2 && !r2c_group_store<N>();
}

template <int N>
__device__ __forceinline__ void r2c_store_linear(const float2 *smem, uint32_t first, uint32_t nf, const FixedArgs &a)
{
   static_assert(Fixed<N>::P.threads == kFixedThreads, "groups of several transforms: 256 work items");
   constexpr int      FPB    = Fixed<N>::P.FPB;
   constexpr uint32_t STRIDE = fixed_stride<N>();
   constexpr uint32_t ROW    = (uint32_t) N + 1u;
   constexpr int      IT     = (int) ((FPB * ROW + kFixedThreads - 1) / kFixedThreads);
   constexpr int      CH     = 8;
   float2            *out    = a.y + (size_t) first * ROW;
   const uint32_t     total  = nf * ROW;
#pragma unroll 1
   for (int c0 = 0; c0 < IT; c0 += CH) {
      float2 zi[CH], zm[CH], A[CH], B[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         // bins past the group (or past IT chunks) re-read bin 0 of transform 0 and store nothing
         uint32_t o = threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads);
         if (c0 + c >= IT || o >= total) o = 0u;
         const uint32_t f  = o / ROW, i = o - f * ROW;
         const uint32_t ii = i < (uint32_t) N ? i : 0u;   // the Nyquist bin reads Z[0]
         const float2  *src = smem + f * STRIDE;
         zi[c]              = src[fslot(fixed_position<N>(ii))];
         zm[c]              = src[fslot(fixed_position<N>(ii == 0u ? 0u : (uint32_t) N - ii))];
         A[c]               = split_factor(a, ii);
         B[c]               = split_factor(a, 2u * N - 1u - ii);
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t o = threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads);
         if (c0 + c >= IT || o >= total) break;
         const uint32_t f = o / ROW, i = o - f * ROW;
         if (i == (uint32_t) N) {
            out[o] = make_float2(zi[c].x - zi[c].y, 0.0f);
            continue;
         }
         float2 t1;
         if (i == 0u) t1 = make_float2(zi[c].x * B[c].x + zi[c].y * B[c].y, zi[c].x * B[c].x - zi[c].y * B[c].y);
         else t1 = make_float2(zm[c].x * B[c].x + zm[c].y * B[c].y, zm[c].x * B[c].y - zm[c].y * B[c].x);
         out[o] = make_float2(zi[c].x * A[c].x - zi[c].y * A[c].y + t1.x, zi[c].x * A[c].y + zi[c].y * A[c].x + t1.y);
      }
   }
}

// Store stage of a group: every work item takes part (R2C with FPB > 1, C2C
// and C2R with contiguous outputs of lengths up to 128), or work item (fl, t) writes
// transform first + fl.
template <int N, int KIND, bool C>
__device__ __forceinline__ void group_store(const float2 *smem, const float2 *data, uint32_t first, uint32_t nf,
                                            uint32_t fl, uint32_t t, const FixedArgs &a,
                                            const StoreMap<N, KIND, C> &map)
{
   if constexpr (KIND == KIND_R2C && r2c_group_store<N>()) r2c_store_group<N>(smem, first, nf, a);
   else if constexpr (KIND == KIND_R2C && r2c_linear_store<N>()) {
      if (a.aux == 0u && a.outOff == nullptr && a.outStride == (uint32_t) N + 1u)
         r2c_store_linear<N>(smem, first, nf, a);
      else if (fl < nf) fixed_store<N, KIND>(data, first + fl, t, a, map);
   }
   else if constexpr ((KIND == KIND_C2C || KIND == KIND_C2R) && c2c_linear_store<N>()) {
      if (a.outOff == nullptr && a.outStride == (uint32_t) N) c2c_store_linear<N>(smem, first, nf, a);
      else if (fl < nf) fixed_store<N, KIND>(data, first + fl, t, a, map);
   }
   else if (fl < nf) fixed_store<N, KIND>(data, first + fl, t, a, map);
}

// Staging read of a group into LDS: work item threadIdx.x takes points
// o = threadIdx.x + 256*c of the group (transform o / N, point o % N, so the
// reads of each transform are coalesced). When a work item has at most 16
// points, all its reads are issued before its LDS writes (one round trip of
// HBM latency per group instead of one per point); lanes past the batch read
// point o % N of the group's first transform and do not write.
template <int N, int KIND, uint32_t STRIDE>
__device__ __forceinline__ void stage_group(const FixedArgs &a, uint32_t first, uint32_t nf, InScale sc, float2 *smem)
{
   constexpr int PTS = Fixed<N>::P.FPB * N;
   constexpr int TH  = Fixed<N>::P.threads;
   constexpr int IT  = (PTS + TH - 1) / TH;
   if constexpr (KIND != KIND_C2R && IT <= 16) {
      // The two address forms are separate uniform paths: with the offset
      // table looked up per point inside one loop, every data load waited
      // for the one before it (the wait for an offset also covers the earlier
      // data loads), 16 HBM round trips per group (N = 256: 123 us per 2^25
      // points).
      float2 v[IT];
      if (a.inOff && Fixed<N>::P.FPB <= 64) {
         // lane j of each wave loads the offset of the group's transform j;
         // the others come by lane shuffle (LDS crossbar, no memory wait)
         const uint32_t lane = threadIdx.x % 64u;
         const uint64_t mine = a.inOff[first + (lane < nf ? lane : 0u)];
         const uint32_t lo = (uint32_t) mine, hi = (uint32_t) (mine >> 32);
#pragma unroll
         for (int c = 0; c < IT; ++c) {
            const uint32_t o  = threadIdx.x + (uint32_t) (c * TH);
            const uint32_t f  = o / (uint32_t) N;
            const int      fs = (int) (f < nf ? f : 0u);
            const uint64_t b  = ((uint64_t) (uint32_t) __shfl((int) hi, fs) << 32) | (uint32_t) __shfl((int) lo, fs);
            v[c]              = fixed_input<N, KIND>(a.x + b, o - f * (uint32_t) N, a);
         }
      }
      else if (a.inOff) {
         uint64_t off[IT];
#pragma unroll
         for (int c = 0; c < IT; ++c) {
            const uint32_t f = (threadIdx.x + (uint32_t) (c * TH)) / (uint32_t) N;
            off[c]           = a.inOff[first + (f < nf ? f : 0u)];
         }
#pragma unroll
         for (int c = 0; c < IT; ++c) {
            const uint32_t o = threadIdx.x + (uint32_t) (c * TH);
            const uint32_t f = o / (uint32_t) N;
            v[c]             = fixed_input<N, KIND>(a.x + off[c], o - f * (uint32_t) N, a);
         }
      }
      else {
#pragma unroll
         for (int c = 0; c < IT; ++c) {
            const uint32_t o = threadIdx.x + (uint32_t) (c * TH);
            const uint32_t f = o / (uint32_t) N;
            const size_t   b = (size_t) (first + (f < nf ? f : 0u)) * a.inStride;
            v[c]             = fixed_input<N, KIND>(a.x + b, o - f * (uint32_t) N, a);
         }
      }
#pragma unroll
      for (int c = 0; c < IT; ++c) {
         const uint32_t o = threadIdx.x + (uint32_t) (c * TH);
         const uint32_t f = o / (uint32_t) N;
         const uint32_t k = o - f * (uint32_t) N;
         if (f < nf) smem[f * STRIDE + fslot(k)] = make_float2(sc.sx * v[c].x, sc.sy * v[c].y);
      }
   }
   else {
      for (uint32_t i = threadIdx.x; i < nf * N; i += TH) {
         const uint32_t f = i / (uint32_t) N;
         const uint32_t k = i - f * (uint32_t) N;
         const float2   z = fixed_input<N, KIND>(a.x + in_base(a, first + f), k, a);
         smem[f * STRIDE + fslot(k)] = make_float2(sc.sx * z.x, sc.sy * z.y);
      }
   }
}

__device__ __forceinline__ float2 c2r_presplit(float2 x, float2 xm, float2 A, float2 B)
{
   return make_float2(x.x * A.x - x.y * A.y + (xm.x * B.x + xm.y * B.y),
                      x.x * A.y + x.y * A.x + (xm.x * B.y - xm.y * B.x));
}

template <int N>
__device__ __forceinline__ void c2r_presplit_group(const FixedArgs &a, uint32_t first, uint32_t nf, InScale sc,
                                                   float2 *smem)
{
   static_assert(!c2r_group_presplit<N>() || Fixed<N>::P.threads == kFixedThreads, "groups of several transforms: 256 work items");
   constexpr int      FPB    = Fixed<N>::P.FPB;
   constexpr uint32_t STRIDE = fixed_stride<N>();
   constexpr uint32_t H      = (uint32_t) N / 2u;
   constexpr int      GP     = (int) ((H + kFixedThreads - 1) / kFixedThreads);
   constexpr int      GCH    = largest_divisor_upto(GP, 4);
#pragma unroll 1
   for (int c0 = 0; c0 < GP; c0 += GCH) {
      float2 Ai[GCH], Bi[GCH], Aj[GCH], Bj[GCH];
#pragma unroll
      for (int c = 0; c < GCH; ++c) {
         // pairs past N/2 (when 256 does not divide N/2) re-read pair N/2 - 1 and store nothing
         const uint32_t i  = min(threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads), H - 1u);
         const uint32_t j  = i == 0u ? H : (uint32_t) N - i;
         Ai[c]             = split_factor(a, i);
         Bi[c]             = split_factor(a, (uint32_t) N + i);
         Aj[c]             = split_factor(a, j);
         Bj[c]             = split_factor(a, (uint32_t) N + j);
      }
#pragma unroll
      for (int f = 0; f < FPB; ++f) {
         if ((uint32_t) f >= nf) break;
         const float2 *in  = a.x + in_base(a, first + (uint32_t) f);
         float2       *dst = smem + f * STRIDE;
         float2        xi[GCH], xm[GCH];
         // bin N/2, the partner of Z[N/2] that work item 0 forms, loaded with
         // the others by every work item (one address per wavefront): loaded
         // where it is used, its round trip held up the whole group
         const float2 xh = in[H];
#pragma unroll
         for (int c = 0; c < GCH; ++c) {
            const uint32_t i = min(threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads), H - 1u);
            xi[c]            = in[i];
            xm[c]            = in[(uint32_t) N - i];   // i = 0: the Nyquist bin
         }
#pragma unroll
         for (int c = 0; c < GCH; ++c) {
            const uint32_t i = threadIdx.x + (uint32_t) ((c0 + c) * kFixedThreads);
            if (H % kFixedThreads != 0 && i >= H) break;
            const float2 zi = c2r_presplit(xi[c], xm[c], Ai[c], Bi[c]);
            dst[fslot(i)]   = make_float2(sc.sx * zi.x, sc.sy * zi.y);
            if (i != 0u) {
               const float2 zj          = c2r_presplit(xm[c], xi[c], Aj[c], Bj[c]);
               dst[fslot((uint32_t) N - i)] = make_float2(sc.sx * zj.x, sc.sy * zj.y);
            }
            else {
               const float2 zh = c2r_presplit(xh, xh, Aj[c], Bj[c]);
               dst[fslot(H)]   = make_float2(sc.sx * zh.x, sc.sy * zh.y);
            }
         }
      }
   }
}

// Prefetching form of the inMode 0 group pre-split (lengths with
// c2r_group_presplit and 4 transforms per group): the four split factors of
// each of a work item's mirror pairs are loaded once per workgroup
// (C2RSplitCache), and the next group's bins (both halves of every pair and
// the Nyquist bin of each transform) are loaded before the current group's
// LDS passes (C2RGroupInput), so the passes overlap the next group's HBM
// reads. Same-box A/B, 537 MB per call: C2R 2048 (inner 1024) 109.9 ->
// 95.9 us; at inner 2048 (2 transforms per group, 168 VGPRs) it was slower
// (104.9 -> 117.7 us), so that length keeps the plain group pre-split.
template <int N>
constexpr bool c2r_prefetch()
{
   return c2r_group_presplit<N>() && Fixed<N>::P.FPB >= 4;
}

template <int N>
struct C2RSplitCache {
   static constexpr uint32_t H  = (uint32_t) N / 2u;
   static constexpr int      GP = (int) ((H + kFixedThreads - 1) / kFixedThreads);
   float2                    Ai[GP], Bi[GP], Aj[GP], Bj[GP];

   __device__ __forceinline__ void init(const FixedArgs &a)
   {
#pragma unroll
      for (int c = 0; c < GP; ++c) {
         const uint32_t i = min(threadIdx.x + (uint32_t) (c * kFixedThreads), H - 1u);
         const uint32_t j = i == 0u ? H : (uint32_t) N - i;
         Ai[c]            = split_factor(a, i);
         Bi[c]            = split_factor(a, (uint32_t) N + i);
         Aj[c]            = split_factor(a, j);
         Bj[c]            = split_factor(a, (uint32_t) N + j);
      }
   }
};

template <int N>
struct C2RGroupInput {
   static constexpr int      FPB = Fixed<N>::P.FPB;
   static constexpr uint32_t H   = C2RSplitCache<N>::H;
   static constexpr int      GP  = C2RSplitCache<N>::GP;
   float2                    xi[FPB][GP], xm[FPB][GP], xh[FPB];

   // transforms past the batch re-read its last one and are not computed
   __device__ __forceinline__ void load(const FixedArgs &a, uint32_t first, uint32_t nf)
   {
#pragma unroll
      for (int f = 0; f < FPB; ++f) {
         const float2 *in = a.x + (size_t) (first + min((uint32_t) f, nf - 1u)) * a.inStride;
#pragma unroll
         for (int c = 0; c < GP; ++c) {
            const uint32_t i = min(threadIdx.x + (uint32_t) (c * kFixedThreads), H - 1u);
            xi[f][c]         = in[i];
            xm[f][c]         = in[(uint32_t) N - i];   // i = 0: the Nyquist bin
         }
         xh[f] = in[H];
      }
   }

   __device__ __forceinline__ void compute(const C2RSplitCache<N> &cc, uint32_t nf, InScale sc, float2 *smem) const
   {
      constexpr uint32_t STRIDE = fixed_stride<N>();
#pragma unroll
      for (int f = 0; f < FPB; ++f) {
         if ((uint32_t) f >= nf) break;
         float2 *dst = smem + f * STRIDE;
#pragma unroll
         for (int c = 0; c < GP; ++c) {
            const uint32_t i = threadIdx.x + (uint32_t) (c * kFixedThreads);
            if (H % kFixedThreads != 0 && i >= H) break;
            const float2 zi = c2r_presplit(xi[f][c], xm[f][c], cc.Ai[c], cc.Bi[c]);
            dst[fslot(i)]   = make_float2(sc.sx * zi.x, sc.sy * zi.y);
            if (i != 0u) {
               const float2 zj              = c2r_presplit(xm[f][c], xi[f][c], cc.Aj[c], cc.Bj[c]);
               dst[fslot((uint32_t) N - i)] = make_float2(sc.sx * zj.x, sc.sy * zj.y);
            }
            else {
               const float2 zh = c2r_presplit(xh[f], xh[f], cc.Aj[c], cc.Bj[c]);
               dst[fslot(H)]   = make_float2(sc.sx * zh.x, sc.sy * zh.y);
            }
         }
      }
   }
};

template <int N>
constexpr bool fixed_direct()
{
   return Pass0<N>::q2 >= 32;
}

// Waves per SIMD of a kernel whose workgroups of `threads` work items use
// `lds` bytes of LDS each (160 KB per CU, four SIMDs), at most 8.
constexpr int lds_waves_per_simd(size_t lds, int threads)
{
   const size_t wgs = lds ? (size_t) 163840 / lds : 8u;
   const size_t w   = (wgs < 1 ? 1 : wgs) * (size_t) (threads / 256);
   return w < 1 ? 1 : w > 8 ? 8 : (int) w;
}

// Points per work item that a group's first pass may hold in registers (all
// its inputs loaded at once: one HBM round trip per group): 16 at four or
// more waves per SIMD, more where fewer waves leave more registers per work
// item.
constexpr int pass0_register_points(int wavesPerSimd)
{
   return wavesPerSimd >= 4 ? 16 : wavesPerSimd >= 2 ? 32 : 64;
}

// The twiddle table is staged in LDS (once per workgroup) when it fits next
// to the data: vector loads of twiddles from L2 would otherwise queue behind
// the prefetched input loads (vmcnt completes in order).
template <int N>
constexpr bool fixed_tw_lds()
{
   constexpr FixedPlan P = Fixed<N>::P;
   return ((size_t) P.FPB * fixed_stride<N>() + (size_t) P.twCount) * sizeof(float2) <= 163840u;
}

template <int N>
constexpr size_t fixed_lds_bytes()
{
   constexpr FixedPlan P = Fixed<N>::P;
   return ((size_t) P.FPB * fixed_stride<N>() + (fixed_tw_lds<N>() ? (size_t) P.twCount : 0u)) * sizeof(float2);
}

// Registers held across the LDS passes by the next group's first-pass inputs
// are bounded: lengths whose first pass needs more points per work item than
// pass0_register_points allows load them right before use instead. The C2R
// pre-split combines four loads per point as they arrive, so C2R does not
// prefetch.
template <int N, int KIND>
constexpr bool fixed_prefetch()
{
   return KIND != KIND_C2R && fixed_direct<N>() &&
          Pass0<N>::iters * Pass0<N>::R <=
              pass0_register_points(lds_waves_per_simd(fixed_lds_bytes<N>(), Fixed<N>::P.threads));
}

// The work item's index within its transform as used inside one group of
// the persistent loop. For the one-transform workgroups of N > 4096 it is
// re-read per group (an empty asm the compiler cannot see through), so that
// the LDS slots and twiddle offsets of every pass are recomputed per group:
// hoisted out of the loop they held about 100 registers (N = 16384: 212
// bytes of spills per work item, whose reloads also waited for the
// prefetched loads).
template <int N>
__device__ __forceinline__ uint32_t group_item(uint32_t t)
{
   if constexpr (Fixed<N>::P.threads > kFixedThreads) asm volatile("" : "+v"(t));
   return t;
}

// The dedicated kernels address their channels by stride only (launch_fixed
// refuses offset tables, which only the class kernels take): with the
// tables cleared here the compiler drops their loads. A table lookup in the
// address of a prefetched load or of a store was a vector-memory load whose
// wait, the counter completing in order, also waited for every load and
// store before it: the prefetched next group before each store phase, the
// previous group's stores before each prefetch.
// C2R kernels are compiled for at least 3 waves per SIMD: with it the
// prefetching pre-split of C2R 2048 (159 VGPRs either way) ran 95.9 instead
// of 99.4 us per 537 MB (same-box A/B).
template <int KIND>
constexpr int fixed_min_waves()
{
   return KIND == KIND_C2R ? 3 : 1;
}

template <int N, int KIND>
__global__ void __launch_bounds__(Fixed<N>::P.threads) __attribute__((amdgpu_waves_per_eu(fixed_min_waves<KIND>())))
fixed_kernel(FixedArgs args)
{
   FixedArgs a = args;
   a.inOff     = nullptr;
   a.outOff    = nullptr;
   constexpr FixedPlan P      = Fixed<N>::P;
   constexpr int       T      = P.T;
   constexpr int       FPB    = P.FPB;
   constexpr uint32_t  STRIDE = fixed_stride<N>();
   extern __shared__ float2 smem[];
   const uint32_t fl   = threadIdx.x / T;
   const uint32_t t    = threadIdx.x % T;
   const InScale  sc   = in_scale(a.mode, 1.0f / (float) N);
   float2        *data = smem + fl * STRIDE;
   const float2  *twp  = fixed_tw_lds<N>() ? (const float2 *) (smem + FPB * STRIDE) : a.tw;
   StoreMap<N, KIND> map;
   map.init(t);

   if constexpr (fixed_prefetch<N, KIND>()) {
      using Q = Pass0<N>;
      constexpr bool kR2CCache = KIND == KIND_R2C && r2c_group_store<N>();
      R2CSplitCache<kR2CCache ? N : 256> rc;
      if constexpr (kR2CCache) rc.init(a);
      float2   v[Q::iters][Q::R];
      uint32_t g = blockIdx.x;
      if (g < a.numGroups) pass0_load<N, KIND>(v, g * FPB, fl, min((uint32_t) FPB, a.numChannels - g * FPB), t, a);
      if constexpr (fixed_tw_lds<N>()) {
         for (uint32_t i = threadIdx.x; i < (uint32_t) P.twCount; i += P.threads) smem[FPB * STRIDE + i] = a.tw[i];
         __syncthreads();
      }
      for (; g < a.numGroups; g += gridDim.x) {
         const uint32_t first = g * FPB;
         const uint32_t nf    = min((uint32_t) FPB, a.numChannels - first);
         const uint32_t tg    = group_item<N>(t);
         pass0_compute<N>(v, data, tg, twp, sc);
         const uint32_t gn = g + gridDim.x;
         if (gn < a.numGroups) pass0_load<N, KIND>(v, gn * FPB, fl, min((uint32_t) FPB, a.numChannels - gn * FPB), tg, a);
         __syncthreads();
         fixed_passes_from<N, 1>(data, tg, twp);
         if constexpr (kR2CCache) r2c_store_group_cached<N>(smem, first, nf, a, rc);
         else group_store<N, KIND>(smem, data, first, nf, fl, tg, a, map);
         __syncthreads();
      }
   }
   else if constexpr (KIND == KIND_C2R && N % T == 0 && N >= 256) {
      // C2R: the pre-split points go from coalesced HBM reads into LDS, then
      // every pass runs from LDS. inMode 0: work item t takes the mirror
      // pairs (i, N-i), i = t + T*it < N/2, reading each input bin once
      // (X[i] and X[N-i] feed both Z[i] and Z[N-i]); i = 0 pairs bin 0 with
      // the Nyquist bin N and also forms the self-mirrored Z[N/2]. inMode 1
      // (V[i] = X[N+i], no mirror): work item t computes points t + T*it, 8
      // at a time.
      if constexpr (fixed_tw_lds<N>()) {
         for (uint32_t i = threadIdx.x; i < (uint32_t) P.twCount; i += P.threads) smem[FPB * STRIDE + i] = a.tw[i];
      }
      if constexpr (c2r_prefetch<N>()) {
         if (a.aux == 0u) {
            C2RSplitCache<N> cc;
            cc.init(a);
            C2RGroupInput<N> in;
            uint32_t         g = blockIdx.x;
            if (g < a.numGroups) in.load(a, g * FPB, min((uint32_t) FPB, a.numChannels - g * FPB));
            for (; g < a.numGroups; g += gridDim.x) {
               const uint32_t first = g * FPB;
               const uint32_t nf    = min((uint32_t) FPB, a.numChannels - first);
               uint32_t       tg    = t;   // per-group LDS slots and offsets, not hoisted into registers
               asm volatile("" : "+v"(tg));
               in.compute(cc, nf, sc, smem);
               const uint32_t gn = g + gridDim.x;
               if (gn < a.numGroups) in.load(a, gn * FPB, min((uint32_t) FPB, a.numChannels - gn * FPB));
               __syncthreads();
               fixed_passes_from<N, 0>(data, tg, twp);
               group_store<N, KIND>(smem, data, first, nf, fl, tg, a, map);
               __syncthreads();
            }
            return;
         }
      }
      constexpr int iters = N / T;
      constexpr int CH    = largest_divisor_upto(iters, 8);
      constexpr int HALF  = ((N / 2) + T - 1) / T;   // pairs per work item
      constexpr int PCH   = largest_divisor_upto(HALF, 4);
      // Z = X A' + conj(Xm) B' for a bin X, its mirror Xm and factors A', B'
      auto presplit = [](float2 x, float2 xm, float2 A, float2 B) __attribute__((always_inline)) {
         return make_float2(x.x * A.x - x.y * A.y + (xm.x * B.x + xm.y * B.y),
                            x.x * A.y + x.y * A.x + (xm.x * B.y - xm.y * B.x));
      };
      for (uint32_t g = blockIdx.x; g < a.numGroups; g += gridDim.x) {
         const uint32_t first = g * FPB;
         const uint32_t nf    = min((uint32_t) FPB, a.numChannels - first);
         if (c2r_group_presplit<N>() && a.aux == 0u) {
            c2r_presplit_group<N>(a, first, nf, sc, smem);
         }
         else if (fl < nf && a.aux == 0u) {
            const float2 *in = a.x + in_base(a, first + fl);
            const float2  xh = in[N / 2];   // bin N/2 for Z[N/2], loaded up front by every work item
#pragma unroll 1
            for (int c0 = 0; c0 < HALF; c0 += PCH) {
               float2 xi[PCH], xm[PCH];
#pragma unroll
               for (int c = 0; c < PCH; ++c) {
                  // pairs past N/2 (when T does not divide N/2) re-read pair
                  // N/2 - 1 and store nothing
                  const uint32_t i = min(t + (uint32_t) ((c0 + c) * T), (uint32_t) N / 2u - 1u);
                  xi[c]            = in[i];
                  xm[c]            = in[(uint32_t) N - i];   // i = 0: the Nyquist bin
               }
#pragma unroll
               for (int c = 0; c < PCH; ++c) {
                  const uint32_t i  = t + (uint32_t) ((c0 + c) * T);
                  if ((N / 2) % T != 0 && i >= (uint32_t) N / 2u) break;
                  const float2   zi = presplit(xi[c], xm[c], split_factor(a, i), split_factor(a, (uint32_t) N + i));
                  data[fslot(i)]    = make_float2(sc.sx * zi.x, sc.sy * zi.y);
                  if (i != 0u) {
                     const uint32_t j  = (uint32_t) N - i;
                     const float2   zj = presplit(xm[c], xi[c], split_factor(a, j), split_factor(a, (uint32_t) N + j));
                     data[fslot(j)]    = make_float2(sc.sx * zj.x, sc.sy * zj.y);
                  }
                  else {
                     constexpr uint32_t H  = (uint32_t) N / 2u;
                     const float2       zh = presplit(xh, xh, split_factor(a, H), split_factor(a, (uint32_t) N + H));
                     data[fslot(H)]        = make_float2(sc.sx * zh.x, sc.sy * zh.y);
                  }
               }
            }
         }
         else if (fl < nf) {
            const float2 *in = a.x + in_base(a, first + fl);
#pragma unroll 1
            for (int c0 = 0; c0 < iters; c0 += CH) {
               float2 z[CH];
#pragma unroll
               for (int c = 0; c < CH; ++c) z[c] = fixed_input<N, KIND>(in, t + (uint32_t) ((c0 + c) * T), a);
#pragma unroll
               for (int c = 0; c < CH; ++c)
                  data[fslot(t + (uint32_t) ((c0 + c) * T))] = make_float2(sc.sx * z[c].x, sc.sy * z[c].y);
            }
         }
         __syncthreads();
         fixed_passes_from<N, 0>(data, t, twp);
         group_store<N, KIND>(smem, data, first, nf, fl, t, a, map);
         __syncthreads();
      }
   }
   else if constexpr (fixed_direct<N>()) {
      using Q = Pass0<N>;
      if constexpr (fixed_tw_lds<N>()) {
         for (uint32_t i = threadIdx.x; i < (uint32_t) P.twCount; i += P.threads) smem[FPB * STRIDE + i] = a.tw[i];
         __syncthreads();
      }
      for (uint32_t g = blockIdx.x; g < a.numGroups; g += gridDim.x) {
         const uint32_t first = g * FPB;
         const uint32_t nf    = min((uint32_t) FPB, a.numChannels - first);
         const uint32_t tg    = group_item<N>(t);
#pragma unroll 1
         for (uint32_t u0 = 0; u0 < (uint32_t) Q::items; u0 += T) {
            float2 v[Q::R];
            pass0_load_one<N, KIND>(v, first, fl, nf, tg + u0, a);
            pass0_compute_one<N>(v, data, tg + u0, twp, sc);
         }
         __syncthreads();
         fixed_passes_from<N, 1>(data, tg, twp);
         group_store<N, KIND>(smem, data, first, nf, fl, tg, a, map);
         __syncthreads();
      }
   }
   else {
      if constexpr (fixed_tw_lds<N>()) {
         for (uint32_t i = threadIdx.x; i < (uint32_t) P.twCount; i += P.threads) smem[FPB * STRIDE + i] = a.tw[i];
      }
      for (uint32_t g = blockIdx.x; g < a.numGroups; g += gridDim.x) {
         const uint32_t first = g * FPB;
         const uint32_t nf    = min((uint32_t) FPB, a.numChannels - first);
         const uint32_t tg    = group_item<N>(t);
         stage_group<N, KIND, STRIDE>(a, first, nf, sc, smem);
         __syncthreads();
         fixed_passes_from<N, 0>(data, tg, twp);
         group_store<N, KIND>(smem, data, first, nf, fl, tg, a, map);
         __syncthreads();
      }
   }
}

// ---- multi-length ("class") kernels -------------------------------------------
// One kernel covers every length of an octave (lengths whose work-item count T
// and transforms per workgroup agree); it runs a list of tasks (a length, its
// channels, twiddles, offsets), each cut into groups of FPB transforms, as a
// persistent loop over (task, group) items. Twiddles are read from global
// memory (L2) rather than staged in LDS, since consecutive items of one
// workgroup may have different lengths. Ragged batches put all lengths of an
// octave in one launch; single lengths without a dedicated kernel use it with
// one task passed by value.
struct ClassTask {
   uint32_t        n;            // one of the class's lengths
   uint32_t        numChannels;  // channels (columns / rows in the two-pass layouts)
   uint32_t        numGroups;
   uint32_t        layout;       // LAYOUT_PLAIN / LAYOUT_COLUMNS / LAYOUT_ROWS
   const float2   *tw;           // twiddle table of length n
   const uint64_t *inOff;        // plain and columns: complex offset of each transform in x
   const uint64_t *outOff;       // plain and rows: ... in y
   uint64_t        scratchOff;   // two-pass: complex offset of the task's scratch region
   uint32_t        span, fullN;  // two-pass: see FixedArgs
   const float2   *colTw;
   const float2   *colTwHi;
   uint32_t        colBits;
   uint32_t        scrStride;    // two-pass: see FixedArgs
   uint32_t        rowAlign;     // two-pass: see FixedArgs
};

// Work unit of the persistent ragged launch: groups [g0, g1) of task, then,
// if task2 != kNoTask, groups [h0, h1) of task2 after a workgroup barrier
// (a long transform's column pass, then its row pass, in one workgroup).
constexpr uint32_t kNoTask = 0xFFFFFFFFu;
struct RaggedUnit {
   uint32_t task, g0, g1;
   uint32_t task2, h0, h1;
};

struct ClassArgs {
   const float2    *x;
   float2          *y;
   float2          *scratch;    // two-pass intermediate (columns write, rows read)
   int              mode;       // C2CMode
   uint32_t         numItems;
   const ClassTask *tasks;      // null: one task, `single`, with item = group index
   const uint32_t  *items;      // (task << 20) | group
   const RaggedUnit *units;     // ragged_class_kernel: the units, in queue order (numItems of them)
   uint32_t        *counters;   // ragged_class_kernel: [0] queue head, [1] workgroups done
   ClassTask        single;
};

constexpr uint32_t kClassGroupBits = 20;
constexpr uint32_t kXcds           = 8;   // accelerator dies (XCDs) of an MI355X, each with its own L2

// Per-transform LDS stride of the class kernels: odd, so that the FPB
// channels of a column or row group (lanes j = 0..FPB-1 at the same point)
// fall in distinct LDS banks.
template <int N>
constexpr uint32_t class_stride()
{
   return fixed_stride<N>() | 1u;
}

// Column group: the FPB adjacent columns of the group are loaded row by row
// (lane j of a row takes column j), transformed in LDS, multiplied by
// W_fullN^(p*k1) (product of two small-table entries) and stored row by row
// into the intermediate copy, whose
// rows have scrStride points. Channels number scrStride columns per
// transform so that a group never straddles two transforms; columns from
// span on are idle (they read column span-1 and store nothing).
// TW_LDS: the column tables are the LDS copy at colLds (low table, then the
// high table at 2^colBits), else a.colTw / a.colTwHi in global memory.
template <int N, bool TW_LDS>
__device__ __forceinline__ void column_group(const FixedArgs &a, uint32_t g, float2 *smem, const float2 *colLds)
{
   constexpr FixedPlan P      = Fixed<N>::P;
   constexpr int       T      = P.T;
   constexpr int       FPB    = P.FPB;
   constexpr uint32_t  STRIDE = class_stride<N>();
   constexpr int       IT     = (N + T - 1) / T;   // points per lane
   constexpr int       CH     = IT < 16 ? IT : 16;
   const uint32_t      j      = threadIdx.x % FPB;
   const uint32_t      i0     = threadIdx.x / FPB;
   const uint32_t      ch     = g * FPB + j;
   const uint32_t      chc    = ch < a.numChannels ? ch : g * FPB;   // lanes past the batch re-read the first column
   const uint32_t      f      = chc / a.scrStride;
   const uint32_t      p      = chc - f * a.scrStride;
   const bool          live   = ch < a.numChannels && p < a.span;
   const float2       *in     = a.x + (a.inOff ? (size_t) a.inOff[f] : (size_t) f * a.fullN) + min(p, a.span - 1u);
   const InScale       sc     = in_scale(a.mode, 1.0f / (float) a.fullN);
   float2             *col    = smem + j * STRIDE;
#pragma unroll 1
   for (int c0 = 0; c0 < IT; c0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t i = min(i0 + (uint32_t) ((c0 + c) * T), (uint32_t) N - 1u);
         v[c]             = load_stream(in + (size_t) a.span * i);
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t i = i0 + (uint32_t) ((c0 + c) * T);
         if ((N % T == 0 && IT % CH == 0) || (c0 + c < IT && i < (uint32_t) N))
            col[fslot(i)] = make_float2(sc.sx * v[c].x, sc.sy * v[c].y);
      }
   }
   __syncthreads();
   fixed_passes_from<N, 0>(smem + (threadIdx.x / T) * STRIDE, threadIdx.x % T, a.tw);
   if (live) {
      const size_t   obase = a.inOff ? (size_t) a.inOff[f] : (size_t) f * a.fullN;   // the transform's place in yOut
      const uint32_t rot   = row_rotation(a, a.yOut + obase, (uint32_t) N);
      float2        *out   = a.y + (size_t) f * ((size_t) N * a.scrStride) + p;
#pragma unroll 1
      for (int c0 = 0; c0 < IT; c0 += CH) {
         float2 v[CH], w[CH];
#pragma unroll
         for (int c = 0; c < CH; ++c) {
            const uint32_t k = min(i0 + (uint32_t) ((c0 + c) * T), (uint32_t) N - 1u);
            v[c]             = col[fslot(fixed_position<N>(k))];
            const uint32_t e = p * k;   // < fullN
            float2         h, l;
            if constexpr (TW_LDS) {
               h = colLds[(1u << a.colBits) + (e >> a.colBits)];
               l = colLds[e & ((1u << a.colBits) - 1u)];
            }
            else {
               h = a.colTwHi[e >> a.colBits];
               l = a.colTw[e & ((1u << a.colBits) - 1u)];
            }
            w[c]             = make_float2(h.x * l.x - h.y * l.y, h.x * l.y + h.y * l.x);
         }
#pragma unroll
         for (int c = 0; c < CH; ++c) {
            const uint32_t k = i0 + (uint32_t) ((c0 + c) * T);
            const uint32_t row = k >= rot ? k - rot : k + (uint32_t) N - rot;   // k1 = (row + rot) mod N
            if ((N % T == 0 && IT % CH == 0) || (c0 + c < IT && k < (uint32_t) N))
               store_stream(out + (size_t) a.scrStride * row, make_float2(v[c].x * w[c].x - v[c].y * w[c].y,
                                                                         v[c].x * w[c].y + v[c].y * w[c].x));
         }
      }
   }
   __syncthreads();
}

// Row store: lane j of the group writes row first + j; the FPB rows of a
// group are adjacent frequencies k1, so each output k2 is a run of FPB
// consecutive points. `map` holds the LDS slots of outputs i0 + T*it,
// i0 = threadIdx.x / FPB.
template <int N, bool C>
__device__ __forceinline__ void rows_store(const float2 *smem, uint32_t first, const FixedArgs &a,
                                           const StoreMap<N, KIND_C2C, C> &map)
{
   constexpr FixedPlan P      = Fixed<N>::P;
   constexpr int       T      = P.T;
   constexpr int       FPB    = P.FPB;
   constexpr uint32_t  STRIDE = class_stride<N>();
   constexpr int       IT     = StoreMap<N, KIND_C2C, C>::iters;
   constexpr int       CH     = IT < 16 ? IT : 16;
   const uint32_t      j      = threadIdx.x % FPB;
   const uint32_t      i0     = threadIdx.x / FPB;
   const uint32_t      ch     = first + j;
   if (ch >= a.numChannels) return;
   const uint32_t f    = ch / a.span;
   const uint32_t r    = ch - f * a.span;
   float2        *outf = a.y + (a.outOff ? (size_t) a.outOff[f] : (size_t) f * a.fullN);
   const uint32_t rot  = row_rotation(a, outf, a.span);
   // Runs on whole lines or 64-byte halves (rows rotated, rowAlign 16 or 8)
   // are stored non-temporally; runs at any other offset (N1 not a multiple
   // of 8) are cached, so that the parts of a line that neighbouring row
   // groups write meet in L2 (adjacent groups run on one die,
   // fixed_class_kernel) before it is written back. Ragged 10000 per step:
   // every row store non-temporal 1.528-1.533 ms, these cached 1.460-1.462
   // ms, the 64-byte-aligned runs cached too 1.477-1.478 ms.
   const bool     full = a.rowAlign != 0u;
   float2        *out  = outf + (r + rot < a.span ? r + rot : r + rot - a.span);   // row r holds k1 = (r + rot) mod N1
   const float2  *src = smem + j * STRIDE;
   const float    sy  = a.mode == C2C_INVERSE_PRESCALE ? -1.0f : 1.0f;
   const bool     pdiv = a.mode == C2C_INVERSE_POSTDIV;
   const float    nf  = (float) a.fullN;
#pragma unroll 1
   for (int c0 = 0; c0 < IT; c0 += CH) {
      float2 v[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) v[c] = src[map.at(c0 + c, i0)];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
         const uint32_t k = i0 + (uint32_t) ((c0 + c) * T);
         if (!(N % T == 0 && IT % CH == 0) && (c0 + c >= IT || k >= (uint32_t) N)) break;
         const float2 o = pdiv ? out_postdiv(v[c], nf) : make_float2(v[c].x, sy * v[c].y);
         if (full) store_stream(out + (size_t) a.span * k, o);
         else out[(size_t) a.span * k] = o;
      }
   }
}

// BUDGET: points per work item the first pass may hold in registers
// (pass0_register_points of the calling kernel's LDS).
template <int N, int KIND, int BUDGET>
__device__ __forceinline__ void fixed_group(const FixedArgs &a, uint32_t g, float2 *smem, const float2 *colLds)
{
   constexpr FixedPlan P      = Fixed<N>::P;
   constexpr int       T      = P.T;
   constexpr int       FPB    = P.FPB;
   constexpr uint32_t  STRIDE = class_stride<N>();
   if (a.layout == LAYOUT_COLUMNS) {
      if (colLds) column_group<N, true>(a, g, smem, colLds);
      else column_group<N, false>(a, g, smem, nullptr);
      return;
   }
   const uint32_t      fl     = threadIdx.x / T;
   const uint32_t      t      = threadIdx.x % T;
   const InScale       sc     = a.layout == LAYOUT_ROWS ? InScale{1.0f, 1.0f} : in_scale(a.mode, 1.0f / (float) N);
   float2             *data   = smem + fl * STRIDE;
   const uint32_t      first  = g * FPB;
   const uint32_t      nf     = min((uint32_t) FPB, a.numChannels - first);
   if constexpr (fixed_direct<N>()) {
      using Q = Pass0<N>;
      if constexpr (Q::iters * Q::R <= BUDGET) {
         float2 v[Q::iters][Q::R];
         pass0_load<N, KIND>(v, first, fl, nf, t, a);
         pass0_compute<N>(v, data, t, a.tw, sc);
      }
      else {
#pragma unroll 1
         for (uint32_t u0 = 0; u0 < (uint32_t) Q::items; u0 += T) {
            float2 v[Q::R];
            pass0_load_one<N, KIND>(v, first, fl, nf, t + u0, a);
            pass0_compute_one<N>(v, data, t + u0, a.tw, sc);
         }
      }
      __syncthreads();
      fixed_passes_from<N, 1>(data, t, a.tw);
   }
   else {
      stage_group<N, KIND, STRIDE>(a, first, nf, sc, smem);
      __syncthreads();
      fixed_passes_from<N, 0>(data, t, a.tw);
   }
   StoreMap<N, KIND> map;
   if (a.layout == LAYOUT_ROWS) {
      map.init(threadIdx.x / FPB);
      rows_store<N>(smem, first, a, map);
   }
   else {
      map.init(t);
      if (fl < nf) fixed_store<N, KIND>(data, first + fl, t, a, map);
   }
   __syncthreads();
}

// C and K: ClassArgs and ClassTask, or the same read through the constant
// address space (scalar loads).
template <class C, class K>
__device__ __forceinline__ FixedArgs class_fixed_args(const C &c, const K &k)
{
   FixedArgs a{};
   a.x           = k.layout == LAYOUT_ROWS ? c.scratch + k.scratchOff : c.x;
   a.y           = k.layout == LAYOUT_COLUMNS ? c.scratch + k.scratchOff : c.y;
   a.tw          = k.tw;
   a.numChannels = k.numChannels;
   a.numGroups   = k.numGroups;
   a.mode        = c.mode;
   a.inStride    = k.layout == LAYOUT_ROWS ? k.scrStride : k.n;   // rows of the intermediate copy
   a.outStride   = k.n;
   a.inOff       = k.layout == LAYOUT_ROWS ? nullptr : k.inOff;
   a.outOff      = k.layout == LAYOUT_COLUMNS ? nullptr : k.outOff;
   a.layout      = k.layout;
   a.span        = k.span;
   a.fullN       = k.fullN;
   a.colTw       = k.colTw;
   a.colTwHi     = k.colTwHi;
   a.colBits     = k.colBits;
   a.scrStride   = k.scrStride;
   a.rowAlign    = k.rowAlign;
   a.yOut        = c.y;
   return a;
}

// Group g of task *tk, compiled per length as a separate function so that
// each length gets its own register allocation (all lengths inlined into one
// kernel body need several times the registers of the largest one). The task
// and the kernel arguments are passed by address (kernel-argument or global
// memory) and read through the constant address space: scalar loads, uniform
// branches, and buffer pointers the compiler knows to be global (global
// instead of flat loads and stores, which would also count against the LDS
// wait counter). Passing the address of a stack copy of the kernel argument
// instead made every field a per-lane flat load (N = 4096: 149 -> 126 us per
// 2^25 points). The data goes through the kernel's dynamic LDS. Twiddle
// tables of at most 511 entries are held in LDS across the groups of a task
// (N = 256: 176 -> 166 us per 2^25 points); longer ones are read from global
// memory (L2), since LDS room for them would cost occupancy. Measured
// alternatives that were slower on MI355X: every table staged per group or
// per task, a loop over a task's groups inside one call (its hoisted
// per-thread indexing spills at the register budget), and prefetching the
// next group's points into registers (measured with per-lane arguments).
// Lengths whose twiddle table (at most 511 entries, 4 KB) the class kernel
// keeps in LDS: staged when a workgroup's item belongs to a task with another
// table than the one held, so consecutive groups of one task load it once.
template <int N>
constexpr bool class_tw_lds()
{
   return Fixed<N>::P.twCount <= 511;
}

template <int N>
constexpr size_t class_length_lds()
{
   return (size_t) Fixed<N>::P.FPB * class_stride<N>() * sizeof(float2);
}

// LDS of a class kernel, in float2 slots: the data region of its largest
// length, then the twiddle region of its largest staged table, then one
// slot holding the address of the table staged there; then, where it costs
// no resident workgroup, the column-pass twiddle tables (kColTwSlots) and
// one slot holding their address.
template <int... Ls>
constexpr uint32_t class_data_slots()
{
   size_t m = 0;
   ((m = m > class_length_lds<Ls>() ? m : class_length_lds<Ls>()), ...);
   return (uint32_t) (m / sizeof(float2));
}

template <int... Ls>
constexpr uint32_t class_tw_slots()
{
   uint32_t m = 0;
   ((m = (class_tw_lds<Ls>() && (uint32_t) Fixed<Ls>::P.twCount > m) ? (uint32_t) Fixed<Ls>::P.twCount : m), ...);
   return m;
}

// The two column twiddle tables of a two-pass length (W_fullN^e for e <
// 2^colBits, then W_fullN^(h 2^colBits) for h < fullN / 2^colBits) held in
// LDS for the lengths up to 65536 (colBits = 8: at most 512 entries): read
// from L2 they were two gathers per output point over up to 64 lines per
// wavefront (N = 65536 = 256 x 256, 2^27 points: column pass 515 us against
// 390 us for the row pass).
constexpr uint32_t kColTwSlots = 512;

template <int... Ls>
constexpr size_t class_lds_base_bytes()
{
   return ((size_t) class_data_slots<Ls...>() + class_tw_slots<Ls...>() + 1u) * sizeof(float2);
}

template <int... Ls>
constexpr bool class_coltw_lds()
{
   constexpr size_t base = class_lds_base_bytes<Ls...>();
   constexpr size_t with = base + (kColTwSlots + 1u) * sizeof(float2);
   return (size_t) 163840 / with >= 1u && (size_t) 163840 / with == (size_t) 163840 / base;
}

template <int... Ls>
constexpr size_t class_lds_bytes()
{
   return class_lds_base_bytes<Ls...>() + (class_coltw_lds<Ls...>() ? (kColTwSlots + 1u) * sizeof(float2) : 0u);
}

// Arguments of a called function arrive in vector registers; the class
// kernel's are the same in every work item, and taking them (and the task
// addresses) back into scalar registers lets the task and the kernel
// arguments be read with scalar loads and keeps the control flow uniform.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t wave_uniform(uint64_t v)
{
   return ((uint64_t) wave_uniform((uint32_t) (v >> 32)) << 32) | wave_uniform((uint32_t) v);
}
using ClassArgsK = const __attribute__((address_space(4))) ClassArgs;
using ClassTaskK = const __attribute__((address_space(4))) ClassTask;

// task: address of the ClassTask (in the task table, or the kernel argument's
// `single`); args: address of the kernel argument (its kernarg segment).
// not_tail_called: a call marked as a tail call makes the callee save and
// restore the callee-saved registers it uses (up to 50 VGPRs, 20-30% of a
// group's time); with plain calls from the one kernel the compiler leaves
// them unsaved and the kernel does not rely on them.
// colSlot: first slot of the kernel's column-table region (0: none).
template <int N, int BUDGET>
__device__ __attribute__((noinline, not_tail_called)) void class_group(uint64_t task, uint64_t args, uint32_t g, uint32_t twSlot,
                                                      uint32_t heldSlot, uint32_t colSlot)
{
   extern __shared__ float2 smem_class[];
   ClassTaskK *tk = (ClassTaskK *) wave_uniform(task);
   ClassArgsK *c  = (ClassArgsK *) wave_uniform(args);
   g              = wave_uniform(g);
   twSlot         = wave_uniform(twSlot);
   heldSlot       = wave_uniform(heldSlot);
   colSlot        = wave_uniform(colSlot);
   FixedArgs a    = class_fixed_args(*c, *tk);
   const float2 *colLds = nullptr;
   if (colSlot != 0u && a.layout == LAYOUT_COLUMNS) {
      const uint32_t L = 1u << a.colBits, nh = (a.fullN + L - 1u) >> a.colBits;
      if (L + nh <= kColTwSlots) {
         const float2 **held = (const float2 **) (smem_class + colSlot + kColTwSlots);
         float2        *ct   = smem_class + colSlot;
         if ((const float2 *) wave_uniform((uint64_t) *held) != a.colTw) {   // staged once per task
            for (uint32_t i = threadIdx.x; i < L; i += Fixed<N>::P.threads) ct[i] = a.colTw[i];
            for (uint32_t i = threadIdx.x; i < nh; i += Fixed<N>::P.threads) ct[L + i] = a.colTwHi[i];
            __syncthreads();
            if (threadIdx.x == 0) *held = a.colTw;   // read again only after this group's closing barrier
         }
         colLds = ct;
      }
   }
   if constexpr (class_tw_lds<N>()) {
      const float2 **held = (const float2 **) (smem_class + heldSlot);
      float2        *twl  = smem_class + twSlot;
      if ((const float2 *) wave_uniform((uint64_t) *held) != a.tw) {   // one LDS word read by every work item
         for (uint32_t i = threadIdx.x; i < (uint32_t) Fixed<N>::P.twCount; i += Fixed<N>::P.threads) twl[i] = a.tw[i];
         __syncthreads();
         if (threadIdx.x == 0) *held = a.tw;   // read again only after this group's closing barrier
      }
      a.tw = twl;
   }
   fixed_group<N, KIND_C2C, BUDGET>(a, g, smem_class, colLds);
}

// Work items per workgroup of a class kernel: the same for all its lengths
// (an octave shares T and FPB).
template <int L0, int... Ls>
constexpr int class_threads()
{
   static_assert(((Fixed<Ls>::P.threads == Fixed<L0>::P.threads) && ...), "a class has one workgroup size");
   return Fixed<L0>::P.threads;
}

// Waves per SIMD the class's LDS allows (threads / 256 waves of each
// resident workgroup per SIMD, 160 KB of LDS per CU), at least one
// workgroup's and at most 4: the register budget of the kernel and of every
// per-length function it calls follows from it (512 / waves VGPRs).
template <int... Ls>
constexpr int class_waves()
{
   const int w = lds_waves_per_simd(class_lds_bytes<Ls...>(), class_threads<Ls...>());
   const int m = class_threads<Ls...>() / 256;
   return w < m ? m : w > 4 ? (m > 4 ? m : 4) : w;
}

template <int... Ls>
__global__ void __launch_bounds__(class_threads<Ls...>()) __attribute__((amdgpu_waves_per_eu(class_waves<Ls...>())))
fixed_class_kernel(ClassArgs c)
{
   extern __shared__ float2 smem_class[];
   constexpr uint32_t twSlot   = class_data_slots<Ls...>();
   constexpr uint32_t heldSlot = twSlot + class_tw_slots<Ls...>();
   constexpr uint32_t colSlot  = class_coltw_lds<Ls...>() ? heldSlot + 1u : 0u;
   constexpr int      kBudget  = pass0_register_points(class_waves<Ls...>());
   if (threadIdx.x == 0) {
      *(const float2 **) (smem_class + heldSlot) = nullptr;
      if constexpr (colSlot != 0u) *(const float2 **) (smem_class + colSlot + kColTwSlots) = nullptr;
   }
   __syncthreads();
   // Items in XCD-contiguous order: workgroups are dispatched round robin over
   // the kXcds dies (workgroup w on die w % kXcds), so the items are dealt in
   // kXcds contiguous runs, one per die. Adjacent groups share the cache lines
   // at their edges when rows or columns are not line aligned; this way they
   // meet in the same L2.
   uint32_t first = blockIdx.x, end = c.numItems, step = gridDim.x;
   if (gridDim.x % kXcds == 0u && c.numItems >= gridDim.x) {
      const uint32_t per   = (c.numItems + kXcds - 1u) / kXcds;
      const uint32_t begin = min((blockIdx.x % kXcds) * per, c.numItems);
      end                  = min(begin + per, c.numItems);
      first                = begin + blockIdx.x / kXcds;
      step                 = gridDim.x / kXcds;
   }
   // c is the kernel's only argument: at offset 0 of the kernarg segment.
   const uint64_t args = (uint64_t) __builtin_amdgcn_kernarg_segment_ptr();
   for (uint32_t w = first; w < end; w += step) {
      uint32_t g    = w;
      uint64_t task = args + offsetof(ClassArgs, single);
      uint32_t n    = c.single.n;
      if (c.tasks) {
         const uint32_t item = c.items[w];
         task                = (uint64_t) (c.tasks + (item >> kClassGroupBits));
         g                   = item & ((1u << kClassGroupBits) - 1u);
         n                   = c.tasks[item >> kClassGroupBits].n;
      }
      (void) ((n == (uint32_t) Ls ? (class_group<Ls, kBudget>(task, args, g, twSlot, heldSlot, colSlot), true) : false) ||
              ...);
   }
}

template <int... Ls>
uint32_t class_resident_groups()
{
   static uint32_t cached = 0;
   if (cached == 0) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
         cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fixed_class_kernel<Ls...>, class_threads<Ls...>(),
                                                       class_lds_bytes<Ls...>()) != hipSuccess ||
          per <= 0)
         per = 1;
      cached = (uint32_t) cus * (uint32_t) per;
   }
   return cached;
}

template <int... Ls>
hipError_t launch_class(const ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   if (c.numItems == 0) return hipSuccess;
   constexpr size_t lds    = class_lds_bytes<Ls...>();
   static bool      ldsSet = false;
   if (lds > 65536 && !ldsSet) {
      const hipError_t e = hipFuncSetAttribute((const void *) fixed_class_kernel<Ls...>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      if (e != hipSuccess) return e;
      ldsSet = true;
   }
   // Grid: up to 8 x the resident workgroups, so that large launches run one
   // item (or a few) per workgroup and the dispatcher refills a CU as soon as
   // one of its workgroups ends (2^25 points on MI355X against the resident
   // grid: class 256 151.6 -> 146.5 us, octave 2160..4096 at N = 4096 125.8 ->
   // 117.1 us, at 4000 147.6 -> 139.8 us). A nonzero `grid` (the ragged
   // autotuner's choice) replaces it; never more workgroups than items.
   uint32_t target = (uint32_t) std::min<uint64_t>(c.numItems, 8ull * class_resident_groups<Ls...>());
   if (grid != 0) target = min(c.numItems, grid);
   hipLaunchKernelGGL((fixed_class_kernel<Ls...>), dim3(target), dim3(class_threads<Ls...>()), lds, stream, c);
   return hipGetLastError();
}

// ---- persistent ragged launch ------------------------------------------------
// One launch runs every transform of a ragged direction: the lengths of all
// the kernel's classes (one workgroup size, LDS of the largest) and a work
// queue of units instead of a fixed item-to-workgroup deal. A unit is one
// group of a plain transform task, or a whole long transform: all groups of
// its column pass, a workgroup barrier, all groups of its row pass. The row
// pass reads only what the same workgroup's column pass wrote, so the
// dependency between the passes is inside the workgroup: the barrier orders
// the column stores before the row loads (one CU, one L1), with no
// cross-workgroup waits, counters or device-scope fences.
// counters: [kRqQueue] next unit, [kRqExit] workgroups done; the last
// workgroup out clears both for the next launch of the same plan (a graph
// replays it with no memset node).
enum : uint32_t { kRqQueue = 0u, kRqExit = 1u, kRqCounters = 2u };

// The column-pass twiddle tables of the long transforms held in LDS (staged
// once per task as in the octave kernels), even though the extra 4 KB leaves
// room for one workgroup fewer per CU than the class's data alone (ragged
// 10000 on MI355X: 1.903 ms per step with the tables read from L2 at 4
// workgroups per CU, 1.827 ms with them in LDS at 3).
constexpr bool ragged_coltw_lds() { return true; }

template <int... Ls>
__global__ void __launch_bounds__(class_threads<Ls...>()) __attribute__((amdgpu_waves_per_eu(class_waves<Ls...>())))
ragged_class_kernel(ClassArgs c)
{
   extern __shared__ float2 smem_class[];
   constexpr uint32_t twSlot   = class_data_slots<Ls...>();
   constexpr uint32_t heldSlot = twSlot + class_tw_slots<Ls...>();
   constexpr uint32_t colSlot  = ragged_coltw_lds() ? heldSlot + 1u : 0u;
   // queue entries (two slots) and the last-workgroup flag after the tables
   uint32_t *sItem = (uint32_t *) (smem_class + heldSlot + 1u + (colSlot != 0u ? kColTwSlots + 1u : 0u));
   uint32_t *sLast = sItem + 2;
   constexpr int      kBudget  = pass0_register_points(class_waves<Ls...>());
   uint32_t          *ctr      = c.counters;
   if (threadIdx.x == 0) {
      *(const float2 **) (smem_class + heldSlot) = nullptr;
      if constexpr (colSlot != 0u) *(const float2 **) (smem_class + colSlot + kColTwSlots) = nullptr;
   }
   const uint64_t args = (uint64_t) __builtin_amdgcn_kernarg_segment_ptr();
   auto           run  = [&](uint32_t ti, uint32_t g) __attribute__((always_inline)) {
      const uint64_t task = (uint64_t) (c.tasks + ti);
      const uint32_t n    = c.tasks[ti].n;
      (void) ((n == (uint32_t) Ls ? (class_group<Ls, kBudget>(task, args, g, twSlot, heldSlot, colSlot), true) : false) ||
              ...);
   };
   for (uint32_t it = 0;; ++it) {
      // two slots: thread 0 writes slot it & 1 while no work item can still
      // be reading it (they all passed the previous unit's barriers)
      if (threadIdx.x == 0)
         sItem[it & 1u] = __hip_atomic_fetch_add(ctr + kRqQueue, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      const uint32_t w = sItem[it & 1u];
      if (w >= c.numItems) break;
      const RaggedUnit u = c.units[w];
      for (uint32_t g = u.g0; g < u.g1; ++g) run(u.task, g);
      if (u.task2 != kNoTask) {
         __syncthreads();   // the column pass's stores before the row pass's loads
         for (uint32_t h = u.h0; h < u.h1; ++h) run(u.task2, h);
      }
   }
   // the last workgroup out (all others have taken their last unit) clears
   // the queue head and the exit count
   if (threadIdx.x == 0)
      *sLast = __hip_atomic_fetch_add(ctr + kRqExit, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1u;
   __syncthreads();
   if (*sLast && threadIdx.x < kRqCounters) __hip_atomic_store(ctr + threadIdx.x, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The class's data, twiddle and (ragged_coltw_lds) column-table slots, plus
// 4 words: two queue-entry slots and the last-workgroup flag.
template <int... Ls>
constexpr size_t ragged_lds_bytes()
{
   return class_lds_base_bytes<Ls...>() + (ragged_coltw_lds() ? (kColTwSlots + 1u) * sizeof(float2) : 0u) +
          4u * sizeof(uint32_t);
}

template <int... Ls>
uint32_t ragged_resident_groups()
{
   static uint32_t cached = 0;
   if (cached == 0) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
         cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, ragged_class_kernel<Ls...>, class_threads<Ls...>(),
                                                       ragged_lds_bytes<Ls...>()) != hipSuccess ||
          per <= 0)
         per = 1;
      cached = (uint32_t) cus * (uint32_t) per;
   }
   return cached;
}

// grid 0: the resident workgroups (a larger grid only adds workgroups that
// find the queue empty).
template <int... Ls>
hipError_t launch_ragged_class(const ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   if (c.numItems == 0) return hipSuccess;
   if (c.counters == nullptr || c.units == nullptr || c.tasks == nullptr) return hipErrorInvalidValue;
   constexpr size_t lds    = ragged_lds_bytes<Ls...>();
   static bool      ldsSet = false;
   if (lds > 65536 && !ldsSet) {
      const hipError_t e = hipFuncSetAttribute((const void *) ragged_class_kernel<Ls...>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      if (e != hipSuccess) return e;
      ldsSet = true;
   }
   uint32_t target = grid != 0 ? grid : ragged_resident_groups<Ls...>();
   target          = min(target, c.numItems);
   hipLaunchKernelGGL((ragged_class_kernel<Ls...>), dim3(target), dim3(class_threads<Ls...>()), lds, stream, c);
   return hipGetLastError();
}

// Workgroups resident at once on the device (queried once per kernel), used
// to size the persistent grid.
template <int N, int KIND>
uint32_t fixed_resident_groups(size_t lds)
{
   static uint32_t cached = 0;
   if (cached == 0) {
      int dev = 0, cus = 0, per = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
         cus = 256;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fixed_kernel<N, KIND>, Fixed<N>::P.threads, lds) !=
              hipSuccess ||
          per <= 0)
         per = 1;
      cached = (uint32_t) cus * (uint32_t) per;
   }
   return cached;
}

// Twice the groups per workgroup a large launch aims for; 0 = the resident
// grid (every workgroup loops over about numGroups / resident groups). Per
// kernel, from a sweep on MI355X at 503 MB per call (resident grid, 16, 8,
// 6, 4, 3, 2, 1.5 and 1 groups per workgroup; kept where the gain exceeded
// 2.5%): these kernels run 3-11% faster with more, shorter-lived workgroups
// (C2C 960: 185.5 -> 170.3 us at 2 groups per workgroup, 1024: 183.9 ->
// 168.3 us at 3), the others are fastest with the resident grid (256, 4096
// and longer, the lengths with radix-5 stages beyond the first pass, the
// real kernels of 1024-8192).
template <int N, int KIND>
constexpr uint32_t fixed_half_groups_per_wg()
{
   if constexpr (KIND == KIND_C2C) {
      switch (N) {
      case 480: case 512: return 3;
      case 960: return 4;
      case 1024: case 1536: case 1920: return 6;
      case 2048: return 8;
      default: return 0;
      }
   }
   else if constexpr (KIND == KIND_R2C) {
      return N == 32 ? 2u : N == 480 ? 4u : N == 512 ? 6u : N == 1024 ? 8u : 0u;
   }
   else {
      switch (N) {
      case 256: case 480: case 512: return 3;
      default: return 0;
      }
   }
}

template <int N, int KIND>
hipError_t launch_fixed(FixedArgs a, hipStream_t stream, uint32_t maxGrid = 0)
{
   constexpr FixedPlan P   = Fixed<N>::P;
   if (a.inOff || a.outOff) return hipErrorInvalidValue;   // strided channels only
   const size_t        lds = fixed_lds_bytes<N>();
   static bool         ldsSet = false;   // set once, so launches can be captured in a graph
   if (lds > 65536 && !ldsSet) {
      const hipError_t e = hipFuncSetAttribute((const void *) fixed_kernel<N, KIND>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
      if (e != hipSuccess) return e;
      ldsSet = true;
   }
   a.numGroups             = (a.numChannels + P.FPB - 1) / P.FPB;
   const uint32_t resident = fixed_resident_groups<N, KIND>(lds);
   uint32_t       grid     = min(a.numGroups, resident);
   constexpr uint32_t q2   = fixed_half_groups_per_wg<N, KIND>();
   if constexpr (q2 != 0) {
      const uint64_t aim = (2ull * a.numGroups + q2 - 1u) / q2;
      grid               = (uint32_t) std::min<uint64_t>(a.numGroups, std::max<uint64_t>(resident, aim));
   }
   if (maxGrid != 0 && grid > maxGrid) grid = maxGrid;
   hipLaunchKernelGGL((fixed_kernel<N, KIND>), dim3(grid), dim3(Fixed<N>::P.threads), lds, stream, a);
   return hipGetLastError();
}

}  // namespace fixedk
}  // namespace fftlib
