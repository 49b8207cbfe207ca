// Length-specialised gfx950 kernels for the fp32 transforms of the hot path:
// batched complex FFT / IFFT, and the fused real-to-complex and
// complex-to-real transforms whose inner complex FFT has a specialised length.
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
// N/16, at most 256) and 256/T transforms ("a group") per workgroup. A
// register pass of radices (r1, r2) is executed by each work item for
// ceil((N/(r1 r2))/T) groups of r1*r2 points; between passes the data stays
// in LDS (one padding slot per 32 points, which keeps the pass and gather
// access patterns of these lengths within 2-3 ways of the 64 LDS banks). The
// twiddle table is staged in LDS once per workgroup when it fits.
//
// The grid is persistent (as many workgroups as fit on the device; each
// takes groups g, g + gridDim.x, ...). When the first pass reads HBM in
// coalesced rows, the reads of the next group are issued into registers
// before the current group's LDS passes and stores, so HBM latency overlaps
// the LDS work. The result is read from LDS in digit-reversed order and
// written in natural order with coalesced stores: HBM sees one read and one
// write per point.
//
// Arithmetic and twiddle use are those of the generic kernels (reference
// src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/..._cn.c:240-491): the
// same butterflies, slot order and twiddle-table entries; the real-transform
// split steps follow src/fft_c7x/FFTLIB_fft1dBatched_i32f_c32fc_o32fc/
// ..._cn.c:40-121 and src/fft_c7x/FFTLIB_ifft1dBatched_i32fc_c32fc_o32f/
// ..._cn.c:35-87.
#include <hip/hip_runtime.h>

#include <math.h>

#include <map>
#include <mutex>
#include <vector>

#include "fft_fixed_impl.h"
#include "fft_fixed_class.h"
#include "fftlib_gpu.h"

namespace fftlib {
using namespace fixedk;
namespace {

// Specialised lengths: C2C list first; the real transforms use the entries
// marked for their inner complex length.
#define FFTLIB_FIXED_C2C_LENGTHS(X)                                                                                 \
   X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192) X(16384) X(48) X(96) X(192) X(240)         \
   X(480) X(960) X(1920) X(3840) X(1200) X(1536) X(1000) X(2000) X(4000) X(3200) X(6400)
#define FFTLIB_FIXED_REAL_LENGTHS(X) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192) X(480) X(960)

template <int KIND>
hipError_t dispatch_fixed_real(uint32_t n, const FixedArgs &a, hipStream_t stream)
{
   switch (n) {
#define FFTLIB_FIXED_CASE(L)                                                                                        \
   case L: return launch_fixed<L, KIND>(a, stream);
      FFTLIB_FIXED_REAL_LENGTHS(FFTLIB_FIXED_CASE)
#undef FFTLIB_FIXED_CASE
   default: return hipErrorInvalidValue;
   }
}

}  // namespace

namespace {

bool smooth235(uint32_t n)
{
   for (uint32_t p : {2u, 3u, 5u})
      while (n % p == 0u) n /= p;
   return n == 1u;
}

bool main_c2c_has(uint32_t n)
{
   switch (n) {
#define FFTLIB_FIXED_CASE(L) case L:
      FFTLIB_FIXED_C2C_LENGTHS(FFTLIB_FIXED_CASE)
#undef FFTLIB_FIXED_CASE
      return true;
   default: return false;
   }
}

// Work items per transform and transforms per workgroup of the specialised
// kernels (make_fixed_plan's rule).
uint32_t fixed_fpb(uint32_t n)
{
   uint32_t T = 1;
   while (T < (uint32_t) kMaxFixedThreads && T * 16u < n) T *= 2u;
   return T >= (uint32_t) kFixedThreads ? 1u : (uint32_t) kFixedThreads / T;
}

// One task of a class kernel covering `count` channels of length n.
ClassTask single_task(uint32_t n, uint32_t count, const float2 *tw, const uint64_t *inOff, const uint64_t *outOff)
{
   ClassTask k{};
   k.n = n, k.numChannels = count, k.tw = tw, k.inOff = inOff, k.outOff = outOff;
   k.numGroups = (count + fixed_fpb(n) - 1u) / fixed_fpb(n);
   return k;
}

}  // namespace

int fixed_class_of(uint32_t n)
{
   if (n < 16u || n > 16384u || !smooth235(n)) return -1;
   int k = 0;
   while ((16u << k) < n) ++k;
   return k;
}

hipError_t fixed_class_launch(int cls, const ClassArgs &c, uint32_t grid, hipStream_t stream)
{
   switch (cls) {
   case 0: return fixed_class_launch_tu1(cls, c, grid, stream);
   case 1: return fixed_class_launch_tu0(cls, c, grid, stream);
   case 2: return fixed_class_launch_tu2(cls, c, grid, stream);
   case 3: return fixed_class_launch_tu3(cls, c, grid, stream);
   case 4: return fixed_class_launch_tu1(cls, c, grid, stream);
   case 5: return fixed_class_launch_tu2(cls, c, grid, stream);
   case 6: return fixed_class_launch_tu3(cls, c, grid, stream);
   case 7: return fixed_class_launch_tu3(cls, c, grid, stream);
   case 8: return fixed_class_launch_tu2(cls, c, grid, stream);
   case 9: return fixed_class_launch_tu1(cls, c, grid, stream);
   case 10: return fixed_class_launch_tu0(cls, c, grid, stream);
   default: return hipErrorInvalidValue;
   }
}

namespace {

// Complex FFT of length n on its specialised kernel: the dedicated one for
// the lengths listed above, else its class kernel with one task.
hipError_t dispatch_c2c_any(uint32_t n, const FixedArgs &a, hipStream_t stream)
{
   if (main_c2c_has(n)) {
      switch (n) {
#define FFTLIB_FIXED_CASE(L)                                                                                        \
   case L: return launch_fixed<L, KIND_C2C>(a, stream);
         FFTLIB_FIXED_C2C_LENGTHS(FFTLIB_FIXED_CASE)
#undef FFTLIB_FIXED_CASE
      default: break;
      }
   }
   const int cls = fixed_class_of(n);
   if (cls < 0) return hipErrorInvalidValue;
   ClassArgs c{};
   c.x = a.x, c.y = a.y, c.scratch = nullptr, c.mode = a.mode;
   c.single   = single_task(n, a.numChannels, a.tw, a.inOff, a.outOff);
   c.numItems = c.single.numGroups;
   return fixed_class_launch(cls, c, 0, stream);
}

}  // namespace

bool fixed_c2c_supported(uint32_t n) { return main_c2c_has(n) || fixed_class_of(n) >= 0; }

bool fixed_real_supported(uint32_t nc)
{
   switch (nc) {
#define FFTLIB_FIXED_CASE(L) case L:
      FFTLIB_FIXED_REAL_LENGTHS(FFTLIB_FIXED_CASE)
#undef FFTLIB_FIXED_CASE
      return true;
   default: return false;
   }
}

hipError_t launch_c2c_fixed(const float *x, float *y, const float *tw, uint32_t n, uint32_t numChannels, int mode,
                            hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   FixedArgs a{};
   a.x = (const float2 *) x, a.y = (float2 *) y, a.tw = (const float2 *) tw, a.sf = nullptr;
   a.numChannels = numChannels, a.mode = mode, a.aux = 0, a.inStride = n, a.outStride = n;
   return dispatch_c2c_any(n, a, stream);
}

hipError_t launch_r2c_fixed(const float *x, float *y, const float *tw, const float *sf, uint32_t nc,
                            uint32_t numChannels, uint32_t outmode, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   FixedArgs a{};
   a.x = (const float2 *) x, a.y = (float2 *) y, a.tw = (const float2 *) tw, a.sf = sf;
   a.sf2 = ((uintptr_t) sf & 7u) == 0 ? (const float2 *) sf : nullptr;
   a.numChannels = numChannels, a.mode = C2C_FORWARD, a.aux = outmode, a.inStride = nc;
   a.outStride = outmode ? 2u * nc : nc + 1u;
   return dispatch_fixed_real<KIND_R2C>(nc, a, stream);
}

hipError_t launch_c2r_fixed(const float *x, float *y, const float *tw, const float *sf, uint32_t nc,
                            uint32_t numChannels, uint32_t inMode, int postdiv, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   FixedArgs a{};
   a.x = (const float2 *) x, a.y = (float2 *) y, a.tw = (const float2 *) tw, a.sf = sf;
   a.sf2 = ((uintptr_t) sf & 7u) == 0 ? (const float2 *) sf : nullptr;
   a.numChannels = numChannels, a.mode = postdiv ? C2C_INVERSE_POSTDIV : C2C_INVERSE_PRESCALE, a.aux = inMode;
   a.inStride = inMode ? 2u * nc : nc + 1u, a.outStride = nc;
   return dispatch_fixed_real<KIND_C2R>(nc, a, stream);
}

// ---- two-pass transforms longer than 16384 points -------------------------------
namespace {

constexpr uint32_t kTwoPassMaxN = 1u << 22;   // column twiddle table: 8 B per point

struct TwoPassKey {
   int      dev;
   uint32_t n;
   bool     operator<(const TwoPassKey &o) const { return dev != o.dev ? dev < o.dev : n < o.n; }
};
std::mutex                          g_twoPassMu;
std::map<TwoPassKey, TwoPassPlan>   g_twoPass;   // device tables live until the process ends

}  // namespace

// Split n = n1 * m2 (m2 = columns of the column pass, n1 = their length).
// The column pass moves runs of FPB(n1) adjacent columns at offsets
// p + m2*i, the row pass runs of FPB(m2) adjacent rows at k1 + n1*k2; a run
// is a whole number of 128-byte lines when both factors are at most 256
// (16 transforms per workgroup) and its start is line aligned when the
// stride is a multiple of 16 points. Measured on MI355X (2^25 points, forward
// + inverse): 18432 = 144 x 128 runs at 1.44 TB/s, 18750 = 150 x 125 at
// 1.16. The split maximises: both factors <= 256 (8), m2 a multiple of 16
// (4) or 8 (2), n1 a multiple of 16 (2) or 8 (1); then the most balanced.
bool two_pass_split(uint32_t n, uint32_t *n1, uint32_t *m2) { return two_pass_split_above(n, 16384u, n1, m2); }

bool two_pass_split_above(uint32_t n, uint32_t above, uint32_t *n1, uint32_t *m2)
{
   if (n <= above || n < 256u || n > kTwoPassMaxN || !smooth235(n)) return false;
   int      a = 0, b = 0, c = 0;
   uint32_t t = n;
   while (t % 2u == 0u) t /= 2u, ++a;
   while (t % 3u == 0u) t /= 3u, ++b;
   while (t % 5u == 0u) t /= 5u, ++c;
   int      bestScore = -1;
   double   bestBal   = 0.0;
   uint32_t best      = 0;
   for (int i = 0; i <= a; ++i) {
      for (int j = 0; j <= b; ++j) {
         for (int k = 0; k <= c; ++k) {
            uint64_t d = 1u << i;
            for (int r = 0; r < j; ++r) d *= 3u;
            for (int r = 0; r < k; ++r) d *= 5u;
            const uint64_t e = n / d;
            if (d < 16u || d > 16384u || e < 16u || e > 16384u) continue;
            // N1 = e a multiple of 16 first (row-pass outputs on whole
            // lines, see rowAlign), then M2 = d (column reads and the
            // intermediate rows without idle columns)
            const int score = ((d <= 256u && e <= 256u) ? 8 : 0) + (e % 16u == 0u ? 4 : e % 8u == 0u ? 2 : 0) +
                              (d % 16u == 0u ? 2 : d % 8u == 0u ? 1 : 0);
            const double bal = fabs(log((double) d) - log((double) e));
            if (score > bestScore || (score == bestScore && bal < bestBal)) {
               bestScore = score, bestBal = bal, best = (uint32_t) d;
            }
         }
      }
   }
   if (best == 0) return false;
   *m2 = best;
   *n1 = n / best;
   return true;
}

size_t two_pass_scratch_points(uint32_t n, uint32_t count)
{
   uint32_t n1 = 0, m2 = 0;
   if (!two_pass_split(n, &n1, &m2)) return 0;
   return (size_t) n1 * ((m2 + 15u) & ~15u) * count;
}

hipError_t two_pass_plan(uint32_t n, TwoPassPlan *plan)
{
   uint32_t n1 = 0, m2 = 0;
   if (!two_pass_split_above(n, kRaggedClassMaxN, &n1, &m2)) return hipErrorInvalidValue;
   int        dev = 0;
   hipError_t e   = hipGetDevice(&dev);
   if (e != hipSuccess) return e;
   std::lock_guard<std::mutex> lock(g_twoPassMu);
   auto it = g_twoPass.find(TwoPassKey{dev, n});
   if (it != g_twoPass.end()) {
      *plan = it->second;
      return hipSuccess;
   }
   // [table of n1 | table of m2 | W_n^e for e < L | W_n^(h L) for h < ceil(n / L)],
   // L = 2^bits ~ sqrt(n): the column twiddle W_n^e is the product of one
   // entry of each of the last two tables (a few KB per length, so the
   // tables of many lengths stay in L2), complex fp32 from double
   uint32_t bits = 0;
   while ((1ull << (2u * bits)) < n) ++bits;
   const uint32_t     L  = 1u << bits;
   const uint32_t     nh = (n + L - 1u) / L;
   std::vector<float> host(2u * ((size_t) n1 + m2 + L + nh));
   FFTLIB_tw_gen_c2c(host.data(), (int32_t) n1);
   FFTLIB_tw_gen_c2c(host.data() + 2u * n1, (int32_t) m2);
   float *col = host.data() + 2u * ((size_t) n1 + m2);
   for (uint32_t k = 0; k < L; ++k) {
      const double a = 2.0 * M_PI * (double) k / (double) n;
      col[2u * k]      = (float) cos(a);
      col[2u * k + 1u] = (float) -sin(a);
   }
   float *colHi = col + 2u * L;
   for (uint32_t h = 0; h < nh; ++h) {
      const double a = 2.0 * M_PI * (double) ((uint64_t) h * L) / (double) n;
      colHi[2u * h]      = (float) cos(a);
      colHi[2u * h + 1u] = (float) -sin(a);
   }
   float *d = nullptr;
   if ((e = hipMalloc((void **) &d, host.size() * sizeof(float))) != hipSuccess) return e;
   if ((e = hipMemcpy(d, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess) {
      (void) hipFree(d);
      return e;
   }
   TwoPassPlan p;
   p.n = n, p.n1 = n1, p.m2 = m2, p.s2 = (m2 + 15u) & ~15u;
   p.tw1 = d, p.tw2 = d + 2u * n1, p.colTw = d + 2u * ((size_t) n1 + m2);
   p.colTwHi = p.colTw + 2u * L, p.colBits = bits;
   g_twoPass[TwoPassKey{dev, n}] = p;
   *plan = p;
   return hipSuccess;
}

ClassTask two_pass_task(const TwoPassPlan &p, int layout, uint32_t count, const uint64_t *off, uint64_t scratchOff)
{
   ClassTask k{};
   const bool cols = layout == (int) LAYOUT_COLUMNS;
   k.n           = cols ? p.n1 : p.m2;
   k.numChannels = count * (cols ? p.s2 : p.n1);   // columns m2..s2-1 of each transform idle
   k.numGroups   = (k.numChannels + fixed_fpb(k.n) - 1u) / fixed_fpb(k.n);
   k.layout      = (uint32_t) layout;
   k.tw          = (const float2 *) (cols ? p.tw1 : p.tw2);
   k.inOff       = cols ? off : nullptr;
   k.outOff      = cols ? nullptr : off;
   k.scratchOff  = scratchOff;
   k.span        = cols ? p.m2 : p.n1;
   k.scrStride   = p.s2;
   k.fullN       = p.n;
   k.colTw       = (const float2 *) p.colTw;
   k.colTwHi     = (const float2 *) p.colTwHi;
   k.colBits     = p.colBits;
   // rows rotated so that the row groups' output runs fill 128-byte lines
   // (A = 16 points), else 64-byte halves (A = 8): needs N1 a multiple of A
   // (every output row k2 starts at the same offset within a line, and each
   // A-row block of a transform's rows lies in one row group) and row groups
   // of a multiple of A rows
   const uint32_t rowFpb = fixed_fpb(p.m2);
   k.rowAlign    = (p.n1 % 16u == 0u && rowFpb % 16u == 0u) ? 16u : (p.n1 % 8u == 0u && rowFpb % 8u == 0u) ? 8u : 0u;
   return k;
}

uint32_t class_group_size(uint32_t n) { return fixed_fpb(n); }

// ---- two-pass tables in caller memory (no allocation) ---------------------------
namespace {

// One block of the reference twiddle layout (FFTLIB_tw_gen_c2c's loops): the
// radix-r stage with stride j of the radix group that starts at length ng,
// t = ng / (r j) butterflies; entry (i', m) of the block, i' < t, m = 1..r-1,
// is W^(m i' j) over ng at complex position k0 + (m-1) t + i'.
struct TwBlock {
   uint32_t r, ng, j, t, k0, first;   // first: index of the block's first thread
};
constexpr int kMaxTwBlocks = 48;       // two tables of at most 2 + 2 + 11 stages each
struct TwGenArgs {
   float2  *dst;
   uint32_t nBlocks;
   uint32_t twEntries;                 // threads of the stage tables (both lengths)
   uint32_t n, L, nh, colOff;          // column tables: W_n^e (e < L) at colOff, W_n^(hL) (h < nh) after
   TwBlock  b[kMaxTwBlocks];
};

// Blocks of length n's stage table at complex offset base; same loop bounds
// and index arithmetic as FFTLIB_tw_gen_c2c.
bool tw_blocks(int32_t numPoints, uint32_t base, TwGenArgs &a)
{
   const int32_t rads[3] = {5, 3, 4};
   int32_t       k = 0, ng = numPoints;
   for (int32_t g = 0; g < 3; ++g) {
      const int32_t r = rads[g];
      if (ng <= 0 || ng % r != 0) continue;
      int32_t len = ng, t = ng / r;
      for (int32_t j = 1; j <= ng / r && len % r == 0; j *= r) {
         if (a.nBlocks >= (uint32_t) kMaxTwBlocks) return false;
         a.b[a.nBlocks++] = TwBlock{(uint32_t) r, (uint32_t) ng, (uint32_t) j, (uint32_t) t, base + (uint32_t) k / 2u,
                                    a.twEntries};
         a.twEntries += (uint32_t) t * (uint32_t) (r - 1);
         k += 2 * t * (r - 1);
         t /= r;
         len /= r;
      }
      if (r != 4) ng = len;
   }
   return true;
}

// One work item per table entry: the stage tables of both lengths in the
// host generator's arithmetic (PI = 3.141592654, angle 2 m pi i / ng in
// double), then the column tables (angle 2 pi e / n with the double-precision
// pi, as two_pass_plan builds them).
__global__ void __launch_bounds__(256) two_pass_tables_kernel(TwGenArgs a)
{
   const uint32_t e = blockIdx.x * 256u + threadIdx.x;
   if (e < a.twEntries) {
      uint32_t bi = 0;
      while (bi + 1u < a.nBlocks && a.b[bi + 1u].first <= e) ++bi;
      const TwBlock  b  = a.b[bi];
      const uint32_t q  = e - b.first;
      const uint32_t ip = q / (b.r - 1u), m = 1u + q % (b.r - 1u);
      const double   pi = 3.141592654;
      const double   an = 2.0 * (double) m * pi * (double) (ip * b.j) / (double) b.ng;
      a.dst[b.k0 + (m - 1u) * b.t + ip] = make_float2((float) cos(an), (float) -sin(an));
      return;
   }
   const uint32_t c = e - a.twEntries;
   if (c < a.L + a.nh) {
      const uint64_t ex = c < a.L ? (uint64_t) c : (uint64_t) (c - a.L) * a.L;
      const double   an = 2.0 * M_PI * (double) ex / (double) a.n;
      a.dst[a.colOff + c] = make_float2((float) cos(an), (float) -sin(an));
   }
}

void col_table_sizes(uint32_t n, uint32_t *bits, uint32_t *L, uint32_t *nh)
{
   uint32_t b = 0;
   while ((1ull << (2u * b)) < n) ++b;
   *bits = b;
   *L    = 1u << b;
   *nh   = (n + *L - 1u) / *L;
}

}  // namespace

size_t two_pass_table_bytes(uint32_t n)
{
   uint32_t n1 = 0, m2 = 0, bits, L, nh;
   if (!two_pass_split(n, &n1, &m2)) return 0;
   col_table_sizes(n, &bits, &L, &nh);
   return ((size_t) n1 + m2 + L + nh) * sizeof(float2);
}

hipError_t two_pass_plan_at(uint32_t n, float *dst, hipStream_t stream, TwoPassPlan *plan)
{
   uint32_t n1 = 0, m2 = 0, bits, L, nh;
   if (!two_pass_split(n, &n1, &m2)) return hipErrorInvalidValue;
   col_table_sizes(n, &bits, &L, &nh);
   TwGenArgs a{};
   a.dst = (float2 *) dst;
   if (!tw_blocks((int32_t) n1, 0u, a) || !tw_blocks((int32_t) m2, n1, a)) return hipErrorInvalidValue;
   a.n = n, a.L = L, a.nh = nh, a.colOff = n1 + m2;
   const uint32_t threads = a.twEntries + L + nh;
   hipLaunchKernelGGL(two_pass_tables_kernel, dim3((threads + 255u) / 256u), dim3(256), 0, stream, a);
   hipError_t e = hipGetLastError();
   if (e != hipSuccess) return e;
   TwoPassPlan p;
   p.n = n, p.n1 = n1, p.m2 = m2, p.s2 = (m2 + 15u) & ~15u;
   p.tw1 = dst, p.tw2 = dst + 2u * n1, p.colTw = dst + 2u * ((size_t) n1 + m2);
   p.colTwHi = p.colTw + 2u * L, p.colBits = bits;
   *plan = p;
   return hipSuccess;
}

hipError_t launch_c2c_two_pass(const float *x, float *y, uint32_t n, uint32_t numChannels, int mode, float *scratch,
                               hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   TwoPassPlan p;
   hipError_t  e = two_pass_plan(n, &p);
   if (e != hipSuccess) return e;
   return launch_c2c_two_pass_plan(p, x, y, numChannels, mode, scratch, stream);
}

hipError_t launch_c2c_two_pass_plan(const TwoPassPlan &p, const float *x, float *y, uint32_t numChannels, int mode,
                                    float *scratch, hipStream_t stream)
{
   if (numChannels == 0) return hipSuccess;
   hipError_t e;
   if ((uint64_t) numChannels * p.s2 >= (1ull << 32) || (uint64_t) numChannels * p.n1 >= (1ull << 32))
      return hipErrorInvalidValue;
   ClassArgs c{};
   c.x = (const float2 *) x, c.y = (float2 *) y, c.scratch = (float2 *) scratch, c.mode = mode;
   c.single   = two_pass_task(p, LAYOUT_COLUMNS, numChannels, nullptr, 0);
   c.numItems = c.single.numGroups;
   if ((e = fixed_class_launch(fixed_class_of(p.n1), c, 0, stream)) != hipSuccess) return e;
   c.single   = two_pass_task(p, LAYOUT_ROWS, numChannels, nullptr, 0);
   c.numItems = c.single.numGroups;
   return fixed_class_launch(fixed_class_of(p.m2), c, 0, stream);
}

}  // namespace fftlib
