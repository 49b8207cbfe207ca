// GPU runtime of the C-ABI layer, the exported GPU controls (fftlib_gpu.h),
// the twiddle/split-factor table generators and the stage-table builder.
#include "fftlib_runtime.h"
#include "fft_kernels.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

#include "fftlib_gpu.h"

namespace fftlib {
namespace {

std::atomic<int>         g_style{FFTLIB_FUNCTION_OPTIMIZED};
std::atomic<hipStream_t> g_stream{nullptr};
std::atomic<int>         g_sync{1};
std::atomic<int>         g_specialized{1};
std::atomic<int>         g_variant{0};
thread_local char        t_lastError[256] = "";

// Device staging for host-memory callers: one grow-only buffer per role,
// guarded by a mutex (host-pointer calls are serialized).
struct Staging {
   std::mutex mu;
   void      *buf[4]  = {nullptr, nullptr, nullptr, nullptr};
   size_t     cap[4]  = {0, 0, 0, 0};
};
Staging &staging()
{
   static Staging s;
   return s;
}

hipError_t ensure(int slot, size_t bytes)
{
   Staging &s = staging();
   if (bytes <= s.cap[slot]) return hipSuccess;
   if (s.buf[slot]) {
      hipError_t e = hipFree(s.buf[slot]);
      s.buf[slot]  = nullptr;
      s.cap[slot]  = 0;
      if (e != hipSuccess) return e;
   }
   size_t     want = (bytes + 4095) & ~(size_t) 4095;
   hipError_t e    = hipMalloc(&s.buf[slot], want);
   if (e == hipSuccess) s.cap[slot] = want;
   return e;
}

}  // namespace

FFTLIB_FUNCTION_STYLE current_style() { return (FFTLIB_FUNCTION_STYLE) g_style.load(); }
hipStream_t           current_stream() { return g_stream.load(); }
bool                  synchronous() { return g_sync.load() != 0; }
bool                  specialized_kernels() { return g_specialized.load() != 0; }
int                   c2c_variant() { return g_variant.load(); }

void set_last_error(const char *what, hipError_t err)
{
   snprintf(t_lastError, sizeof t_lastError, "%s: %s", what, hipGetErrorString(err));
}

void set_last_error_text(const char *text) { snprintf(t_lastError, sizeof t_lastError, "%s", text); }

hipError_t device_scratch(size_t bytes, hipStream_t stream, int role, void **ptr)
{
   struct Slot {
      void  *p   = nullptr;
      size_t cap = 0;
   };
   static std::mutex                                          mu;
   static std::map<std::tuple<int, hipStream_t, int>, Slot>   slots;
   int                                                        dev = 0;
   hipError_t                                                 e   = hipGetDevice(&dev);
   if (e != hipSuccess) return e;
   std::lock_guard<std::mutex> lock(mu);
   Slot &s = slots[std::make_tuple(dev, stream, role)];
   if (bytes > s.cap) {
      if (s.p) {
         // a kernel of an earlier call on this stream may still read it
         if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
         if ((e = hipFree(s.p)) != hipSuccess) return e;
         s.p = nullptr, s.cap = 0;
      }
      const size_t want = (bytes + 65535) & ~(size_t) 65535;
      if ((e = hipMalloc(&s.p, want)) != hipSuccess) return e;
      s.cap = want;
   }
   *ptr = s.p;
   return hipSuccess;
}

static FFTLIB_STATUS fail(const char *what, hipError_t err)
{
   set_last_error(what, err);
   return FFTLIB_ERR_FAILURE;
}

FFTLIB_STATUS run_on_gpu(const Plan &plan, const CallBuffers &b, LaunchFn launch)
{
   return run_buffers(plan.style, b, [&](const DeviceBuffers &d, hipStream_t s) { return launch(plan, d, s); });
}

FFTLIB_STATUS run_buffers(uint32_t style, const CallBuffers &b,
                          const std::function<hipError_t(const DeviceBuffers &, hipStream_t)> &launch)
{
   hipStream_t stream = current_stream();
   if (style == FFTLIB_FUNCTION_GPU) {
      DeviceBuffers d{b.in.ptr, b.tw.ptr, b.aux0.ptr, b.out};
      hipError_t    e = launch(d, stream);
      if (e != hipSuccess) return fail("kernel launch", e);
      if (synchronous()) {
         e = hipStreamSynchronize(stream);
         if (e != hipSuccess) return fail("hipStreamSynchronize", e);
      }
      return FFTLIB_SUCCESS;
   }
   // FFTLIB_FUNCTION_OPTIMIZED: stage through device memory. hipMemcpyDefault
   // lets unified addressing pick the direction, so device pointers work too.
   Staging                    &s = staging();
   std::lock_guard<std::mutex> lock(s.mu);
   const IoBuffer              ins[3] = {b.in, b.tw, b.aux0};
   hipError_t                  e;
   for (int i = 0; i < 3; ++i) {
      if (ins[i].ptr == nullptr || ins[i].bytes == 0) continue;
      if ((e = ensure(i, ins[i].bytes)) != hipSuccess) return fail("hipMalloc (staging)", e);
      if ((e = hipMemcpyAsync(s.buf[i], ins[i].ptr, ins[i].bytes, hipMemcpyDefault, stream)) != hipSuccess)
         return fail("hipMemcpyAsync (to device)", e);
   }
   if ((e = ensure(3, b.outBytes)) != hipSuccess) return fail("hipMalloc (staging)", e);
   DeviceBuffers d{b.in.ptr ? s.buf[0] : nullptr, b.tw.ptr ? s.buf[1] : nullptr, b.aux0.ptr ? s.buf[2] : nullptr,
                   s.buf[3]};
   if ((e = launch(d, stream)) != hipSuccess) return fail("kernel launch", e);
   if ((e = hipMemcpyAsync(b.out, s.buf[3], b.outBytes, hipMemcpyDefault, stream)) != hipSuccess)
      return fail("hipMemcpyAsync (to host)", e);
   if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail("hipStreamSynchronize", e);
   return FFTLIB_SUCCESS;
}

// Stage sequence of the reference's natural-C kernels: every radix-5 stage,
// then radix-3, then radix-4, then one radix-2 for an odd power of two
// (src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/..._cn.c:240-260); the
// twiddle block of a radix-r stage with sub-length M holds (r-1)*(M/r) complex
// values (..._cn.c:305, :335, :385). The output permutation uses the digits
// [5..., 3..., 2,2 per radix-4, 2] (..._cn.c:430-458) or bit reversal.
// Consecutive stages are then paired into register passes of at most 25
// points per work item.
bool build_stage_table(uint32_t n, StageTable *st)
{
   memset(st, 0, sizeof *st);
   if (n == 0) return false;
   st->n          = n;
   uint32_t t     = n, M = n, off = 0;
   auto     stage = [&](uint32_t r) {
      if (st->nStages >= (uint32_t) kMaxStages) return false;
      const uint32_t q          = M / r;
      st->radix[st->nStages]    = (uint8_t) r;
      st->twOff[st->nStages]    = off;
      st->stageM[st->nStages]   = M;
      st->nStages++;
      if (r != 2) off += (r - 1) * q;
      M = q;
      return true;
   };
   auto digit = [&](uint32_t r) {
      if (st->nDigits >= (uint32_t) kMaxStages) return false;
      st->digit[st->nDigits++] = (uint8_t) r;
      return true;
   };
   bool odd = false;
   while (t % 5u == 0u) {
      if (!stage(5) || !digit(5)) return false;
      t /= 5u;
      odd = true;
   }
   while (t % 3u == 0u) {
      if (!stage(3) || !digit(3)) return false;
      t /= 3u;
      odd = true;
   }
   while (t % 4u == 0u) {
      if (!stage(4) || !digit(2) || !digit(2)) return false;
      t /= 4u;
   }
   if (t == 2u) {
      if (!stage(2) || !digit(2)) return false;
      t = 1u;
   }
   if (t != 1u) return false;
   st->twCount = off;
   st->pow2    = odd ? 0u : 1u;
   if (st->pow2) {
      uint32_t l = 0;
      while ((1u << l) < n) ++l;
      st->log2n = l;
   }
   for (uint32_t s = 0; s < st->nStages;) {
      const bool pair = s + 1 < st->nStages && (uint32_t) st->radix[s] * st->radix[s + 1] <= 25u;
      st->passFirst[st->nPasses] = (uint8_t) s;
      st->passLen[st->nPasses]   = pair ? 2 : 1;
      st->nPasses++;
      s += pair ? 2u : 1u;
   }
   return true;
}

}  // namespace fftlib

using namespace fftlib;

extern "C" {

FFTLIB_STATUS FFTLIB_setFunctionStyle(FFTLIB_FUNCTION_STYLE style)
{
   if (style != FFTLIB_FUNCTION_NATC && style != FFTLIB_FUNCTION_OPTIMIZED && style != FFTLIB_FUNCTION_GPU)
      return FFTLIB_ERR_INVALID_VALUE;
   g_style.store((int) style);
   return FFTLIB_SUCCESS;
}

FFTLIB_FUNCTION_STYLE FFTLIB_getFunctionStyle(void) { return current_style(); }

FFTLIB_STATUS FFTLIB_GPU_setStream(void *hipStream)
{
   g_stream.store((hipStream_t) hipStream);
   return FFTLIB_SUCCESS;
}

void *FFTLIB_GPU_getStream(void) { return (void *) current_stream(); }

FFTLIB_STATUS FFTLIB_GPU_setSynchronous(int32_t sync)
{
   g_sync.store(sync != 0);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_setSpecializedKernels(int32_t enable)
{
   g_specialized.store(enable != 0);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_setC2CVariant(int32_t variant)
{
   if (variant < C2C_VARIANT_FP32 || variant > C2C_VARIANT_BF16X2) return FFTLIB_ERR_INVALID_VALUE;
   g_variant.store(variant);
   return FFTLIB_SUCCESS;
}

int32_t FFTLIB_GPU_getC2CVariant(void) { return g_variant.load(); }

FFTLIB_STATUS FFTLIB_GPU_synchronize(void)
{
   hipError_t e = hipStreamSynchronize(current_stream());
   if (e != hipSuccess) {
      set_last_error("hipStreamSynchronize", e);
      return FFTLIB_ERR_FAILURE;
   }
   return FFTLIB_SUCCESS;
}

const char *FFTLIB_GPU_getLastError(void) { return t_lastError; }

const char *FFTLIB_GPU_version(void) { return "fftlib-mi355x 0.1 (gfx950)"; }

// Angles are formed in double with PI = 3.141592654 (the reference's constant,
// not M_PI) as (2*m*PI*i)/Ng, Ng the length of the radix group, and cos/-sin
// rounded to float -- the exact expression order of the reference generator,
// so tables are bit-identical to it.
void FFTLIB_tw_gen_c2c(FFTLIB_F32 *pW, int32_t numPoints)
{
   const double   pi      = 3.141592654;
   const int32_t  rads[3] = {5, 3, 4};
   int32_t        k = 0, ng = numPoints;
   for (int32_t g = 0; g < 3; ++g) {
      const int32_t r = rads[g];
      if (ng <= 0 || ng % r != 0) continue;
      int32_t len = ng, t = ng / r;
      for (int32_t j = 1; j <= ng / r && len % r == 0; j *= r) {
         for (int32_t i = 0; i < ng / r; i += j) {
            for (int32_t m = 1; m < r; ++m) {
               const double a              = 2.0 * (double) m * pi * i / ng;
               pW[k + 2 * (m - 1) * t]     = (FFTLIB_F32) cos(a);
               pW[k + 2 * (m - 1) * t + 1] = (FFTLIB_F32) -sin(a);
            }
            k += 2;
         }
         k += 2 * t * (r - 2);
         t /= r;
         len /= r;
      }
      if (r != 4) ng = len;
   }
}

static int16_t q15_of(double v)
{
   const double x = floor(0.5 + v);
   if (x >= 32767.0) return 32767;
   if (x <= -32768.0) return -32768;
   return (int16_t) x;
}

void FFTLIB_tw_gen_i16sc(int16_t *pW, uint32_t numPoints)
{
   const double pi = 3.141592654;
   const int32_t q = (int32_t) (numPoints >> 2);
   int32_t       k = 0, t = q;
   for (int32_t j = 1; j <= q; j <<= 2) {
      for (int32_t i = 0; i < q; i += j) {
         for (int32_t m = 1; m < 4; ++m) {
            const double a              = 2.0 * (double) m * pi * i / numPoints;
            pW[k + 2 * (m - 1) * t]     = q15_of(32767.5 * cos(a));
            pW[k + 2 * (m - 1) * t + 1] = q15_of(32767.5 * (-sin(a)));
         }
         k += 2;
      }
      k += 4 * t;
      t >>= 2;
   }
}

// N/2-point tables and split factors of the real transforms
// (src/fft_c7x/FFTLIB_fft1dBatched_i32f_c32fc_o32fc/..._tw_gen.cpp:9-107,
// src/fft_c7x/FFTLIB_ifft1dBatched_i32fc_c32fc_o32f/..._tw_gen.cpp:9-95).
void tw_gen_r2c(FFTLIB_F32 *pW, int32_t numPoints) { FFTLIB_tw_gen_c2c(pW, numPoints >> 1); }
void ifft_c2r_tw_gen(FFTLIB_F32 *pW, int32_t numPoints) { FFTLIB_tw_gen_c2c(pW, numPoints >> 1); }

void split_factor_gen_r2c(FFTLIB_F32 *pSf, int32_t numPoints)
{
   const double pi = 3.141592654;
   for (int32_t i = 0; i < numPoints / 2; ++i) {
      const double a                 = 2 * pi * i / numPoints;
      const double s = sin(a), c = cos(a);
      pSf[2 * i]                     = (FFTLIB_F32) (0.5 * (1 - s));
      pSf[2 * i + 1]                 = (FFTLIB_F32) (0.5 * (-c));
      pSf[2 * numPoints - 2 * i - 2] = (FFTLIB_F32) (0.5 * (1 + s));
      pSf[2 * numPoints - 2 * i - 1] = (FFTLIB_F32) (0.5 * c);
   }
}

void ifft_c2r_split_factor_gen(FFTLIB_F32 *pSf, int32_t numPoints)
{
   const double pi = 3.141592654;
   for (int32_t i = 0; i < numPoints / 2; ++i) {
      const double a             = 2 * pi * i / numPoints;
      const double s = sin(a), c = cos(a);
      pSf[2 * i]                 = (FFTLIB_F32) (0.5 * (1 - s));
      pSf[2 * i + 1]             = (FFTLIB_F32) (0.5 * c);
      pSf[2 * i + numPoints]     = (FFTLIB_F32) (0.5 * (1 + s));
      pSf[2 * i + numPoints + 1] = (FFTLIB_F32) (0.5 * (-c));
   }
}

}  // extern "C"
