// Allocation-free asynchronous calls (FFTLIB_GPU_kernelAsyncNoAlloc,
// fftlib_gpu.h): FFTLIB_GPU_kernelAsync with every scratch buffer and
// per-length table of the call taken from `scratch`, a statically allocated
// device array (one copy per GPU, from the library's load until the process
// ends). The entry point calls no hipMalloc / hipFree.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is taken once per device with hipGetSymbolAddress (with
// that device current) and only that address is handed to the kernels.
#include <hip/hip_runtime.h>

#include <stdio.h>

#include <mutex>

#include "fftlib.h"
#include "fftlib_runtime.h"

using namespace fftlib;

namespace {

constexpr size_t kScratchBytes = FFTLIB_GPU_NOALLOC_SCRATCH_BYTES;
constexpr int    kMaxDevices   = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void      *g_scratch[kMaxDevices];   // device address of `scratch` on each device, once taken

// Device address of the current device's scratch array.
hipError_t scratch_address(void **ptr)
{
   int        dev = 0;
   hipError_t e   = hipGetDevice(&dev);
   if (e != hipSuccess) return e;
   if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
   std::lock_guard<std::mutex> lock(g_mu);
   if (g_scratch[dev] == nullptr) {
      void *p = nullptr;
      if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
      if (p == nullptr) return hipErrorInvalidSymbol;
      g_scratch[dev] = p;
   }
   *ptr = g_scratch[dev];
   return hipSuccess;
}

FFTLIB_STATUS fail(const char *what, hipError_t err)
{
   set_last_error(what, err);
   return FFTLIB_ERR_FAILURE;
}

}  // namespace

extern "C" {

FFTLIB_STATUS FFTLIB_GPU_kernelAsyncNoAlloc(void *pBlock, const void *pX, const void *pW, const void *pAux, void *pY,
                                            void *hipStream)
{
   Plan          p;
   FFTLIB_STATUS st = load_plan(pBlock, &p);
   if (st != FFTLIB_SUCCESS) return st;
   CallBuffers b;
   LaunchFn    launch;
   if ((st = describe_call(p, pX, pW, pAux, pY, true, &b, &launch)) != FFTLIB_SUCCESS) return st;
   if (b.in.bytes == 0 || b.outBytes == 0) return FFTLIB_SUCCESS;
   size_t need = 0;
   if ((st = workspace_bytes(p, &need)) != FFTLIB_SUCCESS) return st;
   if (need > kScratchBytes) {
      char msg[160];
      snprintf(msg, sizeof msg, "call needs %zu bytes of scratch; the allocation-free entry point holds %zu", need,
               kScratchBytes);
      set_last_error_text(msg);
      return FFTLIB_ERR_INVALID_DIMENSION;
   }
   Workspace ws;
   if (need > 0) {
      void      *base = nullptr;
      hipError_t e    = scratch_address(&base);
      if (e != hipSuccess) return fail("hipGetSymbolAddress (scratch)", e);
      ws.base  = (char *) base;
      ws.bytes = kScratchBytes;
   }
   const DeviceBuffers d{b.in.ptr, b.tw.ptr, b.aux0.ptr, b.out, &ws};
   const hipError_t    e = launch(p, d, (hipStream_t) hipStream);
   if (e != hipSuccess) return fail("kernel launch", e);
   return FFTLIB_SUCCESS;
}

}  // extern "C"
