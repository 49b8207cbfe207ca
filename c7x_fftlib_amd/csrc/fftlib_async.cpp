// Asynchronous execution of fft_c7x plans (fftlib_gpu.h).
//
// FFTLIB_GPU_kernelAsync enqueues the transform an _init call planned in a
// pBlock on the caller's HIP stream, with device buffers, and returns.
//
// The host pipeline serves callers whose data lives in (pageable) host
// memory. Each of its `depth` slots owns a HIP stream, pinned staging for the
// call's inputs and output, and device buffers. A submit copies the inputs
// into the slot's pinned staging (CPU), then enqueues host-to-device copies,
// the transform and the device-to-host copy on the slot's stream and
// returns; the output reaches the caller's pY when the slot is next reused
// or at pipelineWait. With two or more slots the PCIe copies of one call
// overlap the transform of the previous one.
#include <string.h>

#include <vector>

#include "fftlib.h"
#include "fftlib_runtime.h"

using namespace fftlib;

namespace {

FFTLIB_STATUS fail(const char *what, hipError_t err)
{
   set_last_error(what, err);
   return FFTLIB_ERR_FAILURE;
}

// Plan of pBlock as _init wrote it (any of the ten kernels).
FFTLIB_STATUS load_plan(const void *pBlock, Plan *p)
{
   if (pBlock == nullptr) return FFTLIB_ERR_NULL_POINTER;
   memcpy(p, pBlock, sizeof *p);
   if (p->magic != kPlanMagic || p->kernel < K_C2C_BATCHED || p->kernel > K_C2R_SINGLE) return FFTLIB_ERR_INVALID_VALUE;
   if (p->style == FFTLIB_FUNCTION_NATC) return FFTLIB_ERR_NOT_IMPLEMENTED;
   return FFTLIB_SUCCESS;
}

size_t round_up(size_t v, size_t a) { return (v + a - 1u) / a * a; }

struct Slot {
   hipStream_t stream = nullptr;
   hipEvent_t  done   = nullptr;
   char       *hostIn = nullptr;  // pinned: input | twiddles | aux
   size_t      hostInCap = 0;
   char       *hostOut   = nullptr;  // pinned output
   size_t      hostOutCap = 0;
   char       *devIn      = nullptr;  // device: input | twiddles | aux
   size_t      devInCap   = 0;
   char       *devOut     = nullptr;
   size_t      devOutCap  = 0;
   bool        busy       = false;
   void       *pendingY   = nullptr;
   size_t      pendingBytes = 0;
};

// Grow-only buffer of at least `bytes` (contents are not kept).
template <typename Alloc, typename Free>
hipError_t grow(char **buf, size_t *cap, size_t bytes, Alloc alloc, Free release)
{
   if (bytes <= *cap) return hipSuccess;
   if (*buf) {
      hipError_t e = release(*buf);
      *buf = nullptr, *cap = 0;
      if (e != hipSuccess) return e;
   }
   const size_t want = round_up(bytes, 1u << 16);
   hipError_t   e    = alloc((void **) buf, want);
   if (e == hipSuccess) *cap = want;
   else *buf = nullptr;
   return e;
}

// Waits for the slot's call and hands its output to the caller.
hipError_t finish(Slot &s)
{
   if (!s.busy) return hipSuccess;
   hipError_t e = hipEventSynchronize(s.done);
   s.busy       = false;
   if (e != hipSuccess) return e;
   memcpy(s.pendingY, s.hostOut, s.pendingBytes);
   return hipSuccess;
}

}  // namespace

struct FFTLIB_GPU_Pipeline_s {
   int               device = 0;
   uint32_t          next   = 0;  // slot of the oldest call (round robin)
   std::vector<Slot> slots;
};

extern "C" {

FFTLIB_STATUS FFTLIB_GPU_kernelAsync(void *pBlock, const void *pX, const void *pW, const void *pAux, void *pY,
                                     void *hipStream)
{
   Plan          p;
   FFTLIB_STATUS st = load_plan(pBlock, &p);
   if (st != FFTLIB_SUCCESS) return st;
   CallBuffers b;
   LaunchFn    launch;
   if ((st = describe_call(p, pX, pW, pAux, pY, true, &b, &launch)) != FFTLIB_SUCCESS) return st;
   if (b.in.bytes == 0 || b.outBytes == 0) return FFTLIB_SUCCESS;
   const DeviceBuffers d{b.in.ptr, b.tw.ptr, b.aux0.ptr, b.out};
   hipError_t          e = launch(p, d, (hipStream_t) hipStream);
   if (e != hipSuccess) return fail("kernel launch", e);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_pipelineCreate(FFTLIB_GPU_Pipeline **pipeline, uint32_t depth)
{
   if (pipeline == nullptr) return FFTLIB_ERR_NULL_POINTER;
   *pipeline = nullptr;
   if (depth < 1u || depth > 8u) return FFTLIB_ERR_INVALID_VALUE;
   FFTLIB_GPU_Pipeline *q = new FFTLIB_GPU_Pipeline;
   hipError_t           e = hipGetDevice(&q->device);
   q->slots.resize(depth);
   for (Slot &s : q->slots) {
      if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
   }
   if (e != hipSuccess) {
      FFTLIB_GPU_pipelineDestroy(q);
      return fail("pipeline create", e);
   }
   *pipeline = q;
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_pipelineSubmit(FFTLIB_GPU_Pipeline *pipeline, void *pBlock, const void *pX, const void *pW,
                                        const void *pAux, void *pY)
{
   if (pipeline == nullptr) return FFTLIB_ERR_NULL_POINTER;
   Plan          p;
   FFTLIB_STATUS st = load_plan(pBlock, &p);
   if (st != FFTLIB_SUCCESS) return st;
   CallBuffers b;
   LaunchFn    launch;
   if ((st = describe_call(p, pX, pW, pAux, pY, false, &b, &launch)) != FFTLIB_SUCCESS) return st;
   if (b.in.bytes == 0 || b.outBytes == 0) return FFTLIB_SUCCESS;

   int        dev = 0;
   hipError_t e   = hipGetDevice(&dev);
   if (e == hipSuccess && dev != pipeline->device) e = hipSetDevice(pipeline->device);
   if (e != hipSuccess) return fail("hipSetDevice", e);

   Slot &s = pipeline->slots[pipeline->next];
   if ((e = finish(s)) != hipSuccess) return fail("pipeline wait", e);

   // staging layout: input | twiddles | aux, each 256-byte aligned
   const size_t offTw  = round_up(b.in.bytes, 256);
   const size_t offAux = offTw + round_up(b.tw.bytes, 256);
   const size_t inAll  = offAux + round_up(b.aux0.bytes, 256);
   auto hostAlloc = [](void **ptr, size_t n) { return hipHostMalloc(ptr, n, hipHostMallocDefault); };
   auto hostFree  = [](void *ptr) { return hipHostFree(ptr); };
   auto devAlloc  = [](void **ptr, size_t n) { return hipMalloc(ptr, n); };
   auto devFree   = [](void *ptr) { return hipFree(ptr); };
   if ((e = grow(&s.hostIn, &s.hostInCap, inAll, hostAlloc, hostFree)) != hipSuccess ||
       (e = grow(&s.hostOut, &s.hostOutCap, b.outBytes, hostAlloc, hostFree)) != hipSuccess ||
       (e = grow(&s.devIn, &s.devInCap, inAll, devAlloc, devFree)) != hipSuccess ||
       (e = grow(&s.devOut, &s.devOutCap, b.outBytes, devAlloc, devFree)) != hipSuccess)
      return fail("pipeline staging allocation", e);

   memcpy(s.hostIn, b.in.ptr, b.in.bytes);
   memcpy(s.hostIn + offTw, b.tw.ptr, b.tw.bytes);
   if (b.aux0.ptr) memcpy(s.hostIn + offAux, b.aux0.ptr, b.aux0.bytes);
   if ((e = hipMemcpyAsync(s.devIn, s.hostIn, inAll, hipMemcpyHostToDevice, s.stream)) != hipSuccess)
      return fail("hipMemcpyAsync (to device)", e);
   const DeviceBuffers d{s.devIn, s.devIn + offTw, b.aux0.ptr ? s.devIn + offAux : nullptr, s.devOut};
   if ((e = launch(p, d, s.stream)) != hipSuccess) return fail("kernel launch", e);
   if ((e = hipMemcpyAsync(s.hostOut, s.devOut, b.outBytes, hipMemcpyDeviceToHost, s.stream)) != hipSuccess)
      return fail("hipMemcpyAsync (to host)", e);
   if ((e = hipEventRecord(s.done, s.stream)) != hipSuccess) return fail("hipEventRecord", e);
   s.busy          = true;
   s.pendingY      = pY;
   s.pendingBytes  = b.outBytes;
   pipeline->next  = (pipeline->next + 1u) % (uint32_t) pipeline->slots.size();
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_pipelineWait(FFTLIB_GPU_Pipeline *pipeline)
{
   if (pipeline == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const uint32_t n     = (uint32_t) pipeline->slots.size();
   hipError_t     first = hipSuccess;
   // oldest first, so that two calls writing the same pY leave the later result
   for (uint32_t i = 0; i < n; ++i) {
      hipError_t e = finish(pipeline->slots[(pipeline->next + i) % n]);
      if (first == hipSuccess) first = e;
   }
   if (first != hipSuccess) return fail("pipeline wait", first);
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_GPU_pipelineDestroy(FFTLIB_GPU_Pipeline *pipeline)
{
   if (pipeline == nullptr) return FFTLIB_ERR_NULL_POINTER;
   hipError_t first = hipSuccess;
   auto       keep  = [&first](hipError_t e) {
      if (first == hipSuccess) first = e;
   };
   for (Slot &s : pipeline->slots) {
      if (s.stream) keep(hipStreamSynchronize(s.stream));
      s.busy = false;  // outputs of calls never waited for are dropped
      if (s.hostIn) keep(hipHostFree(s.hostIn));
      if (s.hostOut) keep(hipHostFree(s.hostOut));
      if (s.devIn) keep(hipFree(s.devIn));
      if (s.devOut) keep(hipFree(s.devOut));
      if (s.done) keep(hipEventDestroy(s.done));
      if (s.stream) keep(hipStreamDestroy(s.stream));
   }
   delete pipeline;
   if (first != hipSuccess) return fail("pipeline destroy", first);
   return FFTLIB_SUCCESS;
}

}  // extern "C"
