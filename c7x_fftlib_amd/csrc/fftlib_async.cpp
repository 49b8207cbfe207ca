// Asynchronous execution of fft_c7x plans (fftlib_gpu.h).
//
// FFTLIB_GPU_kernelAsync enqueues the transform an _init call planned in a
// pBlock on the caller's HIP stream, with device buffers, and returns.
//
// The host pipeline serves callers whose data lives in host memory. Each of
// its `depth` slots owns a HIP stream, pinned staging for the call's inputs
// and output, and device buffers. A submit enqueues host-to-device copies,
// the transform and the device-to-host copy on the slot's stream and
// returns. Buffers the runtime can reach directly (pinned host memory from
// hipHostMalloc, or device memory) are copied by DMA straight from and to
// the caller's pointers; pageable inputs are first copied into the slot's
// pinned staging on the CPU, and a pageable output reaches the caller's pY
// when the slot is next reused or at pipelineWait. With two or more slots
// the PCIe copies of one call overlap the transform of the previous one.
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

size_t round_up(size_t v, size_t a) { return (v + a - 1u) / a * a; }

// True for memory the GPU's copy engines can address directly: pinned host
// memory or device memory. Pageable memory reports no HIP allocation (an
// error status, cleared here so that it is not taken for a launch error).
bool dma_reachable(const void *p)
{
   hipPointerAttribute_t attr;
   if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
      (void) hipGetLastError();
      return false;
   }
   return attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

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
   if (s.pendingY) memcpy(s.pendingY, s.hostOut, s.pendingBytes);
   return hipSuccess;
}

}  // namespace

FFTLIB_STATUS fftlib::load_plan(const void *pBlock, Plan *p)
{
   if (pBlock == nullptr) return FFTLIB_ERR_NULL_POINTER;
   memcpy(p, pBlock, sizeof *p);
   if (p->magic != kPlanMagic || p->kernel < K_C2C_BATCHED || p->kernel > K_C2R_SINGLE) return FFTLIB_ERR_INVALID_VALUE;
   if (p->style == FFTLIB_FUNCTION_NATC) return FFTLIB_ERR_NOT_IMPLEMENTED;
   return FFTLIB_SUCCESS;
}

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

   // staging layout: input | twiddles | aux, each 256-byte aligned; pinned
   // staging is only needed for buffers the copy engines cannot reach
   const IoBuffer ins[3] = {b.in, b.tw, b.aux0};
   const size_t   off[3] = {0, round_up(b.in.bytes, 256), round_up(b.in.bytes, 256) + round_up(b.tw.bytes, 256)};
   const size_t   inAll  = off[2] + round_up(b.aux0.bytes, 256);
   bool           reach[3];
   bool           staged = false;
   for (int i = 0; i < 3; ++i) {
      reach[i] = ins[i].ptr == nullptr || dma_reachable(ins[i].ptr);
      staged   = staged || !reach[i];
   }
   const bool direct    = dma_reachable(pY);
   auto       hostAlloc = [](void **ptr, size_t n) { return hipHostMalloc(ptr, n, hipHostMallocDefault); };
   auto       hostFree  = [](void *ptr) { return hipHostFree(ptr); };
   auto       devAlloc  = [](void **ptr, size_t n) { return hipMalloc(ptr, n); };
   auto       devFree   = [](void *ptr) { return hipFree(ptr); };
   if ((staged && (e = grow(&s.hostIn, &s.hostInCap, inAll, hostAlloc, hostFree)) != hipSuccess) ||
       (!direct && (e = grow(&s.hostOut, &s.hostOutCap, b.outBytes, hostAlloc, hostFree)) != hipSuccess) ||
       (e = grow(&s.devIn, &s.devInCap, inAll, devAlloc, devFree)) != hipSuccess ||
       (e = grow(&s.devOut, &s.devOutCap, b.outBytes, devAlloc, devFree)) != hipSuccess)
      return fail("pipeline staging allocation", e);

   for (int i = 0; i < 3; ++i) {
      if (ins[i].ptr == nullptr || ins[i].bytes == 0) continue;
      const void *src = ins[i].ptr;
      if (!reach[i]) {
         memcpy(s.hostIn + off[i], src, ins[i].bytes);
         src = s.hostIn + off[i];
      }
      if ((e = hipMemcpyAsync(s.devIn + off[i], src, ins[i].bytes, hipMemcpyDefault, s.stream)) != hipSuccess)
         return fail("hipMemcpyAsync (to device)", e);
   }
   const DeviceBuffers d{s.devIn, s.devIn + off[1], b.aux0.ptr ? s.devIn + off[2] : nullptr, s.devOut};
   if ((e = launch(p, d, s.stream)) != hipSuccess) return fail("kernel launch", e);
   // outputs leave in submission order (two calls may write one pY)
   const uint32_t n    = (uint32_t) pipeline->slots.size();
   const Slot    &prev = pipeline->slots[(pipeline->next + n - 1u) % n];
   if (n > 1u && prev.busy && (e = hipStreamWaitEvent(s.stream, prev.done, 0)) != hipSuccess)
      return fail("hipStreamWaitEvent", e);
   if ((e = hipMemcpyAsync(direct ? pY : (void *) s.hostOut, s.devOut, b.outBytes, hipMemcpyDefault, s.stream)) !=
       hipSuccess)
      return fail("hipMemcpyAsync (to host)", e);
   if ((e = hipEventRecord(s.done, s.stream)) != hipSuccess) return fail("hipEventRecord", e);
   s.busy          = true;
   s.pendingY      = direct ? nullptr : pY;
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
