// Process-wide GPU runtime state of the C-ABI layer: selected function style,
// HIP stream, completion mode, last error text, and the device staging
// buffers used when callers hand over host memory.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <functional>

#include "FFTLIB_types.h"
#include "fft_plan.h"

namespace fftlib {

// One buffer of a transform call: caller pointer and byte count. `out`
// buffers are written back to the caller when staged.
struct IoBuffer {
   const void *ptr;
   size_t      bytes;
};

struct CallBuffers {
   IoBuffer in;     // transform input
   IoBuffer tw;     // twiddle table
   IoBuffer aux0;   // split factors (real transforms) or shift vector (int16)
   void    *out;    // transform output
   size_t   outBytes;
};

// Device memory a call takes its scratch and per-length tables from in
// place of the library's own buffers (the allocation-free entry point): a
// region of `bytes` at device address `base`, handed out front to back in
// 256-byte aligned pieces. A launcher given a workspace allocates nothing.
struct Workspace {
   char  *base  = nullptr;
   size_t bytes = 0;
   size_t used  = 0;
   // Next `n` bytes of the region (hipErrorOutOfMemory when they do not fit).
   hipError_t take(size_t n, void **ptr)
   {
      const size_t at = (used + 255u) & ~(size_t) 255u;
      if (at > bytes || n > bytes - at) return hipErrorOutOfMemory;
      *ptr = base + at;
      used = at + n;
      return hipSuccess;
   }
};

// Device pointers handed to a launcher; ws: null = the library's own scratch
// buffers and tables.
struct DeviceBuffers {
   const void *in;
   const void *tw;
   const void *aux0;
   void       *out;
   Workspace  *ws = nullptr;
};

using LaunchFn = hipError_t (*)(const Plan &plan, const DeviceBuffers &d, hipStream_t stream);

// Runs `launch` with the caller's buffers according to the plan's style:
// FFTLIB_FUNCTION_GPU uses them in place; FFTLIB_FUNCTION_OPTIMIZED stages
// them through device memory. Returns a FFTLIB status and records HIP errors.
FFTLIB_STATUS run_on_gpu(const Plan &plan, const CallBuffers &b, LaunchFn launch);

// The same for any launcher: `style` FFTLIB_FUNCTION_GPU uses the buffers in
// place, FFTLIB_FUNCTION_OPTIMIZED stages them (the staged input buffer is
// library memory the launcher may overwrite).
FFTLIB_STATUS run_buffers(uint32_t style, const CallBuffers &b,
                          const std::function<hipError_t(const DeviceBuffers &, hipStream_t)> &launch);

FFTLIB_FUNCTION_STYLE current_style();
hipStream_t           current_stream();
bool                  synchronous();
bool                  specialized_kernels();
int                   c2c_variant();
void                  set_last_error(const char *what, hipError_t err);
void                  set_last_error_text(const char *text);

// Buffers and launcher of one call of an fft_c7x plan (numChannels and mode
// as the call gives them). `device`: the buffers are used in place and must
// be aligned for the kernels. Returns the status the kernel's _kernel reports
// for bad buffers.
FFTLIB_STATUS describe_call(const Plan &p, const void *pX, const void *pW, const void *pAux, void *pY, bool device,
                            CallBuffers *b, LaunchFn *launch);

// Library-owned device scratch of at least `bytes` for work enqueued on
// `stream` of the current device: one grow-only buffer per (device, stream,
// role), kept until the process ends, so calls on different streams never
// share one. Growing it waits for the stream's earlier work. Roles: the
// intermediate of long complex transforms, and the complex half-length
// spectrum of long real transforms (which itself runs a long complex one).
enum ScratchRole : int { SCRATCH_COMPLEX = 0, SCRATCH_REAL = 1 };
hipError_t device_scratch(size_t bytes, hipStream_t stream, int role, void **ptr);

// Bytes of workspace one call of plan p takes when it runs with a Workspace
// (the same route choices as its launcher; 256-byte alignment of every piece
// included). FFTLIB_ERR_NOT_IMPLEMENTED when the call's route keeps device
// tables that are built on first use (the split-precision MFMA variants).
FFTLIB_STATUS workspace_bytes(const Plan &p, size_t *bytes);

// Plan of pBlock as an _init call wrote it (any of the ten fft_c7x kernels).
FFTLIB_STATUS load_plan(const void *pBlock, Plan *p);

}  // namespace fftlib
