// C-ABI entry points of the ten fft_c7x kernels: _init writes a GPU plan into
// the caller's pBlock, _kernel runs the plan on the GPU, _checkParams mirrors
// the reference's optimized-kernel parameter checks (c7504 variant where the
// reference has one, c71 otherwise; the reference build picks the c7504
// sources for the C7504/C7524 devices its published numbers come from).
//
// Behaviour beyond the reference: _kernel validates the plan and refuses NULL
// buffers instead of faulting, and pX is never written (the reference uses it
// as scratch). The real transforms do not write pXFFT (half lengths that do
// not fit one workgroup's LDS go through library scratch instead).
#include <stdint.h>
#include <string.h>

#include "fftlib.h"
#include "fft_fixed_class.h"
#include "fft_kernels.h"
#include "fftlib_runtime.h"

using namespace fftlib;

namespace {

constexpr uint32_t kMaxComplexPoints = 1u << 20;

bool misaligned(const void *p, uintptr_t a) { return ((uintptr_t) p & (a - 1u)) != 0; }

bool smooth235(uint32_t n)
{
   // same loop bounds as the reference checks (at most 32 divisions per prime)
   for (int i = 0; i < 32 && n % 5u == 0u; ++i) n /= 5u;
   for (int i = 0; i < 32 && n % 3u == 0u; ++i) n /= 3u;
   for (int i = 0; i < 32 && n % 2u == 0u; ++i) n /= 2u;
   return n == 1u;
}

uint32_t radix4_stages(const StageTable &st)
{
   uint32_t c = 0;
   for (uint32_t s = 0; s < st.nStages; ++s) c += st.radix[s] == 4 ? 1u : 0u;
   return c;
}

FFTLIB_STATUS write_plan(void *pBlock, KernelId kernel, uint32_t numPoints, uint32_t complexN, uint32_t numChannels,
                         uint32_t mode, uint32_t numShifts)
{
   if (pBlock == nullptr) return FFTLIB_ERR_NULL_POINTER;
   Plan p;
   memset(&p, 0, sizeof p);
   p.magic       = kPlanMagic;
   p.kernel      = kernel;
   p.style       = (uint32_t) current_style();
   p.numPoints   = numPoints;
   p.numChannels = numChannels;
   p.mode        = mode;
   p.numShifts   = numShifts;
   if (complexN == 0 || complexN > kMaxComplexPoints) return FFTLIB_ERR_INVALID_DIMENSION;
   if (!build_stage_table(complexN, &p.st)) return FFTLIB_ERR_INVALID_DIMENSION;
   if ((kernel == K_I16_BATCHED || kernel == K_I16_SINGLE) && !p.st.pow2) return FFTLIB_ERR_INVALID_DIMENSION;
   memcpy(pBlock, &p, sizeof p);
   return FFTLIB_SUCCESS;
}

// The int16 kernels keep a whole transform in one workgroup's LDS; longer
// ones are refused at _kernel time.
FFTLIB_STATUS refuse_too_long(const Plan &p)
{
   if (fits_lds(p.st)) return FFTLIB_SUCCESS;
   set_last_error_text("int16 transform longer than one workgroup's LDS");
   return FFTLIB_ERR_NOT_IMPLEMENTED;
}

// Reads and validates the plan an _init call left in pBlock.
FFTLIB_STATUS read_plan(const void *pBlock, KernelId kernel, uint32_t numPoints, Plan *p)
{
   if (pBlock == nullptr) return FFTLIB_ERR_NULL_POINTER;
   memcpy(p, pBlock, sizeof *p);
   if (p->magic != kPlanMagic || p->kernel != (uint32_t) kernel) return FFTLIB_ERR_INVALID_VALUE;
   if (p->numPoints != numPoints) return FFTLIB_ERR_INVALID_DIMENSION;
   if (p->style == FFTLIB_FUNCTION_NATC) return FFTLIB_ERR_NOT_IMPLEMENTED;
   return FFTLIB_SUCCESS;
}

// ---- launchers bound to plans ---------------------------------------------
// Routes of a batched complex FFT of st.n points: the MFMA kernel when
// selected, the specialised (dedicated or class) kernel of the length, two
// passes on the class kernels for long lengths, else the generic LDS or
// outer-pass kernels.
enum ComplexRoute { ROUTE_MFMA, ROUTE_FIXED, ROUTE_TWO_PASS, ROUTE_LARGE, ROUTE_LDS };

ComplexRoute complex_route(const StageTable &st)
{
   uint32_t n1 = 0, m2 = 0;
   if (c2c_variant() != C2C_VARIANT_FP32 && mfma_c2c_supported(st.n)) return ROUTE_MFMA;
   if (specialized_kernels() && fixed_c2c_supported(st.n)) return ROUTE_FIXED;
   if (specialized_kernels() && two_pass_split(st.n, &n1, &m2)) return ROUTE_TWO_PASS;
   if (!fits_lds(st)) return ROUTE_LARGE;
   return ROUTE_LDS;
}

size_t aligned256(size_t v) { return (v + 255u) & ~(size_t) 255u; }

// Workspace bytes of one complex FFT call (0 for the routes without scratch;
// SIZE_MAX for the MFMA route, whose tables are built on first use).
size_t complex_workspace(const StageTable &st, uint32_t numChannels)
{
   switch (complex_route(st)) {
   case ROUTE_MFMA: return SIZE_MAX;
   case ROUTE_TWO_PASS:
      return aligned256(two_pass_table_bytes(st.n)) + aligned256(two_pass_scratch_points(st.n, numChannels) * 8u);
   case ROUTE_LARGE: return aligned256((size_t) st.n * numChannels * 8u);
   default: return 0;
   }
}

// Scratch of a launcher: from the call's workspace when it has one, else the
// library's grow-only buffer of (device, stream, role).
hipError_t call_scratch(Workspace *ws, size_t bytes, hipStream_t s, int role, void **ptr)
{
   if (ws) return ws->take(bytes, ptr);
   return device_scratch(bytes, s, role, ptr);
}

hipError_t run_complex(const StageTable &st, const float *x, float *y, const float *tw, uint32_t numChannels,
                       int mode, hipStream_t s, Workspace *ws)
{
   switch (complex_route(st)) {
   case ROUTE_MFMA:
      if (ws) return hipErrorNotSupported;   // device tables built on first use
      return launch_c2c_mfma(x, y, st.n, numChannels, mode, c2c_variant(), s);
   case ROUTE_FIXED: return launch_c2c_fixed(x, y, tw, st.n, numChannels, mode, s);
   case ROUTE_TWO_PASS: {
      void      *scratch = nullptr;
      hipError_t e       = call_scratch(ws, two_pass_scratch_points(st.n, numChannels) * 8u, s, SCRATCH_COMPLEX,
                                        &scratch);
      if (e != hipSuccess) return e;
      if (ws == nullptr) return launch_c2c_two_pass(x, y, st.n, numChannels, mode, (float *) scratch, s);
      void       *tables = nullptr;
      TwoPassPlan plan;
      if ((e = ws->take(two_pass_table_bytes(st.n), &tables)) != hipSuccess) return e;
      if ((e = two_pass_plan_at(st.n, (float *) tables, s, &plan)) != hipSuccess) return e;
      return launch_c2c_two_pass_plan(plan, x, y, numChannels, mode, (float *) scratch, s);
   }
   case ROUTE_LARGE: {
      void      *scratch = nullptr;
      hipError_t e       = call_scratch(ws, (size_t) st.n * numChannels * 8u, s, SCRATCH_COMPLEX, &scratch);
      if (e != hipSuccess) return e;
      return launch_c2c_large(x, y, tw, st, numChannels, mode, (float *) scratch, s);
   }
   default: return launch_c2c(x, y, tw, st, numChannels, mode, s);
   }
}

hipError_t run_c2c(const Plan &p, const DeviceBuffers &d, hipStream_t s)
{
   int mode = C2C_FORWARD;
   if (p.kernel == K_IC2C_BATCHED) mode = C2C_INVERSE_PRESCALE;
   if (p.kernel == K_IC2C_SINGLE) mode = C2C_INVERSE_POSTDIV;
   return run_complex(p.st, (const float *) d.in, (float *) d.out, (const float *) d.tw, p.numChannels, mode, s, d.ws);
}

// Real transforms: the fused kernel of the half length (complex FFT in LDS
// plus the split) where one is specialised; otherwise, from half length
// kRealSplitMin on, the complex FFT on the multi-length / two-pass kernels
// and the split as a second pass through library scratch; the generic fused
// kernel for shorter half lengths, or when the specialised kernels are
// switched off and the half length fits one workgroup. Measured on MI355X
// (half lengths 20-17280, 1 to 2^25/N channels) the two-pass route is 2-15x
// faster than the generic fused kernel, which runs one transform per
// workgroup with its twiddles and index arithmetic unspecialised, except for
// calls of at most kRealFusedMaxPoints half-length points, where the second
// launch costs more (about 1 us) than it saves.
constexpr uint32_t kRealSplitMin       = 16;
constexpr uint64_t kRealFusedMaxPoints = 256;

bool real_fused_generic(const Plan &p)
{
   if (!fits_lds(p.st)) return false;
   return !specialized_kernels() || p.st.n < kRealSplitMin ||
          (uint64_t) p.st.n * p.numChannels <= kRealFusedMaxPoints;
}

// True when a real transform runs as a complex pass plus a split pass
// through the half-length spectrum buffer.
bool real_two_launch(const Plan &p)
{
   return !(specialized_kernels() && fixed_real_supported(p.st.n)) && !real_fused_generic(p);
}

hipError_t run_r2c(const Plan &p, const DeviceBuffers &d, hipStream_t s)
{
   const float *x = (const float *) d.in, *tw = (const float *) d.tw, *sf = (const float *) d.aux0;
   float       *y = (float *) d.out;
   if (specialized_kernels() && fixed_real_supported(p.st.n))
      return launch_r2c_fixed(x, y, tw, sf, p.st.n, p.numChannels, p.mode, s);
   if (real_fused_generic(p)) return launch_r2c(x, y, tw, sf, p.st, p.numChannels, p.mode, s);
   void      *z = nullptr;
   hipError_t e = call_scratch(d.ws, (size_t) p.st.n * p.numChannels * 8u, s, SCRATCH_REAL, &z);
   if (e == hipSuccess) e = run_complex(p.st, x, (float *) z, tw, p.numChannels, C2C_FORWARD, s, d.ws);
   if (e == hipSuccess) e = launch_r2c_split((const float *) z, y, sf, p.st.n, p.numChannels, p.mode, s);
   return e;
}

hipError_t run_c2r(const Plan &p, const DeviceBuffers &d, hipStream_t s)
{
   const float *x = (const float *) d.in, *tw = (const float *) d.tw, *sf = (const float *) d.aux0;
   float       *y       = (float *) d.out;
   const int    postdiv = p.kernel == K_C2R_SINGLE ? 1 : 0;
   if (specialized_kernels() && fixed_real_supported(p.st.n))
      return launch_c2r_fixed(x, y, tw, sf, p.st.n, p.numChannels, p.mode, postdiv, s);
   if (real_fused_generic(p)) return launch_c2r(x, y, tw, sf, p.st, p.numChannels, p.mode, postdiv, s);
   void      *z = nullptr;
   hipError_t e = call_scratch(d.ws, (size_t) p.st.n * p.numChannels * 8u, s, SCRATCH_REAL, &z);
   if (e == hipSuccess) e = launch_c2r_presplit(x, (float *) z, sf, p.st.n, p.numChannels, p.mode, s);
   if (e == hipSuccess)
      e = run_complex(p.st, (const float *) z, y, tw, p.numChannels,
                      postdiv ? C2C_INVERSE_POSTDIV : C2C_INVERSE_PRESCALE, s, d.ws);
   return e;
}

hipError_t run_i16(const Plan &p, const DeviceBuffers &d, hipStream_t s)
{
   return launch_i16((const int16_t *) d.in, (int16_t *) d.out, (const int16_t *) d.tw, (const uint32_t *) d.aux0,
                     p.st, p.numChannels, s);
}

}  // namespace

// Buffers and launcher of one call of plan p (numChannels and mode as the
// call gives them). `device`: the buffers are used in place, so they must be
// aligned for the kernels' vector accesses. The checks run in the order the
// entry points of each kernel family report them.
FFTLIB_STATUS fftlib::describe_call(const Plan &p, const void *pX, const void *pW, const void *pAux, void *pY,
                                    bool device, CallBuffers *b, LaunchFn *launch)
{
   const size_t  tw = (size_t) p.st.twCount * 2u;
   FFTLIB_STATUS st;
   switch (p.kernel) {
   case K_C2C_BATCHED:
   case K_C2C_SINGLE:
   case K_IC2C_BATCHED:
   case K_IC2C_SINGLE: {
      if (pX == nullptr || pW == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
      if (device && (misaligned(pX, 8) || misaligned(pW, 8) || misaligned(pY, 8)))
         return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
      const size_t data = (size_t) p.numPoints * p.numChannels * 2u * sizeof(float);
      *b                = CallBuffers{{pX, data}, {pW, tw * sizeof(float)}, {nullptr, 0}, pY, data};
      *launch           = run_c2c;
      return FFTLIB_SUCCESS;
   }
   case K_I16_BATCHED:
   case K_I16_SINGLE: {
      if (pX == nullptr || pW == nullptr || pY == nullptr || pAux == nullptr) return FFTLIB_ERR_NULL_POINTER;
      if ((st = refuse_too_long(p)) != FFTLIB_SUCCESS) return st;
      if (device && (misaligned(pX, 4) || misaligned(pW, 4) || misaligned(pY, 4) || misaligned(pAux, 4)))
         return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
      const uint32_t need = radix4_stages(p.st);
      if (p.numShifts < need) return FFTLIB_ERR_INVALID_DIMENSION;
      const size_t data = (size_t) p.numPoints * p.numChannels * 2u * sizeof(int16_t);
      *b = CallBuffers{{pX, data}, {pW, tw * sizeof(int16_t)}, {pAux, (size_t) need * sizeof(uint32_t)}, pY, data};
      *launch = run_i16;
      return FFTLIB_SUCCESS;
   }
   case K_R2C_BATCHED:
   case K_R2C_SINGLE:
   case K_C2R_BATCHED:
   case K_C2R_SINGLE: {
      const bool r2c = p.kernel == K_R2C_BATCHED || p.kernel == K_R2C_SINGLE;
      if (p.mode > 1u) return r2c ? FFTLIB_ERR_INVALID_OUTMODE : FFTLIB_ERR_INVALID_VALUE;
      if (pX == nullptr || pW == nullptr || pAux == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
      if (device && (misaligned(pX, 8) || misaligned(pW, 8) || misaligned(pY, 8) || misaligned(pAux, 4)))
         return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
      const uint32_t nSpec = p.mode ? p.numPoints : p.numPoints / 2u + 1u;
      const size_t   real  = (size_t) p.numPoints * p.numChannels * sizeof(float);
      const size_t   spec  = (size_t) nSpec * p.numChannels * 2u * sizeof(float);
      const IoBuffer sf{pAux, (size_t) p.numPoints * 2u * sizeof(float)};
      if (r2c) *b = CallBuffers{{pX, real}, {pW, tw * sizeof(float)}, sf, pY, spec};
      else *b = CallBuffers{{pX, spec}, {pW, tw * sizeof(float)}, sf, pY, real};
      *launch = r2c ? run_r2c : run_c2r;
      return FFTLIB_SUCCESS;
   }
   default:
      return FFTLIB_ERR_INVALID_VALUE;
   }
}

FFTLIB_STATUS fftlib::workspace_bytes(const Plan &p, size_t *bytes)
{
   size_t need = 0;
   switch (p.kernel) {
   case K_C2C_BATCHED:
   case K_C2C_SINGLE:
   case K_IC2C_BATCHED:
   case K_IC2C_SINGLE: need = complex_workspace(p.st, p.numChannels); break;
   case K_R2C_BATCHED:
   case K_R2C_SINGLE:
   case K_C2R_BATCHED:
   case K_C2R_SINGLE:
      if (real_two_launch(p)) {
         const size_t inner = complex_workspace(p.st, p.numChannels);
         need = inner == SIZE_MAX ? SIZE_MAX : aligned256((size_t) p.st.n * p.numChannels * 8u) + inner;
      }
      break;
   default: break;   // int16: no scratch
   }
   if (need == SIZE_MAX) {
      set_last_error_text("the split-precision C2C variants keep device tables built on first use");
      return FFTLIB_ERR_NOT_IMPLEMENTED;
   }
   *bytes = need;
   return FFTLIB_SUCCESS;
}

namespace {

// _kernel of every fft_c7x kernel: the plan of pBlock with the call's
// channel count (and outmode / inMode where the entry point takes one).
FFTLIB_STATUS exec_plan(KernelId kernel, uint32_t numPoints, uint32_t numChannels, const uint32_t *mode,
                        const void *pX, const void *pW, const void *pAux, void *pY, void *pBlock)
{
   Plan          p;
   FFTLIB_STATUS st = read_plan(pBlock, kernel, numPoints, &p);
   if (st != FFTLIB_SUCCESS) return st;
   p.numChannels = numChannels;
   if (mode) p.mode = *mode;
   CallBuffers b;
   LaunchFn    launch;
   if ((st = describe_call(p, pX, pW, pAux, pY, p.style == FFTLIB_FUNCTION_GPU, &b, &launch)) != FFTLIB_SUCCESS)
      return st;
   if (b.in.bytes == 0 || b.outBytes == 0) return FFTLIB_SUCCESS;
   return run_on_gpu(p, b, launch);
}

FFTLIB_STATUS exec_c2c(KernelId kernel, const FFTLIB_F32 *pX, const FFTLIB_F32 *pW, FFTLIB_F32 *pY,
                       uint32_t numPoints, uint32_t numChannels, void *pBlock)
{
   return exec_plan(kernel, numPoints, numChannels, nullptr, pX, pW, nullptr, pY, pBlock);
}

FFTLIB_STATUS exec_i16(KernelId kernel, const int16_t *pX, const int16_t *pW, int16_t *pY, const uint32_t *pShift,
                       uint32_t numPoints, uint32_t numChannels, void *pBlock)
{
   return exec_plan(kernel, numPoints, numChannels, nullptr, pX, pW, pShift, pY, pBlock);
}

FFTLIB_STATUS exec_r2c(KernelId kernel, const FFTLIB_F32 *pX, const FFTLIB_F32 *pW, const FFTLIB_F32 *pSf,
                       FFTLIB_F32 *pY, uint32_t numPoints, uint32_t numChannels, uint32_t outmode, void *pBlock)
{
   return exec_plan(kernel, numPoints, numChannels, &outmode, pX, pW, pSf, pY, pBlock);
}

FFTLIB_STATUS exec_c2r(KernelId kernel, const FFTLIB_F32 *pX, const FFTLIB_F32 *pW, const FFTLIB_F32 *pSf,
                       FFTLIB_F32 *pY, uint32_t numPoints, uint32_t numChannels, uint32_t inMode, void *pBlock)
{
   return exec_plan(kernel, numPoints, numChannels, &inMode, pX, pW, pSf, pY, pBlock);
}

// ---- parameter checks (reference _ci checkParams) ---------------------------
// fp32 C2C batched, forward and inverse: $C2CB/c7504/..._ci.cpp:2806-2852 and
// $IC2CB/c7504/..._ci.cpp:2898-2946.
FFTLIB_STATUS check_c2c_batched(const void *pX, const FFTLIB_bufParams1D_t *bx, const void *pW,
                                const FFTLIB_bufParams1D_t *bw, const void *pY, const FFTLIB_bufParams1D_t *by,
                                uint32_t numPoints, const void *pBlock)
{
   if (pX == nullptr || pW == nullptr || pY == nullptr || pBlock == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (bx->dim_x != by->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bw->dim_x != numPoints * 2u) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bx->data_type != FFTLIB_FLOAT32 || bw->data_type != FFTLIB_FLOAT32 || by->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   if (misaligned(pX, 16)) return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
   return smooth235(numPoints) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

// Single-channel fp32 checks (C2C1, IC2C1, R2C1, C2R1): X, W and Y descriptors
// share dim_x; N = dim_x / 2 must be 2/3/5-smooth ($C2C1/c7504/..._ci.cpp:
// 2279-2323, $IC2C1/c7504/..._ci.cpp:2381-2425, $R2C1/c71/..._ci.cpp:277-322,
// $C2R1/c71/..._ci.cpp:166-210).
FFTLIB_STATUS check_single_f32(const void *pX, const FFTLIB_bufParams1D_t *bx, const void *pW,
                               const FFTLIB_bufParams1D_t *bw, const void *pY, const FFTLIB_bufParams1D_t *by)
{
   if (pX == nullptr || pW == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (bx->dim_x != bw->dim_x || bx->dim_x != by->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bx->data_type != FFTLIB_FLOAT32 || bw->data_type != FFTLIB_FLOAT32 || by->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   if (misaligned(pX, 16)) return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
   return smooth235(bx->dim_x / 2u) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

// Batched real transforms: X, W, Y share dim_x; N/2 must be smooth; R2C also
// validates outmode ($R2CB/c71/..._ci.cpp:442-517, $C2RB/c71/..._ci.cpp:
// 299-346).
FFTLIB_STATUS check_real_batched(const void *pX, const FFTLIB_bufParams1D_t *bx, const void *pW,
                                 const FFTLIB_bufParams1D_t *bw, const void *pY, const FFTLIB_bufParams1D_t *by,
                                 uint32_t numPoints, int checkOutmode, uint32_t outmode)
{
   if (pX == nullptr || pW == nullptr || pY == nullptr) return FFTLIB_ERR_NULL_POINTER;
   if (bx->dim_x != bw->dim_x || bx->dim_x != by->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bx->data_type != FFTLIB_FLOAT32 || bw->data_type != FFTLIB_FLOAT32 || by->data_type != FFTLIB_FLOAT32)
      return FFTLIB_ERR_INVALID_TYPE;
   if (misaligned(pX, 16)) return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
   if (checkOutmode && outmode != 0u && outmode != 1u) return FFTLIB_ERR_INVALID_OUTMODE;
   return smooth235(numPoints / 2u) ? FFTLIB_SUCCESS : FFTLIB_ERR_INVALID_DIMENSION;
}

}  // namespace

extern "C" {

// ===== fp32 C2C batched forward =============================================
FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         uint32_t numPoints, uint32_t numChannels, void *pBlock)
{
   return write_plan(pBlock, K_C2C_BATCHED, numPoints, numPoints, numChannels, 0, 0);
}

FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                           FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                           FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                           uint32_t numPoints, uint32_t numChannels, void *pBlock)
{
   return exec_c2c(K_C2C_BATCHED, pX, pW, pY, numPoints, numChannels, pBlock);
}

FFTLIB_STATUS FFTLIB_fft1dBatched_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                uint32_t numPoints, uint32_t numChannels,
                                                                void *pBlock)
{
   return check_c2c_batched(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY, numPoints, pBlock);
}

// ===== fp32 C2C batched inverse =============================================
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                          FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                          FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                          uint32_t numPoints, uint32_t numChannels, void *pBlock)
{
   return write_plan(pBlock, K_IC2C_BATCHED, numPoints, numPoints, numChannels, 0, 0);
}

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                            FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                            FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                            uint32_t numPoints, uint32_t numChannels, void *pBlock)
{
   return exec_c2c(K_IC2C_BATCHED, pX, pW, pY, numPoints, numChannels, pBlock);
}

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                 FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                 FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                 uint32_t numPoints, uint32_t numChannels,
                                                                 void *pBlock)
{
   return check_c2c_batched(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY, numPoints, pBlock);
}

// ===== fp32 C2C single channel forward / inverse ============================
FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                  FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                  FFTLIB_bufParams1D_t *bufParamsY, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const uint32_t n = bufParamsX->dim_x >> 1;
   return write_plan(pBlock, K_C2C_SINGLE, n, n, 1, 0, 0);
}

FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                    FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                    FFTLIB_bufParams1D_t *bufParamsY, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   return exec_c2c(K_C2C_SINGLE, pX, pW, pY, bufParamsX->dim_x >> 1, 1, pBlock);
}

FFTLIB_STATUS FFTLIB_fft1d_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         void *pBlock)
{
   return check_single_f32(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY);
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                   FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                   FFTLIB_bufParams1D_t *bufParamsY, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const uint32_t n = bufParamsX->dim_x >> 1;
   return write_plan(pBlock, K_IC2C_SINGLE, n, n, 1, 0, 0);
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_kernel(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                     FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pY,
                                                     FFTLIB_bufParams1D_t *bufParamsY, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   return exec_c2c(K_IC2C_SINGLE, pX, pW, pY, bufParamsX->dim_x >> 1, 1, pBlock);
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                          FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                          FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                          void *pBlock)
{
   return check_single_f32(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY);
}

// ===== int16 batched / single ===============================================
FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_init(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                         FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                         FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                         FFTLIB_bufParams1D_t *bufParamsShift, uint32_t numPoints,
                                                         uint32_t numChannels, void *pBlock)
{
   const uint32_t ns = bufParamsShift ? bufParamsShift->dim_x : 0u;
   return write_plan(pBlock, K_I16_BATCHED, numPoints, numPoints, numChannels, 0, ns);
}

FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_kernel(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                           FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                           FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                           FFTLIB_bufParams1D_t *bufParamsShift, uint32_t numPoints,
                                                           uint32_t numChannels, void *pBlock)
{
   return exec_i16(K_I16_BATCHED, pX, pW, pY, pShift, numPoints, numChannels, pBlock);
}

// $I16B/c7504/..._ci.cpp:794-840: only the NULL checks are active.
FFTLIB_STATUS FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_checkParams(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                int16_t *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                int16_t *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                uint32_t *pShift, FFTLIB_bufParams1D_t *bufParamsShift,
                                                                uint32_t numPoints, uint32_t numChannels,
                                                                void *pBlock)
{
   if (pX == nullptr || pW == nullptr || pY == nullptr || pShift == nullptr || pBlock == nullptr)
      return FFTLIB_ERR_NULL_POINTER;
   return FFTLIB_SUCCESS;
}

FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_init(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                  FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                  FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                  FFTLIB_bufParams1D_t *bufParamsShift, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const uint32_t n  = bufParamsX->dim_x >> 1;
   const uint32_t ns = bufParamsShift ? bufParamsShift->dim_x : 0u;
   return write_plan(pBlock, K_I16_SINGLE, n, n, 1, 0, ns);
}

FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_kernel(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                    FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                    FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                    FFTLIB_bufParams1D_t *bufParamsShift, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   return exec_i16(K_I16_SINGLE, pX, pW, pY, pShift, bufParamsX->dim_x >> 1, 1, pBlock);
}

// $I16/c7504/..._ci.cpp:659-699
FFTLIB_STATUS FFTLIB_fft1d_i16sc_c16sc_o16sc_checkParams(int16_t *pX, FFTLIB_bufParams1D_t *bufParamsX, int16_t *pW,
                                                         FFTLIB_bufParams1D_t *bufParamsW, int16_t *pY,
                                                         FFTLIB_bufParams1D_t *bufParamsY, uint32_t *pShift,
                                                         FFTLIB_bufParams1D_t *bufParamsShift, void *pBlock)
{
   if (pX == nullptr || pW == nullptr || pY == nullptr || pShift == nullptr || pBlock == nullptr)
      return FFTLIB_ERR_NULL_POINTER;
   if (bufParamsX->dim_x != bufParamsW->dim_x || bufParamsX->dim_x != bufParamsY->dim_x)
      return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->dim_x < 64u * 2u) return FFTLIB_ERR_INVALID_DIMENSION;
   if (bufParamsX->data_type != FFTLIB_INT16 || bufParamsW->data_type != FFTLIB_INT16 ||
       bufParamsY->data_type != FFTLIB_INT16)
      return FFTLIB_ERR_INVALID_TYPE;
   if (bufParamsShift->data_type != FFTLIB_UINT32) return FFTLIB_ERR_INVALID_TYPE;
   if (misaligned(pX, 16)) return FFTLIB_ERR_NOT_ALIGNED_PTRS_STRIDES;
   const uint32_t pow4 = 1u << ((bufParamsShift->dim_x * 2u) & 31u);
   if (pow4 * 2u != bufParamsX->dim_x && pow4 * 4u != bufParamsX->dim_x) return FFTLIB_ERR_INVALID_DIMENSION;
   return FFTLIB_SUCCESS;
}

// ===== fp32 R2C batched / single ============================================
FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                        FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                        FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                        FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                        FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                        uint32_t numPoints, uint32_t numChannels, uint8_t outmode,
                                                        void *pBlock)
{
   if (outmode > 1u) return FFTLIB_ERR_INVALID_OUTMODE;
   if (numPoints & 1u) return FFTLIB_ERR_INVALID_DIMENSION;
   return write_plan(pBlock, K_R2C_BATCHED, numPoints, numPoints / 2u, numChannels, outmode, 0);
}

FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX,
                                                          FFTLIB_bufParams1D_t *bufParamsX,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pW,
                                                          FFTLIB_bufParams1D_t *bufParamsW,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                          FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pSf,
                                                          FFTLIB_bufParams1D_t *bufParamsSf,
                                                          FFTLIB_F32 *FFTLIB_RESTRICT pY,
                                                          FFTLIB_bufParams1D_t *bufParamsY, uint32_t numPoints,
                                                          uint32_t numChannels, uint8_t outmode, void *pBlock)
{
   return exec_r2c(K_R2C_BATCHED, pX, pW, pSf, pY, numPoints, numChannels, outmode, pBlock);
}

FFTLIB_STATUS FFTLIB_fft1dBatched_i32f_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                               FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                               FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                               uint32_t numPoints, uint32_t numChannels,
                                                               uint8_t outmode, void *pBlock)
{
   return check_real_batched(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY, numPoints, 1, outmode);
}

FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                 FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                                 FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                                 FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                                 FFTLIB_bufParams1D_t *bufParamsY, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const uint32_t n = bufParamsX->dim_x;
   if (n & 1u) return FFTLIB_ERR_INVALID_DIMENSION;
   return write_plan(pBlock, K_R2C_SINGLE, n, n / 2u, 1, 1, 0);
}

FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                   FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                   FFTLIB_F32 *FFTLIB_RESTRICT pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                   void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   return exec_r2c(K_R2C_SINGLE, pX, pW, pSf, pY, bufParamsX->dim_x, 1, 1, pBlock);
}

FFTLIB_STATUS FFTLIB_fft1d_i32f_c32fc_o32fc_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                        FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                        FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                        void *pBlock)
{
   return check_single_f32(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY);
}

// ===== fp32 C2R batched / single ============================================
FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pXFFT, FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                         FFTLIB_F32 *pSf, FFTLIB_bufParams1D_t *bufParamsSf,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         uint32_t numPoints, uint32_t numChannels, uint8_t inMode,
                                                         void *pBlock)
{
   if (inMode > 1u) return FFTLIB_ERR_INVALID_VALUE;
   if (numPoints & 1u) return FFTLIB_ERR_INVALID_DIMENSION;
   return write_plan(pBlock, K_C2R_BATCHED, numPoints, numPoints / 2u, numChannels, inMode, 0);
}

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX,
                                                           FFTLIB_bufParams1D_t *bufParamsX,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pW,
                                                           FFTLIB_bufParams1D_t *bufParamsW,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                           FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pSf,
                                                           FFTLIB_bufParams1D_t *bufParamsSf,
                                                           FFTLIB_F32 *FFTLIB_RESTRICT pY,
                                                           FFTLIB_bufParams1D_t *bufParamsY, uint32_t numPoints,
                                                           uint32_t numChannels, uint8_t inMode, void *pBlock)
{
   return exec_c2r(K_C2R_BATCHED, pX, pW, pSf, pY, numPoints, numChannels, inMode, pBlock);
}

FFTLIB_STATUS FFTLIB_ifft1dBatched_i32fc_c32fc_o32f_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                                FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                                FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                                uint32_t numPoints, uint32_t numChannels,
                                                                uint8_t inMode, void *pBlock)
{
   return check_real_batched(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY, numPoints, 0, inMode);
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_init(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX, FFTLIB_F32 *pW,
                                                  FFTLIB_bufParams1D_t *bufParamsW, FFTLIB_F32 *pXFFT,
                                                  FFTLIB_bufParams1D_t *bufParamsXFFT, FFTLIB_F32 *pSf,
                                                  FFTLIB_bufParams1D_t *bufParamsSf, FFTLIB_F32 *pY,
                                                  FFTLIB_bufParams1D_t *bufParamsY, void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   const uint32_t n = bufParamsX->dim_x;
   if (n & 1u) return FFTLIB_ERR_INVALID_DIMENSION;
   return write_plan(pBlock, K_C2R_SINGLE, n, n / 2u, 1, 1, 0);
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_kernel(FFTLIB_F32 *FFTLIB_RESTRICT pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pXFFT,
                                                    FFTLIB_bufParams1D_t *bufParamsXFFT,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pSf,
                                                    FFTLIB_bufParams1D_t *bufParamsSf,
                                                    FFTLIB_F32 *FFTLIB_RESTRICT pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                    void *pBlock)
{
   if (bufParamsX == nullptr) return FFTLIB_ERR_NULL_POINTER;
   return exec_c2r(K_C2R_SINGLE, pX, pW, pSf, pY, bufParamsX->dim_x, 1, 1, pBlock);
}

FFTLIB_STATUS FFTLIB_ifft1d_i32fc_c32fc_o32f_checkParams(FFTLIB_F32 *pX, FFTLIB_bufParams1D_t *bufParamsX,
                                                         FFTLIB_F32 *pW, FFTLIB_bufParams1D_t *bufParamsW,
                                                         FFTLIB_F32 *pY, FFTLIB_bufParams1D_t *bufParamsY,
                                                         void *pBlock)
{
   return check_single_f32(pX, bufParamsX, pW, bufParamsW, pY, bufParamsY);
}

}  // extern "C"
