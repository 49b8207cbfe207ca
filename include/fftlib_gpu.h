/*
 * GPU controls of the MI355X build: the function-style selector, the HIP
 * stream and completion mode used by every _kernel/_exec entry point, and the
 * twiddle generators the reference keeps in its test drivers.
 *
 * None of these has a counterpart function in the reference library; the
 * selector extends FFTLIB_FUNCTION_STYLE (src/common/FFTLIB_types.h:216-221),
 * which in the reference only the MMA kernels' InitArgs consult.
 */
#ifndef FFTLIB_GPU_H_
#define FFTLIB_GPU_H_ 1

#include "FFTLIB_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Style recorded by every subsequent fft_c7x _init call (process wide;
 * default FFTLIB_FUNCTION_OPTIMIZED). Returns FFTLIB_ERR_INVALID_VALUE for an
 * unknown style. */
FFTLIB_STATUS FFTLIB_setFunctionStyle(FFTLIB_FUNCTION_STYLE style);
FFTLIB_FUNCTION_STYLE FFTLIB_getFunctionStyle(void);

/* HIP stream (a hipStream_t, NULL = the null stream) on which kernels are
 * enqueued. */
FFTLIB_STATUS FFTLIB_GPU_setStream(void *hipStream);
void *FFTLIB_GPU_getStream(void);

/* 1 (default): _kernel returns after the GPU finished, like the reference's
 * synchronous API. 0: device-pointer (FFTLIB_FUNCTION_GPU) calls return right
 * after enqueueing; the caller synchronizes the stream. Host-pointer calls
 * always complete before returning. */
FFTLIB_STATUS FFTLIB_GPU_setSynchronous(int32_t synchronous);

/* 1 (default): complex FFT lengths with a compile-time specialised kernel use
 * it; 0: every length runs the generic (runtime stage table) kernels. Both
 * give results within fp32 rounding of each other; the switch exists for
 * tuning and for testing the generic kernels at every length. */
FFTLIB_STATUS FFTLIB_GPU_setSpecializedKernels(int32_t enable);

/* Arithmetic of the fp32 complex FFT kernels (batched and single, forward and
 * inverse) for the lengths that have a DFT-as-GEMM matrix-core kernel:
 * N = R1 * 16 * R3 with R1 <= 16 and R3 in {1, 2, 4}, namely 32, 48, 64, 80,
 * 96, 128, 144, 160, 192, 240, 256, 288, 320, 384, 480, 512, 576, 640, 768,
 * 960 and 1024; other lengths always use FFTLIB_C2C_VARIANT_FP32.
 *   FFTLIB_C2C_VARIANT_FP32   (0, default) mixed-radix fp32, the reference's
 *                             algorithm and twiddle table;
 *   FFTLIB_C2C_VARIANT_BF16X3 (1) radix-R1 and radix-16 DFT stages as MFMA
 *                             GEMMs on bf16 hi+mid+lo operand parts (6
 *                             products), FP32 accumulation, then an fp32
 *                             radix-R3 stage;
 *   FFTLIB_C2C_VARIANT_FP16X2 (2) same with fp16 hi+lo parts (3 products);
 *   FFTLIB_C2C_VARIANT_BF16X2 (3) same with bf16 hi+lo parts (3 products).
 * The MFMA variants generate their own DFT matrices and do not read pW.
 * Process wide; returns FFTLIB_ERR_INVALID_VALUE for an unknown value. */
#define FFTLIB_C2C_VARIANT_FP32   0
#define FFTLIB_C2C_VARIANT_BF16X3 1
#define FFTLIB_C2C_VARIANT_FP16X2 2
#define FFTLIB_C2C_VARIANT_BF16X2 3
FFTLIB_STATUS FFTLIB_GPU_setC2CVariant(int32_t variant);
int32_t       FFTLIB_GPU_getC2CVariant(void);

/* Ragged batches (no counterpart in the reference, whose batched kernels take
 * one length per call: a reference caller issues one _kernel per length).
 *
 * FFTLIB_GPU_raggedCreate plans `count` complex transforms of lengths[i]
 * points (2/3/5-smooth, 16 <= lengths[i] <= 2^20) stored back to back:
 * transform i starts at complex element sum_{j<i} lengths[j] of every buffer
 * passed to FFTLIB_GPU_raggedExec. Each transform has the C2C semantics of
 * FFTLIB_fft1dBatched_i32fc_c32fc_o32fc (forward) or
 * FFTLIB_ifft1dBatched_i32fc_c32fc_o32fc (inverse) with the library's own
 * twiddle tables.
 *
 * FFTLIB_GPU_raggedExec runs the whole batch as one HIP graph launch on the
 * library stream (captured on the first call and again whenever the buffers
 * or the direction change): direction 0 forward pX -> pY, 1 inverse
 * pX -> pY, 2 forward pX -> pY then inverse pY -> pZ. Buffers are device
 * memory, 8-byte aligned; pX is not written. Calls on one handle are
 * serialised; calls must not overlap with other calls that use the same
 * handle's buffers.
 *
 * FFTLIB_GPU_raggedTune chooses the grid size (workgroup count) of every
 * multi-length kernel the plan launches (one per octave of lengths) by
 * timing whole forward directions with the caller's buffers (pY is
 * overwritten): one launch at a time, each candidate with the others' grids
 * fixed, a candidate kept when the direction is at least 1% faster; the
 * choices go to a process-wide cache that later plans with the same kernels
 * and work sizes use; FFTLIB_GPU_tuneSave / FFTLIB_GPU_tuneLoad write and read that cache as
 * text ("fftlib-amd-tune 1", then one "key item-count-bucket grid
 * microseconds" line per entry, grid 0 = the kernel's default, microseconds
 * = the tuned direction's time,
 * key = 100000 + 100 * phase + octave index;
 * phase 0 = lengths up to 16384 and the column pass of longer lengths,
 * phase 1 = the row pass of longer lengths).
 *
 * Two forms of plan (FFTLIB_GPU_raggedSetForm; FFTLIB_GPU_raggedTune times
 * both and keeps the faster unless the caller chose one):
 *   FFTLIB_RAGGED_PERSISTENT: one persistent kernel launch per
 *     direction runs every transform from a work queue; each transform
 *     longer than 4096 points is one unit of the queue, run by one workgroup
 *     as a column pass into its own scratch region and a row pass after a
 *     workgroup barrier (no unit waits for another). Setting it returns
 *     FFTLIB_ERR_NOT_IMPLEMENTED for a plan with a length whose two-pass
 *     split needs a factor above 4096 (none up to 65536 points).
 *   FFTLIB_RAGGED_GRAPH (default): one multi-length kernel launch per octave
 *     of lengths and pass, forked over several streams and joined per pass.
 * The tuning cache also holds the persistent launch's workgroup count (key
 * 200000) and the chosen form for plans of a size (key 200001, "grid" =
 * the form, bucket = total points / 1024).
 *
 * A plan owns device memory: offset and twiddle tables per length, a scratch
 * buffer of 8 bytes per point of the lengths above 4096, counters of the
 * persistent launch, and its graph. */
#define FFTLIB_RAGGED_GRAPH      0
#define FFTLIB_RAGGED_PERSISTENT 1
typedef struct FFTLIB_GPU_ragged_s *FFTLIB_GPU_raggedHandle;
FFTLIB_STATUS FFTLIB_GPU_raggedCreate(FFTLIB_GPU_raggedHandle *handle, const uint32_t *lengths, uint32_t count);
FFTLIB_STATUS FFTLIB_GPU_raggedExec(FFTLIB_GPU_raggedHandle handle, const FFTLIB_F32 *pX, FFTLIB_F32 *pY,
                                    FFTLIB_F32 *pZ, int32_t direction);
FFTLIB_STATUS FFTLIB_GPU_raggedTune(FFTLIB_GPU_raggedHandle handle, const FFTLIB_F32 *pX, FFTLIB_F32 *pY);
FFTLIB_STATUS FFTLIB_GPU_raggedDestroy(FFTLIB_GPU_raggedHandle handle);
uint64_t      FFTLIB_GPU_raggedPoints(FFTLIB_GPU_raggedHandle handle);
FFTLIB_STATUS FFTLIB_GPU_raggedSetForm(FFTLIB_GPU_raggedHandle handle, int32_t form);
int32_t       FFTLIB_GPU_raggedGetForm(FFTLIB_GPU_raggedHandle handle);
FFTLIB_STATUS FFTLIB_GPU_tuneSave(const char *path);
FFTLIB_STATUS FFTLIB_GPU_tuneLoad(const char *path);

/* Asynchronous calls (no counterpart in the reference, whose _kernel entry
 * points are synchronous; for applications that overlap FFTs with other GPU
 * work or keep their data in host memory).
 *
 * FFTLIB_GPU_kernelAsync runs the transform an fft_c7x _init call planned in
 * pBlock (any of the ten kernels, with the channel count and outmode/inMode
 * given to _init) on the HIP stream `hipStream` (a hipStream_t; NULL = the
 * null stream) and returns once it is enqueued. pX, pW, pAux and pY are the
 * buffers that kernel's _kernel takes (pAux: pShift for the int16 kernels,
 * pSf for the real-input/real-output kernels, unused for C2C), in device
 * memory with the alignment the FFTLIB_FUNCTION_GPU style requires; pX is
 * not written. The caller keeps them alive until the stream reaches the
 * work. Long complex transforms use library scratch owned per (device,
 * stream), so calls on different streams may overlap. Errors in the call's
 * parameters are reported before anything is enqueued, as _kernel reports
 * them.
 *
 * The host pipeline runs calls whose buffers are host memory through
 * `depth` (1..8) slots, each with its own non-blocking HIP stream and device
 * buffers (grow-only, freed by pipelineDestroy). FFTLIB_GPU_pipelineSubmit
 * enqueues the host-to-device copies, the transform and the device-to-host
 * copy, and returns; when all slots are busy it first waits for the oldest
 * call. Pinned buffers (hipHostMalloc) and device buffers are copied by DMA
 * from and to the caller's memory while the call runs: keep such inputs
 * unchanged until the call has completed. Pageable inputs are copied into
 * the slot's pinned staging before submit returns; a pageable output is
 * written to its pY when the slot is reused or by FFTLIB_GPU_pipelineWait.
 * pipelineWait waits for every submitted call (in submission order), so an
 * input that is an earlier call's output needs a pipelineWait in between.
 * The plan in pBlock is copied at submit time. One pipeline serves one
 * thread at a time; pipelineDestroy drops outputs that were never waited
 * for. */
typedef struct FFTLIB_GPU_Pipeline_s FFTLIB_GPU_Pipeline;
FFTLIB_STATUS FFTLIB_GPU_kernelAsync(void *pBlock, const void *pX, const void *pW, const void *pAux, void *pY,
                                     void *hipStream);
FFTLIB_STATUS FFTLIB_GPU_pipelineCreate(FFTLIB_GPU_Pipeline **pipeline, uint32_t depth);
FFTLIB_STATUS FFTLIB_GPU_pipelineSubmit(FFTLIB_GPU_Pipeline *pipeline, void *pBlock, const void *pX, const void *pW,
                                        const void *pAux, void *pY);
FFTLIB_STATUS FFTLIB_GPU_pipelineWait(FFTLIB_GPU_Pipeline *pipeline);
FFTLIB_STATUS FFTLIB_GPU_pipelineDestroy(FFTLIB_GPU_Pipeline *pipeline);

/* Allocation-free variant of FFTLIB_GPU_kernelAsync (no counterpart in the
 * reference). For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc. Same
 * arguments, checks and results as FFTLIB_GPU_kernelAsync, but every scratch
 * buffer a call needs (the intermediate copy of complex transforms longer
 * than 16384 points, the half-length spectrum of real transforms without a
 * fused kernel, and the per-length twiddle tables of the long transforms,
 * which a small kernel writes on every call) comes from one statically
 * allocated device array of FFTLIB_GPU_NOALLOC_SCRATCH_BYTES per GPU, held
 * from the library's load until the process ends; no hipMalloc, hipFree or
 * allocating library routine is called. A call that needs more scratch
 * returns FFTLIB_ERR_INVALID_DIMENSION before anything is enqueued (a
 * 65536-point complex transform takes 512 KB of it: up to 127 of them per
 * call). The split-precision C2C variants keep device tables built on
 * first use and are refused with FFTLIB_ERR_NOT_IMPLEMENTED.
 *
 * Every call on one GPU shares that GPU's array: calls whose work overlaps on
 * the device (different streams, or several host threads) race on it, so
 * order them on one stream or synchronise between them. Nothing in the array
 * needs clearing between calls: each call writes every scratch value before
 * reading it. */
#define FFTLIB_GPU_NOALLOC_SCRATCH_BYTES (64u << 20)
FFTLIB_STATUS FFTLIB_GPU_kernelAsyncNoAlloc(void *pBlock, const void *pX, const void *pW, const void *pAux, void *pY,
                                            void *hipStream);

/* Waits for all work the library enqueued on its stream. */
FFTLIB_STATUS FFTLIB_GPU_synchronize(void);

/* Text of the last HIP failure seen by the calling thread ("" if none). */
const char *FFTLIB_GPU_getLastError(void);

/* Library version string. */
const char *FFTLIB_GPU_version(void);

/* Mixed-radix fp32 C2C twiddle table of 2*numPoints floats, the layout every
 * fp32 C2C kernel reads (test/fft_c7x/FFTLIB_fft1d_i32fc_c32fc_o32fc/
 * FFTLIB_fft1d_i32fc_c32fc_o32fc_d.c:31-97, where it is a test-driver helper). */
void FFTLIB_tw_gen_c2c(FFTLIB_F32 *pW, int32_t numPoints);

/* Radix-4 int16 Q15 twiddle table of 2*numPoints int16 for the int16 kernels
 * (test/fft_c7x/FFTLIB_fft1dBatched_i16sc_c16sc_o16sc/
 * FFTLIB_fft1dBatched_i16sc_c16sc_o16sc_d.c:29-61). */
void FFTLIB_tw_gen_i16sc(int16_t *pW, uint32_t numPoints);

#ifdef __cplusplus
}
#endif

#endif
