// Plan structures shared by the host C-ABI layer and the gfx950 kernels.
//
// A plan describes one complex FFT length as the sequence of in-place
// decimation-in-frequency stages FFTLIB's kernels use (all radix-5 stages,
// then radix-3, then radix-4, then at most one radix-2; reference
// src/fft_c7x/FFTLIB_fft1dBatched_i32fc_c32fc_o32fc/..._cn.c:240-260), where
// each stage's twiddle block lives in the caller's twiddle table, plus the
// digit sequence that maps a frequency index to its position in the in-place
// result (..._cn.c:409-491).
//
// For the GPU the stages are grouped into passes of one or two consecutive
// stages; a work item keeps all r1*r2 points a pass touches in registers, so
// one LDS round trip serves two stages.
#pragma once

#include <stdint.h>

namespace fftlib {

constexpr int kMaxStages = 20;  // lengths up to 2^20 need at most 20 digits

struct StageTable {
   uint32_t n;        // complex FFT length
   uint32_t nStages;  // number of DIF stages
   uint32_t pow2;     // 1 when n is a power of two (output order = bit reversal)
   uint32_t log2n;    // log2(n) when pow2
   uint32_t nDigits;  // digits of the mixed-radix reversal (when !pow2)
   uint32_t twCount;  // complex twiddles the stages read (table prefix length)
   uint32_t nPasses;  // register passes
   uint8_t  radix[kMaxStages];      // 5, 3, 4 or 2 per stage, in execution order
   uint8_t  digit[kMaxStages];      // reversal digit radices, least significant first
   uint8_t  passFirst[kMaxStages];  // first stage of each pass
   uint8_t  passLen[kMaxStages];    // stages in each pass (1 or 2)
   uint32_t twOff[kMaxStages];      // complex offset of each stage's twiddle block
   uint32_t stageM[kMaxStages];     // sub-DFT length entering each stage
};

// Builds the stage table of a 2/3/5-smooth length; returns false otherwise.
bool build_stage_table(uint32_t n, StageTable *st);

// Kernel families recorded in a plan.
enum KernelId : uint32_t {
   K_C2C_BATCHED   = 1,
   K_C2C_SINGLE    = 2,
   K_IC2C_BATCHED  = 3,
   K_IC2C_SINGLE   = 4,
   K_I16_BATCHED   = 5,
   K_I16_SINGLE    = 6,
   K_R2C_BATCHED   = 7,
   K_R2C_SINGLE    = 8,
   K_C2R_BATCHED   = 9,
   K_C2R_SINGLE    = 10,
};

// Plan stored by every _init in the caller's pBlock and read by _kernel.
// It must fit the smallest reference PBLOCK_SIZE (9 * 64 = 576 bytes).
struct Plan {
   uint32_t   magic;
   uint32_t   kernel;       // KernelId
   uint32_t   style;        // FFTLIB_FUNCTION_STYLE captured at _init
   uint32_t   numPoints;    // as the entry point defines it (real N for R2C/C2R)
   uint32_t   numChannels;
   uint32_t   mode;         // outmode (R2C) / inMode (C2R)
   uint32_t   numShifts;    // int16: entries of pShift
   uint32_t   reserved;
   StageTable st;           // complex FFT of length numPoints or numPoints/2
};

constexpr uint32_t kPlanMagic = 0x46465447u;  // "GTFF"

static_assert(sizeof(Plan) <= 576, "plan must fit the smallest PBLOCK_SIZE");

}  // namespace fftlib
