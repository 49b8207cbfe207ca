// Multi-length ("class") complex FFT kernels: every 2/3/5-smooth length from
// 16 to 16384, grouped by octave (2^k, 2^(k+1)] so that the lengths of a class
// share the work-item count and transforms per workgroup. The classes are
// instantiated in fft_fixed_class0.hip .. fft_fixed_class3.hip so that they
// compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftlib {
namespace fixedk {
struct ClassArgs;
struct ClassTask;
}

constexpr int kFixedClasses = 11;

// Class of length n (-1 when n has none).
int fixed_class_of(uint32_t n);

// Launches class cls over c's items; grid > 0 sets the number of workgroups
// (at most the item count), 0 = the class kernels' default.
hipError_t fixed_class_launch(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream);

// Per translation unit: launch of the classes it instantiates
// (hipErrorInvalidValue for others).
hipError_t fixed_class_launch_tu0(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream);
hipError_t fixed_class_launch_tu1(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream);
hipError_t fixed_class_launch_tu2(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream);
hipError_t fixed_class_launch_tu3(int cls, const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream);

// Persistent ragged launch (fft_ragged_class.hip): every item of c (tasks
// of the lengths up to kRaggedClassMaxN, with their wait / signal counters)
// in one launch; grid 0 = the resident workgroups.
constexpr uint32_t kRaggedClassMaxN = 4096;
hipError_t ragged_class_launch(const fixedk::ClassArgs &c, uint32_t grid, hipStream_t stream);
// Workgroups of the persistent ragged kernel resident at once on the device.
uint32_t ragged_class_resident();

// Workgroup group size (transforms per workgroup) of length n's class kernel.
uint32_t class_group_size(uint32_t n);

// Two-pass transforms longer than 16384 points: n = n1 * m2 with both factors
// class-kernel lengths. The column pass (class of n1) runs the n1-point DFTs
// of the m2 columns and applies W_n^(p*k1); the row pass (class of m2) runs
// the m2-point DFTs of the n1 rows and writes the outputs in natural order.
// The intermediate copy holds, per transform, n1 rows of s2 points (m2
// rounded up to a multiple of 16, so that every row starts on a 128-byte
// line; columns m2..s2-1 are not used).
struct TwoPassPlan {
   uint32_t     n, n1, m2, s2;
   const float *tw1, *tw2;  // twiddle tables of n1 and m2 (device)
   const float *colTw;      // W_n^e for e < 2^colBits (device)
   const float *colTwHi;    // W_n^(h 2^colBits) for h < ceil(n / 2^colBits) (device)
   uint32_t     colBits;    // 2^colBits ~ sqrt(n)
};

// Split of n (false when n <= 16384, n > 2^22 or not 2/3/5-smooth).
bool two_pass_split(uint32_t n, uint32_t *n1, uint32_t *m2);
// The same choice for any n above `above` (at least 256): the persistent
// ragged launch splits every length above kRaggedClassMaxN.
bool two_pass_split_above(uint32_t n, uint32_t above, uint32_t *n1, uint32_t *m2);

// Complex points of the intermediate copy for `count` transforms of n points
// (0 when n has no two-pass split).
size_t two_pass_scratch_points(uint32_t n, uint32_t count);

// Plan of n (> kRaggedClassMaxN) on the current device; its tables are built
// on first use and kept until the process ends.
hipError_t two_pass_plan(uint32_t n, TwoPassPlan *plan);

// Class-kernel task of one pass over `count` transforms: layout
// LAYOUT_COLUMNS (off = input offsets of the transforms, or null) or
// LAYOUT_ROWS (off = output offsets); scratchOff = complex offset of the
// transforms' region of the scratch buffer (two_pass_scratch_points).
fixedk::ClassTask two_pass_task(const TwoPassPlan &p, int layout, uint32_t count, const uint64_t *off,
                                uint64_t scratchOff);

// numChannels contiguous transforms of n points, x -> y, through `scratch`
// (two_pass_scratch_points(n, numChannels) complex values); mode as in
// launch_c2c.
hipError_t launch_c2c_two_pass(const float *x, float *y, uint32_t n, uint32_t numChannels, int mode, float *scratch,
                               hipStream_t stream);

// The same with the tables of a given plan.
hipError_t launch_c2c_two_pass_plan(const TwoPassPlan &p, const float *x, float *y, uint32_t numChannels, int mode,
                                    float *scratch, hipStream_t stream);

// Bytes of n's two-pass tables (0 when n has no two-pass split), and a plan
// whose tables a kernel enqueued on `stream` writes to device memory at dst
// (two_pass_table_bytes(n) bytes, 8-byte aligned): nothing is allocated. The
// values are those two_pass_plan computes on the host (same formulas in
// double precision).
size_t     two_pass_table_bytes(uint32_t n);
hipError_t two_pass_plan_at(uint32_t n, float *dst, hipStream_t stream, TwoPassPlan *plan);

}  // namespace fftlib
