// Host-side launchers of the gfx950 FFT kernels (implemented in
// fft_kernels.hip). All pointers are device-accessible; every launcher only
// enqueues work on `stream` and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_plan.h"

namespace fftlib {

// Scaling variants of the complex inverse transform.
enum C2CMode : int {
   C2C_FORWARD          = 0,
   C2C_INVERSE_PRESCALE = 1,  // batched IFFT: conj, x (1.0f/N) up front, conj on store
   C2C_INVERSE_POSTDIV  = 2,  // single-channel IFFT: conj up front, / (float)N and conj on store
};

// Arithmetic of the batched complex FFT (FFTLIB_GPU_setC2CVariant).
enum C2CVariant : int {
   C2C_VARIANT_FP32   = 0,  // fp32 mixed-radix DIF (reference algorithm)
   C2C_VARIANT_BF16X3 = 1,  // DFT-as-GEMM on MFMA, bf16 hi+mid+lo operands, 6 products
   C2C_VARIANT_FP16X2 = 2,  // DFT-as-GEMM on MFMA, fp16 hi+lo operands, 3 products
   C2C_VARIANT_BF16X2 = 3,  // DFT-as-GEMM on MFMA, bf16 hi+lo operands, 3 products
};

// Lengths with an MFMA DFT-as-GEMM kernel (fft_mfma.hip).
bool mfma_c2c_supported(uint32_t n);

// Batched complex FFT of n points through the MFMA kernel of `variant`
// (one of the MFMA variants above); same modes as launch_c2c. The
// caller's twiddle table is not read.
hipError_t launch_c2c_mfma(const float *x, float *y, uint32_t n, uint32_t numChannels, int mode, int variant,
                           hipStream_t stream);

// Batched complex FFT of st.n points per channel, channels contiguous.
hipError_t launch_c2c(const float *x, float *y, const float *tw, const StageTable &st, uint32_t numChannels,
                      int mode, hipStream_t stream);

// Batched real-to-complex FFT: N = 2*st.n reals per channel -> N/2+1
// (outmode 0) or N (outmode 1) bins per channel.
hipError_t launch_r2c(const float *x, float *y, const float *tw, const float *sf, const StageTable &st,
                      uint32_t numChannels, uint32_t outmode, hipStream_t stream);

// Batched complex-to-real inverse FFT producing N = 2*st.n reals per channel
// from N/2+1 (inMode 0) or N (inMode 1) bins per channel. postdiv selects the
// single-channel scaling (division at the store).
hipError_t launch_c2r(const float *x, float *y, const float *tw, const float *sf, const StageTable &st,
                      uint32_t numChannels, uint32_t inMode, int postdiv, hipStream_t stream);

// Batched int16 fixed-point FFT (st.n a power of two). shift holds one entry
// per radix-4 stage.
hipError_t launch_i16(const int16_t *x, int16_t *y, const int16_t *tw, const uint32_t *shift,
                      const StageTable &st, uint32_t numChannels, hipStream_t stream);

// Real transforms longer than one workgroup's LDS run as a complex FFT of
// nc = N/2 points plus a separate elementwise pass over HBM through `z`
// (nc * numChannels complex values, the reference's pXFFT role):
//   R2C: complex FFT x -> z, then launch_r2c_split z -> y (the split of
//        launch_r2c: Y[i] = Z[i] A[i] + mirrored Z[nc-i] B[i], Nyquist bin,
//        outmode 1 conjugate half);
//   C2R: launch_c2r_presplit x -> z (Z[i] = X[i] A'[i] + V[i] B'[i]), then
//        the complex inverse FFT z -> y.
hipError_t launch_r2c_split(const float *z, float *y, const float *sf, uint32_t nc, uint32_t numChannels,
                            uint32_t outmode, hipStream_t stream);
hipError_t launch_c2r_presplit(const float *x, float *z, const float *sf, uint32_t nc, uint32_t numChannels,
                               uint32_t inMode, hipStream_t stream);

// Largest complex length the single-pass LDS kernels accept.
uint32_t max_lds_points();

// True when one transform of st.n points (with padding) fits one workgroup's
// LDS, i.e. launch_c2c / launch_r2c / launch_c2r can run it.
bool fits_lds(const StageTable &st);

// Complex FFTs longer than one workgroup's LDS on the generic kernels
// (fft_large.hip): outer radix passes over HBM into `scratch`
// (st.n * numChannels complex values), then sub-blocks finished in LDS. Same
// modes as launch_c2c.
hipError_t launch_c2c_large(const float *x, float *y, const float *tw, const StageTable &st, uint32_t numChannels,
                            int mode, float *scratch, hipStream_t stream);

// Lengths with a compile-time specialised complex kernel (fft_fixed.hip).
bool fixed_c2c_supported(uint32_t n);

// Batched complex FFT through the specialised kernel of length n (only for
// fixed_c2c_supported(n)); same semantics as launch_c2c.
hipError_t launch_c2c_fixed(const float *x, float *y, const float *tw, uint32_t n, uint32_t numChannels, int mode,
                            hipStream_t stream);

// Inner complex lengths nc (= N/2) with specialised real-transform kernels.
bool fixed_real_supported(uint32_t nc);

// Specialised counterparts of launch_r2c / launch_c2r for nc = st.n.
hipError_t launch_r2c_fixed(const float *x, float *y, const float *tw, const float *sf, uint32_t nc,
                            uint32_t numChannels, uint32_t outmode, hipStream_t stream);
hipError_t launch_c2r_fixed(const float *x, float *y, const float *tw, const float *sf, uint32_t nc,
                            uint32_t numChannels, uint32_t inMode, int postdiv, hipStream_t stream);

}  // namespace fftlib
