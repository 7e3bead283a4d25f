// The FP32 build of the x/z transforms (namespace sc::fp32): fft_kernels.hip
// with real = float.  Used by the time steps of a mixed-precision context.
#define SC_FFT_F32 1
#include "fft_kernels.hip"
