// hvc_kernels_f32.hip -- neuron integration kernels of the FP32 form.
//
// Membrane, gating and conductance state are integrated in FP32 (hardware exp2,
// reciprocal, multiplication by rounded reciprocals of constant divisors); the
// clock, spike times, noise event times and everything of delivery and STDP
// stay FP64 (hvc_kernels.hip).  The kernels are the FP64 ones of
// hvc_integrate.hip.h instantiated with real = float.
#define HVC_FP32 1
#include "hvc_integrate.hip.h"
