// Host build of the force kernels' arccos (acos_as in oxdna_device.h), for a CPU test of
// its error against FP64 libm over [-1, 1]: the same source compiled for the host.
#include "oxdna_device.h"

extern "C" void acos_as_batch(const float* x, float* theta, float* sine, int n) {
    for (int k = 0; k < n; ++k) theta[k] = oxd::acos_as(x[k], sine[k]);
}
