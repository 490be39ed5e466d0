// ensemble.hip — the batched-replica kernel's instantiations and their launcher, in a
// translation unit of their own (built, like engine.hip, without machine-level
// loop-invariant code motion: with it, the constants and invariant products of the
// integrator and of the pair terms are hoisted out of the step loop, both sets stay live
// across the whole loop, and the kernel needs more than 256 VGPRs, one wavefront per
// SIMD; without it, 243).
#include "ensemble.h"

namespace oxd {

namespace {

template<int NT, bool COUNT, bool MULTI>
hipError_t launch_nt(uint32_t n_rep, uint32_t lds, hipStream_t st, const EnsArgs& a) {
    const void* fn = reinterpret_cast<const void*>(&k_ensemble<NT, COUNT, MULTI>);
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ensemble<NT, COUNT, MULTI>), dim3(n_rep), dim3(NT), lds, st, a);
    return hipSuccess;
}

}  // namespace

hipError_t launch_ensemble(int nt, bool count, bool multi, uint32_t n_rep, uint32_t lds, hipStream_t st,
                           const EnsArgs& a) {
    if (count || multi) {
        if (nt != 512) return hipErrorInvalidValue;
        if (count && multi) return launch_nt<512, true, true>(n_rep, lds, st, a);
        if (count) return launch_nt<512, true, false>(n_rep, lds, st, a);
        return launch_nt<512, false, true>(n_rep, lds, st, a);
    }
    switch (nt) {
    case 64: return launch_nt<64, false, false>(n_rep, lds, st, a);
    case 128: return launch_nt<128, false, false>(n_rep, lds, st, a);
    case 256: return launch_nt<256, false, false>(n_rep, lds, st, a);
    case 512: return launch_nt<512, false, false>(n_rep, lds, st, a);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace oxd
