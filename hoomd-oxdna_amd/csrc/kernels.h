// kernels.h — gfx950 kernels of the oxDNA engine: per-step frames, cell list,
// neighbour list, the fused force kernel, energy reduction, Langevin integration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oxdna_device.h"

namespace oxd {

enum { T_EXCVOL = 0, T_HBOND = 1, T_CROSS = 2, T_COAX = 3, T_DEBYE = 4, T_HBONDCROSS = 5, T_BONDED = 6 };
enum { S_FENE = 0, S_BEXC = 1, S_STACK = 2, S_NEXC = 3, S_HB = 4, S_CR = 5, S_CX = 6, S_DH = 7 };

// Flags / counters shared by the step kernels (device memory, one per context).
struct DevFlags {
    uint32_t rebuild;        // neighbour list must be rebuilt this step
    uint32_t overflow;       // bit0: a cell overflowed, bit1: a neighbour row overflowed
    uint32_t max_cell;       // largest cell occupancy seen
    uint32_t max_neigh;      // largest neighbour count seen
    unsigned long long step; // integration step counter (random stream)
    unsigned long long builds;
};

__device__ __forceinline__ int type_of(double w) { return (int)(uint32_t)(__double_as_longlong(w) & 0xffffffffull); }

__device__ __forceinline__ double min_image(double d, double L, double Linv) {
#pragma clang fp contract(off)
    double img = rint(d * Linv);
    double t = img * L;
    return d - t;
}

// Body axes of a unit quaternion (w, x, y, z): a1 = R e_x, a2 = R e_y, a3 = R e_z.
__device__ __forceinline__ void quat_axes(double w, double x, double y, double z, float4& f0, float4& f1, float4& f2) {
    f0.x = (float)(1.0 - 2.0 * (y * y + z * z)); f0.y = (float)(2.0 * (x * y + z * w)); f0.z = (float)(2.0 * (x * z - y * w));
    f1.x = (float)(2.0 * (x * y - z * w)); f1.y = (float)(1.0 - 2.0 * (x * x + z * z)); f1.z = (float)(2.0 * (y * z + x * w));
    f2.x = (float)(2.0 * (x * z + y * w)); f2.y = (float)(2.0 * (y * z - x * w)); f2.z = (float)(1.0 - 2.0 * (x * x + y * y));
}

struct ForceArgs {
    uint32_t N;
    const double4* pos;    // x, y, z, type (low 32 bits of .w)
    const float4* fr0;     // a1, charge
    const float4* fr1;     // a2, type (int bits)
    const float4* fr2;     // a3, unused
    const uint32_t* n_neigh;
    const uint32_t* nlist;
    const size_t* head;
    const int* bpart;      // 2 per particle: partner or -1
    const int* binfo;      // 2 per particle: bit0 role (0: this is A, 1: this is B), bits 1.. bond type
    const DParams* P;
    double L[3], Linv[3];
    uint32_t mask;
    float4* force;
    float4* torque;
    float* virial;
    size_t vpitch;
    double* epart;         // per block 8 partial energies (WANT_E)
    int accumulate;
    uint32_t* cell_count;  // zeroed here when non-null
    uint32_t ncell_total;
    double4* split;        // optional per-particle (fene, bexc, stack) half energies
    unsigned long long* counts;  // COUNT instantiation: per term (inside cut-off, nonzero)
};

}  // namespace oxd
