// kernels.h — argument blocks and device helpers of the oxDNA engine's kernels:
// per-step frames, cell binning, in-kernel neighbour rows, the fused force kernel,
// energy reduction, Langevin integration.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "integrator.h"
#include "oxdna_device.h"

namespace oxd {

enum { T_EXCVOL = 0, T_HBOND = 1, T_CROSS = 2, T_COAX = 3, T_DEBYE = 4, T_HBONDCROSS = 5, T_BONDED = 6 };
enum { S_FENE = 0, S_BEXC = 1, S_STACK = 2, S_NEXC = 3, S_HB = 4, S_CR = 5, S_CX = 6, S_DH = 7 };

// Flags / counters shared by the step kernels (device memory, one per context). A run
// numbers its steps k = 1, 2, ... from the state it starts from (k = 0); kernels of
// step k use the buffers of index k mod 3 (rebuild requests, cell counts) or k mod 2
// (positions, frames, cell members, step counters), so a kernel never clears or
// overwrites what another workgroup of the same launch still reads.
struct DevFlags {
    uint32_t rebuild[3];     // neighbour rows must be rebuilt at step k (index k mod 3)
    uint32_t overflow;       // bit0: a cell overflowed, bit1: a neighbour row overflowed
    uint32_t max_cell;       // largest cell occupancy seen (run steps: only when a cell overflowed)
    uint32_t max_neigh;      // largest neighbour count seen (run steps: only when a row overflowed)
    unsigned long long step; // integration steps done (kept by the host between runs)
    unsigned long long builds;
    // random-stream counter of step k (index k mod 2): every thread of a step kernel
    // reads it, so it has a cache line of its own, away from the atomics above
    alignas(128) unsigned long long stepc[2];
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

// Cell of a position in a box centred on the origin (cells of side >= r_list).
__device__ __forceinline__ uint32_t cell_index(double x, double y, double z, const double* Linv, const int* nc) {
    double u[3] = {x * Linv[0] + 0.5, y * Linv[1] + 0.5, z * Linv[2] + 0.5};
    int c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double v = u[a] - floor(u[a]);
        int k = (int)(v * nc[a]);
        c[a] = k >= nc[a] ? nc[a] - 1 : (k < 0 ? 0 : k);
    }
    return ((uint32_t)c[2] * nc[1] + c[1]) * nc[0] + c[0];
}

// Every selected nonbonded term of one pair, i half: adds the force on i to Fp and
// the torque on i to T, the per-term energies to et (split_energy slots) and returns
// the pair energy. Rc = r_j - r_i (minimum image), rsq = |Rc|^2, tp the type-pair
// index, P the start of the parameter block. MULTI: some table differs between type
// pairs, so each term's set is found through P.poff (else the base sets, at offsets
// fixed at compile time). With COUNT, cnt[2 t] counts evaluations of term t inside its
// cut-off and cnt[2 t + 1] those with a nonzero energy.
template<bool COUNT, bool MULTI>
__device__ __forceinline__ float pair_terms(const DParams& P, uint32_t mask, int tp, float rsq, V Rc,
                                            const Frame& fi, const Frame& fj, float qiqj, V& Fp, V& T,
                                            float et[8], uint32_t cnt[14]) {
    // byte offsets of this type pair's sets
    uint4 po = make_uint4(0u, 0u, 0u, 0u);
    if constexpr (MULTI) po = *reinterpret_cast<const uint4*>(&P.poff[tp][0]);
    const unsigned char* pb = reinterpret_cast<const unsigned char*>(&P);
    auto at = [&](int t, uint32_t base_off) -> const unsigned char* {
        if constexpr (MULTI) {
            const uint32_t w = t < 2 ? po.x : (t < 4 ? po.y : po.z);
            return pb + ((t & 1) ? w >> 16 : w & 0xffffu);
        } else {
            return pb + base_off;
        }
    };
    float ep = 0.f;
    if ((mask & (1u << T_EXCVOL)) && rsq < P.rcutsq[T_EXCVOL][tp]) {
        float e = ev_pair(Rc, fi, fj, *reinterpret_cast<const DExcVol*>(at(T_EXCVOL, offsetof(DParams, ev))), Fp, T);
        et[S_NEXC] += e; ep += e;
        if (COUNT) { cnt[2 * T_EXCVOL]++; cnt[2 * T_EXCVOL + 1] += e != 0.f; }
    }
    if ((mask & (1u << T_HBOND)) && rsq < P.rcutsq[T_HBOND][tp]) {
        const DBaseBase& S = *reinterpret_cast<const DBaseBase*>(at(T_HBOND, offsetof(DParams, hb)));
        BB g;
        if (bb_geom(Rc, fi, fj, S.base, g)) {
            float eps = P.hb.hb_eps[tp], rv, rd;
            f1_vd(g.rmod, S.f1, eps, eps * S.shift_factor, rv, rd);
            float e = bb_term(g, fi, fj, rv, rd, S.hb_t1, S.hb_t23, S.hb_t4, S.hb_t78,
                              false, true, Fp, T);
            et[S_HB] += e; ep += e;
            if (COUNT) cnt[2 * T_HBOND + 1] += e != 0.f;
        }
        if (COUNT) cnt[2 * T_HBOND]++;
    }
    if ((mask & (1u << T_CROSS)) && rsq < P.rcutsq[T_CROSS][tp]) {
        const DBaseBase& S = *reinterpret_cast<const DBaseBase*>(at(T_CROSS, offsetof(DParams, cs)));
        BB g;
        if (bb_geom(Rc, fi, fj, S.base, g)) {
            float rv, rd;
            f2_vd(g.rmod, S.f2, rv, rd);
            float e = bb_term(g, fi, fj, rv, rd, S.cs_t1, S.cs_t23, S.cs_t4, S.cs_t78,
                              true, S.cs_t4_enabled != 0, Fp, T);
            et[S_CR] += e; ep += e;
            if (COUNT) cnt[2 * T_CROSS + 1] += e != 0.f;
        }
        if (COUNT) cnt[2 * T_CROSS]++;
    }
    if ((mask & (1u << T_HBONDCROSS)) && rsq < P.rcutsq[T_HBONDCROSS][tp]) {
        const DBaseBase& S = *reinterpret_cast<const DBaseBase*>(at(T_HBONDCROSS, offsetof(DParams, hbx)));
        BB g;
        if (bb_geom(Rc, fi, fj, S.base, g)) {
            const float eps = P.hbx.hb_eps[tp];
            float rv1, rd1, rv2, rd2;
            f1_vd(g.rmod, S.f1, eps, eps * S.shift_factor, rv1, rd1);
            f2_vd(g.rmod, S.f2, rv2, rd2);
            const bool t4cs = S.cs_t4_enabled != 0;
            const bool on1 = bb_active(g, rv1, S.hb_t1, S.hb_t23, S.hb_t4, S.hb_t78, false, true);
            const bool on2 = bb_active(g, rv2, S.cs_t1, S.cs_t23, S.cs_t4, S.cs_t78, true, t4cs);
            float e1 = 0.f, e2 = 0.f;
            if (on1 || on2) {
                BBAng an;  // one arccos per cosine for both terms
                bb_angles(g, an);
                BBW w;  // both terms' weights, assembled once
                bbw_zero(w);
                if (on1) {
                    const BBW w1 = bb_weights(g, an, rv1, rd1, S.hb_t1, S.hb_t23, S.hb_t4, S.hb_t78, false, true);
                    e1 = w1.e;
                    bbw_add(w, w1);
                }
                if (on2) {
                    const BBW w2 = bb_weights(g, an, rv2, rd2, S.cs_t1, S.cs_t23, S.cs_t4, S.cs_t78, true, t4cs);
                    e2 = w2.e;
                    bbw_add(w, w2);
                }
                if (e1 != 0.f || e2 != 0.f) bb_assemble(g, fi, fj, w, Fp, T);
            }
            et[S_HB] += e1; et[S_CR] += e2; ep += e1 + e2;
            if (COUNT) cnt[2 * T_HBONDCROSS + 1] += (e1 != 0.f || e2 != 0.f);
        }
        if (COUNT) cnt[2 * T_HBONDCROSS]++;
    }
    if ((mask & (1u << T_COAX)) && rsq < P.rcutsq[T_COAX][tp]) {
        float e = coax_pair(Rc, fi, fj, *reinterpret_cast<const DCoax*>(at(T_COAX, offsetof(DParams, cx))), Fp, T);
        et[S_CX] += e; ep += e;
        if (COUNT) { cnt[2 * T_COAX]++; cnt[2 * T_COAX + 1] += e != 0.f; }
    }
    if ((mask & (1u << T_DEBYE)) && rsq < P.rcutsq[T_DEBYE][tp]) {
        float e = dh_pair(Rc, fi, fj, qiqj, *reinterpret_cast<const DDebye*>(at(T_DEBYE, offsetof(DParams, dh))), Fp, T);
        et[S_DH] += e; ep += e;
        if (COUNT) { cnt[2 * T_DEBYE]++; cnt[2 * T_DEBYE + 1] += e != 0.f; }
    }
    return ep;
}

struct CellArgs {
    uint32_t N;
    const double4* pos;
    double Linv[3];
    int nc[3];
    uint32_t cap;
    uint32_t* count;    // NULL: no cell list (no nonbonded term)
    uint32_t* count_prev;  // step kernels: the counts of two steps back (zeroed as they go)
    uint32_t* members;
    DevFlags* flags;
    uint32_t ncell;     // cells of one count buffer
    // fused run steps bin a particle by its position predicted before the force kick:
    // per particle, the predicted cell of a misprediction two steps back (its count in
    // count_prev is cleared too) or kNoCell
    uint32_t* rebin;
};

constexpr uint32_t kNoCell = 0xffffffffu;
constexpr uint32_t kHole = 0xffffffffu;  // a cell member slot a misprediction left empty

// Bins particle i; returns the occupancy it saw (slot + 1) for the caller's
// wave-reduced update of flags->max_cell.
__device__ __forceinline__ uint32_t bin_particle(const CellArgs& a, uint32_t i, double x, double y, double z) {
    uint32_t c = cell_index(x, y, z, a.Linv, a.nc);
    uint32_t slot = atomicAdd(&a.count[c], 1u);
    if (slot < a.cap) a.members[(size_t)c * a.cap + slot] = i;
    else atomicOr(&a.flags->overflow, 1u);
    return slot + 1;
}

// atomicMax of v over a whole wavefront: one atomic per wave instead of one per lane
// (every lane of the wave must call it; v = 0 contributes nothing).
__device__ __forceinline__ void wave_atomic_max(uint32_t* p, uint32_t v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, m));
    if ((threadIdx.x & 63u) == 0u && v) atomicMax(p, v);
}

// Workgroups go to the 8 XCDs round-robin (workgroup b runs on XCD b mod 8, each XCD
// with its own L2). This renumbering gives XCD x one contiguous range of logical
// workgroups, so particles with close indices (one duplex, one strand) share an L2.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t per = G / 8u, rem = G % 8u, x = b % 8u;
    return x * per + min(x, rem) + b / 8u;
}

// In-kernel neighbour-row construction (scan_cells in engine.hip).
struct NbrArgs {
    const double4* pos;
    double L[3], Linv[3];
    int nc[3];
    uint32_t cap;
    const uint32_t* count;
    const uint32_t* members;
    double rlist2;
    double rbb2;           // closer than this (squared): every term can reach the pair
    // frames of the particles and the base site of the base-base terms: a near pair whose
    // base sites are closer than sqrt(bb_act2) at the rebuild (its base-base radial
    // reaches it now) is class 0, other near pairs class 1 (bb_act2 <= 0: all class 1)
    const float4* fr0;
    const float4* fr1;
    const float4* fr2;
    float bb_base[3];
    float bb_act2;
    double rnear2;         // closer than this: excluded volume + Debye-Hueckel; else Debye-Hueckel only
    uint32_t max_neigh;
    double4* pos_last;
    DevFlags* flags;       // NULL: never rebuild (rows supplied by the caller)
    const uint32_t* rebuild_flag;  // this step's rebuild request (NULL: none)
    int force_rebuild;     // rebuild regardless of the request
};

// First half of step t of a run for every particle (and, for the kernels that also
// finish step t - 1, the Langevin parameters and the random-stream counter of that step).
struct StepArgs {
    uint32_t N;
    double4* bd;       // Langevin force (xyz) of the last second half-kick, lab frame
    double4* bd_tq;    // Langevin torque (xyz) of the last second half-kick, lab frame
    const double4* pos_in;  // positions after step t - 1
    double4* pos_out;       // positions of step t; its old content (step t - 2) names the
                            // cells to clear in cells.count_prev
    double4* vel;
    double4* quat;
    double4* angmom;
    const float4* force;
    const float4* torque;
    const double* charge;
    float4* fr0;       // frames of step t
    float4* fr1;
    float4* fr2;
    const double4* pos_last;
    Langevin lv;
    double L[3], Linv[3];
    double skin_half2;
    unsigned long long seed;
    DevFlags* flags;
    uint32_t* rebuild_set;                 // rebuild request of step t
    const unsigned long long* step_in;     // random-stream counter of step t - 1
    unsigned long long* step_out;          // ... of step t (step_in + 1)
    CellArgs cells;    // count: step t's cells; count_prev: step t - 2's, cleared as its particles leave
};

struct ForceArgs {
    uint32_t N;
    // compute units of the device: a grid of at most two workgroups per compute unit runs
    // in two dispatch batches, and workgroups ncu and up start in the second (in the
    // padding after N: the block stays 13 64-byte lines)
    uint32_t ncu;
    const double4* pos;    // x, y, z, type (low 32 bits of .w)
    const float4* fr0;     // a1, charge
    const float4* fr1;     // a2, type (int bits)
    const float4* fr2;     // a3, unused
    uint32_t* n_neigh;
    uint32_t* n_near;      // per row: near entries (classes 0-1, low 16 bits), of them class 0 (high 16 bits);
                           // NULL: rows without classes, all near
    uint32_t* nlist;
    const size_t* head;
    const int* bpart;      // 2 per particle: partner or -1
    const int* binfo;      // 2 per particle: bit0 role (0: this is A, 1: this is B), bits 1.. bond type
    const DParams* P;
    double L[3], Linv[3];
    uint32_t mask;
    uint32_t xwords;       // size of P's extras block in 16-byte words (in the padding after mask;
                           // nonzero: some table differs between type pairs)
    float4* force;
    float4* torque;
    float* virial;
    size_t vpitch;
    double* epart;         // per block 8 partial energies (WANT_E)
    int accumulate;
    double4* split;        // optional per-particle (fene, bexc, stack) half energies (WANT_E)
    unsigned long long* counts;  // COUNT instantiation: per term (inside cut-off, nonzero)
    NbrArgs nb;
    uint32_t* rebuild_clear;  // run steps: the request of step k + 2, cleared for step k + 1 to set
    // 1: each thread evaluates two far row entries of a particle (latency-bound grids:
    // fewer rounds); 0: one (grids that fill the GPU: fewer instructions per round)
    int far_pairs;
    // 0: forces only; 1: then the Langevin second half-kick of step k (st.step_in); 2: then
    // also the first half of step k + 1 (st), for the particle's own lane
    int integrate;
    StepArgs st;
};

}  // namespace oxd
