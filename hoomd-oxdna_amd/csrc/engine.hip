// engine.hip — oxDNA force engine for gfx950: kernels, context, C ABI.
//
// Data layout in HBM (N particles, one context):
//   pos    double4[N]  x, y, z, type id in the low 32 bits of .w (host-engine layout)
//   quat   double4[N]  w, x, y, z
//   vel    double4[N], angmom double4[N] (quaternion conjugate momentum, NO_SQUISH)
//   frames float4[3][N] a1 (+charge), a2 (+type), a3 — rebuilt every step
//   force, torque float4[N] (force.w = half energy of i), virial float[6][N]
//   nlist  uint32[N][max_neigh] rows sorted by index, n_neigh uint32[N]
//   cells  uint32[ncell] counts + uint32[ncell][cell_cap] members
// One step = 3 kernels on one stream (captured as a HIP graph):
//   step1 (half kick, drift, NO_SQUISH rotation, frames, cell binning, skin test)
//   -> fused forces (neighbour rows rebuilt in-kernel when the skin test fired; every
//      selected nonbonded term + bonded in one traversal)
//   -> step2 (Langevin half kick, translational + rotational; cell counts reset).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ensemble.h"
#include "integrator.h"
#include "kernels.h"
#include "oxdna_hip.h"

using namespace oxd;

// =====================================================================================
// kernels
// =====================================================================================
namespace oxd {

__global__ void k_frames(uint32_t N, const double4* __restrict__ quat, const double* __restrict__ charge,
                         const double4* __restrict__ pos, float4* fr0, float4* fr1, float4* fr2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double4 q = quat[i];
    float4 a, b, c;
    quat_axes(q.x, q.y, q.z, q.w, a, b, c);
    a.w = charge ? (float)charge[i] : 1.0f;
    b.w = __int_as_float(type_of(pos[i].w));
    c.w = 0.f;
    fr0[i] = a; fr1[i] = b; fr2[i] = c;
}

__global__ void k_cell_assign(CellArgs a) {
    uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    uint32_t occ = 0;
    if (i < a.N) {
        double4 p = a.pos[i];
        occ = bin_particle(a, i, p.x, p.y, p.z);
    }
    wave_atomic_max(&a.flags->max_cell, occ);
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) p[k] = 0u;
}

// ------------------------------------------------------------------ neighbour rows
// Rebuilds particle i's neighbour row inside the force kernel. The TPP lanes of the
// particle lay out the members of the (up to) 27 cells around it as one list of
// candidates, take the candidates in turn (FP64 minimum-image distance, no contraction,
// bonded partners excluded), append the matches to an LDS row, and a rank sort writes
// the row to LDS and HBM (deterministic for any order of the cell members and of the
// appends). Row entries are keys: the
// neighbour index and, in bits 30-31, the terms the pair can reach while the particles
// move less than the skin: 0 and 1 every term ("near"; 0: the base-base radial reaches
// the pair at the rebuild already), 2 excluded volume and Debye-Hueckel only, 3
// Debye-Hueckel only. Sorting by key groups the classes, so a wavefront mostly runs one
// class's path instead of the union of all of them; the classes only order the work,
// every entry is evaluated for every selected term. Lanes of one
// particle are adjacent lanes of one wavefront; every thread of the block reaches the
// barriers.
constexpr uint32_t kClassShift = 30;
constexpr uint32_t kIndexMask = (1u << kClassShift) - 1u;

// 0: not a neighbour; else 1 + the pair's class (see above). bsi: i's base site (lab
// frame, relative to i).
template<int TPP>
__device__ __forceinline__ int ref_pair(const NbrArgs& nb, const double4& pi, const V& bsi, uint32_t i, uint32_t j,
                                        int b0, int b1) {
    if (j == i || (int)j == b0 || (int)j == b1) return 0;
    const double4 pj = nb.pos[j];
    double dx = min_image(pi.x - pj.x, nb.L[0], nb.Linv[0]);
    double dy = min_image(pi.y - pj.y, nb.L[1], nb.Linv[1]);
    double dz = min_image(pi.z - pj.z, nb.L[2], nb.Linv[2]);
    double r2;
    {
#pragma clang fp contract(off)
        r2 = dx * dx;
        r2 = r2 + dy * dy;
        r2 = r2 + dz * dz;
    }
    if (!(r2 < nb.rlist2)) return 0;
    if (r2 < nb.rbb2) {
        if (nb.bb_act2 > 0.f) {
            // base site of j minus that of i (bsi), r_j - r_i = -(dx, dy, dz)
            const float4 g0 = nb.fr0[j], g1 = nb.fr1[j], g2 = nb.fr2[j];
            const float b0x = nb.bb_base[0], b1x = nb.bb_base[1], b2x = nb.bb_base[2];
            const float sx = (float)(-dx) + b0x * g0.x + b1x * g1.x + b2x * g2.x - bsi.x;
            const float sy = (float)(-dy) + b0x * g0.y + b1x * g1.y + b2x * g2.y - bsi.y;
            const float sz = (float)(-dz) + b0x * g0.z + b1x * g1.z + b2x * g2.z - bsi.z;
            if (sx * sx + sy * sy + sz * sz < nb.bb_act2) return 1;
        }
        return 2;
    }
    return r2 < nb.rnear2 ? 3 : 4;
}

// The (up to) 27 cells around particle i and where each cell's members start in the
// particle's list of candidates: lane `lane` of the TPP takes CPL consecutive cells, reads
// their occupancies, and a lane prefix sum places them. c_id[q] receives cell q's index,
// c_st[q] its first candidate (c_st[q] = the candidate count for q >= the cell count, up
// to q = 27).
template<int TPP>
__device__ __forceinline__ void cell_starts(const NbrArgs& nb, const double4& pi, uint32_t lane, uint32_t* c_id,
                                            uint32_t* c_st) {
    constexpr int CPL = (27 + TPP - 1) / TPP;
    const int nx = nb.nc[0] >= 3 ? 3 : nb.nc[0], ny = nb.nc[1] >= 3 ? 3 : nb.nc[1], nz = nb.nc[2] >= 3 ? 3 : nb.nc[2];
    const int ox = nb.nc[0] >= 3 ? -1 : 0, oy = nb.nc[1] >= 3 ? -1 : 0, oz = nb.nc[2] >= 3 ? -1 : 0;
    const int ncell = nx * ny * nz;
    const uint32_t c = cell_index(pi.x, pi.y, pi.z, nb.Linv, nb.nc);
    const int cx = c % nb.nc[0], cy = (c / nb.nc[0]) % nb.nc[1], cz = c / (nb.nc[0] * nb.nc[1]);
    uint32_t cnt[CPL], sum = 0;
#pragma unroll
    for (int u = 0; u < CPL; ++u) {
        const int q = (int)lane * CPL + u;
        cnt[u] = 0;
        if (q < ncell) {
            int x = (cx + ox + q % nx + nb.nc[0]) % nb.nc[0];
            int y = (cy + oy + (q / nx) % ny + nb.nc[1]) % nb.nc[1];
            int z = (cz + oz + q / (nx * ny) + nb.nc[2]) % nb.nc[2];
            const uint32_t cc = ((uint32_t)z * nb.nc[1] + y) * nb.nc[0] + x;
            c_id[q] = cc;
            cnt[u] = min(nb.count[cc], nb.cap);
        }
        sum += cnt[u];
    }
    uint32_t v = sum;
#pragma unroll
    for (uint32_t d = 1; d < TPP; d <<= 1) {
        uint32_t y = __shfl_up(v, d, TPP);
        if (lane >= d) v += y;
    }
    uint32_t st = v - sum;
#pragma unroll
    for (int u = 0; u < CPL; ++u) {
        const int q = (int)lane * CPL + u;
        if (q <= 27) c_st[q] = st;
        st += cnt[u];
    }
}

// The neighbours of particle i among its cells' members: the TPP lanes take the
// candidates in turn (a dense cell is shared by all lanes instead of walked by one) and
// append matches to urow through an LDS counter. The order of urow depends on the
// counter; the rank sort that follows makes the row independent of it.
template<int TPP>
__device__ __forceinline__ void row_candidates(const NbrArgs& nb, const double4& pi, const V& bsi, uint32_t i, int b0,
                                               int b1,
                                               uint32_t lane, const uint32_t* c_id, const uint32_t* c_st,
                                               uint32_t* urow, uint32_t* counter) {
    const uint32_t T = c_st[27];
    for (uint32_t t = lane; t < T; t += TPP) {
        // cell of candidate t: the largest q with c_st[q] <= t (empty cells are skipped)
        uint32_t lo = 0, n = 27;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            if (n <= 1) break;
            const uint32_t h = n / 2;
            if (c_st[lo + h] <= t) lo += h;
            n -= h;
        }
        const uint32_t j = nb.members[(size_t)c_id[lo] * nb.cap + (t - c_st[lo])];
        if (j == kHole) continue;  // left by a mispredicted binning
        const int cls = ref_pair<TPP>(nb, pi, bsi, i, j, b0, b1);
        if (cls != 0) {
            const uint32_t p = atomicAdd(counter, 1u);
            if (p < nb.max_neigh) urow[p] = j | ((uint32_t)(cls - 1) << kClassShift);
        }
    }
}

// ------------------------------------------------------------------ integrator
// Pieces of the first half of step t for particle i, shared by the kernels that run it:
// step_clear zeroes, in the counts of step t - 2, the cell the particle was binned into
// then (xold: its position then, what pos_out[i] still holds), for step t + 1 to bin
// into; step_bin bins the new position x (the atomic's slot is consumed by step_finish,
// so its round trip overlaps the rotation) and runs the skin test against plast (the
// position of the last row rebuild); step_finish stores the rotation, the frames of
// step t and the cell member. Only a cell overflow touches the flags (overflow bit,
// largest occupancy for the resize): device-scope atomics on one address from every
// wavefront would serialise the launch.
__device__ __forceinline__ void step_clear(const StepArgs& a, const double4& xold) {
    if (a.cells.count_prev) a.cells.count_prev[cell_index(xold.x, xold.y, xold.z, a.cells.Linv, a.cells.nc)] = 0u;
}

struct CellSlot {
    uint32_t cell, slot;
};

__device__ __forceinline__ CellSlot step_bin(const StepArgs& a, const double4& x, const double4& plast) {
    CellSlot b{0u, 0u};
    if (!a.cells.count) return b;
    b.cell = cell_index(x.x, x.y, x.z, a.cells.Linv, a.cells.nc);
    b.slot = atomicAdd(&a.cells.count[b.cell], 1u);
    const double dx = min_image(x.x - plast.x, a.L[0], a.Linv[0]);
    const double dy = min_image(x.y - plast.y, a.L[1], a.Linv[1]);
    const double dz = min_image(x.z - plast.z, a.L[2], a.Linv[2]);
    if (dx * dx + dy * dy + dz * dz > a.skin_half2) atomicOr(a.rebuild_set, 1u);
    return b;
}

__device__ __forceinline__ void step_finish(const StepArgs& a, uint32_t i, const double4& x, const double4& q,
                                            const double4& p, CellSlot b, float charge) {
    a.quat[i] = q;
    a.angmom[i] = p;
    float4 f0, f1, f2;
    quat_axes(q.x, q.y, q.z, q.w, f0, f1, f2);
    f0.w = charge;
    f1.w = __int_as_float(type_of(x.w));
    f2.w = 0.f;
    a.fr0[i] = f0; a.fr1[i] = f1; a.fr2[i] = f2;
    if (!a.cells.count) return;
    if (b.slot < a.cells.cap) {
        a.cells.members[(size_t)b.cell * a.cells.cap + b.slot] = i;
    } else {
        atomicOr(&a.flags->overflow, 1u);
        atomicMax(&a.flags->max_cell, b.slot + 1);
    }
}

// Second half of step t - 1 merged with the first half of step t (lv_merge_kick after
// lv_merge_pre: v_pre, p_pre), with the force and torque of step t - 1: kick, drift,
// binning and skin test, NO_SQUISH rotation, frames.
__device__ __forceinline__ void merged_particle(const StepArgs& a, uint32_t i, double4 x, double4 v, const double4& q0,
                                                double4 p, const float4& f, const float4& tq, const double4& plast,
                                                float charge) {
    double4 q = q0;
    lv_merge_kick(a.lv, f.x, f.y, f.z, tq.x, tq.y, tq.z, q, v, p);
    lv_drift(a.lv, a.L, a.Linv, v, x);
    a.pos_out[i] = x;
    a.vel[i] = v;
    const CellSlot b = step_bin(a, x, plast);
    lv_rotate(a.lv, q, p);
    step_finish(a, i, x, q, p, b, charge);
}

// merged_particle for the fused run step, in two independent parts for two wavefronts:
// the translational one (the force's kick, drift, binning, skin test, cell member) and
// the rotational one (the torque's kick, NO_SQUISH rotation, quaternion, angular momentum,
// frames). The particle already holds a cell slot bp, taken during the rounds for the
// position predicted before the force kick: the slot takes the particle if the new
// position is in that cell; otherwise it becomes a hole, the particle is binned again,
// and the predicted cell goes to a.cells.rebin for its clearing two steps later.
__device__ __forceinline__ void merged_translate_pred(const StepArgs& a, uint32_t i, double4 x, double4 v,
                                                      const float4& f, const double4& plast, CellSlot bp) {
    const double h = 2.0 * a.lv.hm;
    v.x += h * f.x; v.y += h * f.y; v.z += h * f.z;
    lv_drift(a.lv, a.L, a.Linv, v, x);
    a.pos_out[i] = x;
    a.vel[i] = v;
    CellSlot b = bp;
    if (a.cells.count) {
        const uint32_t c = cell_index(x.x, x.y, x.z, a.cells.Linv, a.cells.nc);
        uint32_t re = kNoCell;
        if (c != bp.cell) {
            if (bp.slot < a.cells.cap) a.cells.members[(size_t)bp.cell * a.cells.cap + bp.slot] = kHole;
            b.cell = c;
            b.slot = atomicAdd(&a.cells.count[c], 1u);
            re = bp.cell;
        }
        a.cells.rebin[i] = re;
    }
    const double dx = min_image(x.x - plast.x, a.L[0], a.Linv[0]);
    const double dy = min_image(x.y - plast.y, a.L[1], a.Linv[1]);
    const double dz = min_image(x.z - plast.z, a.L[2], a.Linv[2]);
    if (dx * dx + dy * dy + dz * dz > a.skin_half2) atomicOr(a.rebuild_set, 1u);
    if (!a.cells.count) return;
    if (b.slot < a.cells.cap) {
        a.cells.members[(size_t)b.cell * a.cells.cap + b.slot] = i;
    } else {
        atomicOr(&a.flags->overflow, 1u);
        atomicMax(&a.flags->max_cell, b.slot + 1);
    }
}
__device__ __forceinline__ void merged_rotate(const StepArgs& a, uint32_t i, double4 q, double4 p, const float4& tq,
                                              float charge, float tbits) {
    const double h2 = 2.0 * a.lv.dt;
    const double4 T = qmul_vq(tq.x, tq.y, tq.z, q);
    p.x += h2 * T.x; p.y += h2 * T.y; p.z += h2 * T.z; p.w += h2 * T.w;
    lv_rotate(a.lv, q, p);
    a.quat[i] = q;
    a.angmom[i] = p;
    float4 f0, f1, f2;
    quat_axes(q.x, q.y, q.z, q.w, f0, f1, f2);
    f0.w = charge;
    f1.w = tbits;
    f2.w = 0.f;
    a.fr0[i] = f0; a.fr1[i] = f1; a.fr2[i] = f2;
}

// The per-particle state a run chunk starts from: what a chunk whose cells or rows
// overflowed is run again from.
struct RunState {
    double4 *pos, *vel, *quat, *angmom, *pos_last, *bd, *bd_tq;
    float4 *force, *torque;
};

// First half of step 1 of a run chunk from the stored state (step 0, the chunk's steps
// numbered from step0), with the Langevin force and torque of the stored second half;
// each particle's starting state is copied to `save` as it is read.
// A workgroup of kStep1Block threads covers kStep1Block / 2 particles: the first half of
// its threads runs their translational part (kick, drift, binning, skin test, cell member),
// the second half their rotational part (kick, NO_SQUISH rotation, frames), side by side
// (the same operations as one thread doing both, so the same results).
constexpr uint32_t kStep1Block = 256;
__global__ __launch_bounds__(kStep1Block) void k_step1(StepArgs a, RunState save, unsigned long long step0) {
    constexpr uint32_t H = kStep1Block / 2;
    const bool rot = threadIdx.x >= H;
    const uint32_t i = xcd_block() * H + (threadIdx.x % H);
    if (i == 0 && !rot) *a.step_out = step0 + 1ull;
    if (i >= a.N) return;
    if (!rot) {
        double4 x = a.pos_in[i], v = a.vel[i];
        const float4 f = a.force[i];
        const double4 bd = a.bd[i], plast = a.pos_last[i];
        save.pos[i] = x; save.vel[i] = v; save.force[i] = f; save.bd[i] = bd; save.pos_last[i] = plast;
        step_clear(a, a.pos_out[i]);
        lv_first_half_translate(a.lv, a.L, a.Linv, f.x, f.y, f.z, bd, x, v);
        a.pos_out[i] = x;
        a.vel[i] = v;
        const CellSlot b = step_bin(a, x, plast);
        if (!a.cells.count) return;
        if (b.slot < a.cells.cap) {
            a.cells.members[(size_t)b.cell * a.cells.cap + b.slot] = i;
        } else {
            atomicOr(&a.flags->overflow, 1u);
            atomicMax(&a.flags->max_cell, b.slot + 1);
        }
    } else {
        double4 q = a.quat[i], p = a.angmom[i];
        const float4 t = a.torque[i];
        const double4 bdt = a.bd_tq[i];
        save.quat[i] = q; save.angmom[i] = p; save.torque[i] = t; save.bd_tq[i] = bdt;
        lv_first_half_rotate(a.lv, t.x, t.y, t.z, bdt, q, p);
        a.quat[i] = q;
        a.angmom[i] = p;
        float4 f0, f1, f2;
        quat_axes(q.x, q.y, q.z, q.w, f0, f1, f2);
        f0.w = a.charge ? (float)a.charge[i] : 1.f;
        f1.w = __int_as_float(type_of(a.pos_in[i].w));
        f2.w = 0.f;
        a.fr0[i] = f0; a.fr1[i] = f1; a.fr2[i] = f2;
    }
}

// Second half of step t - 1 and first half of step t, merged: the kernel between two
// force launches on systems large enough to fill the GPU (smaller ones run it in the
// force kernel's epilogue).
__global__ void k_step21(StepArgs a) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i == 0) *a.step_out = *a.step_in + 1ull;
    if (i >= a.N) return;
    double4 v = a.vel[i], p = a.angmom[i];
    const double4 q = a.quat[i];
    step_clear(a, a.pos_out[i]);
    lv_merge_pre(a.lv, i, *a.step_in, a.seed, q, v, p);
    merged_particle(a, i, a.pos_in[i], v, q, p, a.force[i], a.torque[i], a.pos_last[i],
                    a.charge ? (float)a.charge[i] : 1.f);
}

// Second half of step t - 1 alone (the last step of a run chunk, large systems).
__global__ void k_step2(StepArgs a) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    double4 v = a.vel[i], p = a.angmom[i], bd, bdt;
    const float4 f = a.force[i], t = a.torque[i];
    lv_second_half(a.lv, i, *a.step_in, a.seed, f.x, f.y, f.z, t.x, t.y, t.z, a.quat[i], v, p, bd, bdt);
    a.vel[i] = v; a.angmom[i] = p; a.bd[i] = bd; a.bd_tq[i] = bdt;
}

// ------------------------------------------------------------------ fused forces
// A workgroup of 256 threads owns P consecutive particles (TPP = 256 / P lanes each in
// the per-particle phases). (1) Each particle's neighbour row, rebuilt in-kernel when the
// skin test asked for it; its leading class-0 and class-1 entries are "near" (every
// term), the rest "far" (excluded volume and Debye-Hueckel at most). (2) The workgroup's
// near entries (every particle's class-0 entries, then every particle's class-1
// entries), its bond slots and its far entries, laid end to end with each part starting
// on a wavefront boundary, are evaluated one per thread (far entries two per thread on
// grids too small to fill the GPU: fewer rounds), 256 per round, so the base-base,
// coaxial and bonded paths run on full wavefronts instead of on the few lanes of each
// particle that hold such a pair; each evaluation leaves its force, torque and energy in
// an LDS record. (3) Each particle's lanes sum its records in a fixed order (runs are
// bit-reproducible), then store, or (FUSED) integrate.
constexpr int kRecBase = 7;  // F (3), T (3), half energy

// Sum over the TPP adjacent lanes of a particle, left in every one of them: DPP moves
// within a row of 16 lanes (rotations for 16 lanes; half-row mirror and quad
// permutations for 8 and 4) where the lanes of a particle fit one row, lane shuffles
// otherwise. The order of the additions is fixed. Called with every lane of the
// wavefront active: a DPP move from an inactive lane does not deliver its value.
template<int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}
template<int TPP>
__device__ __forceinline__ float lane_sum(float x) {
    if constexpr (TPP == 16) {
        x += dpp_mov<0x128>(x);  // row_ror:8
        x += dpp_mov<0x124>(x);  // row_ror:4
        x += dpp_mov<0x122>(x);  // row_ror:2
        x += dpp_mov<0x121>(x);  // row_ror:1
    } else if constexpr (TPP == 8) {
        x += dpp_mov<0x141>(x);  // row_half_mirror: lane l of 8 takes 7 - l
        x += dpp_mov<0x4E>(x);   // quad_perm [2, 3, 0, 1]
        x += dpp_mov<0xB1>(x);   // quad_perm [1, 0, 3, 2]
    } else if constexpr (TPP == 4) {
        x += dpp_mov<0x4E>(x);
        x += dpp_mov<0xB1>(x);
    } else {
#pragma unroll
        for (uint32_t m = TPP / 2; m > 0; m >>= 1) x += __shfl_xor(x, m);
    }
    return x;
}

// Largest s in [0, P) with off[s] <= x (off ascending, off[0] = 0).
template<int P>
__device__ __forceinline__ uint32_t owner_of(const uint32_t* off, uint32_t x) {
    uint32_t lo = 0, n = P;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (n <= 1) break;
        const uint32_t h = n / 2;
        if (off[lo + h] <= x) lo += h;
        n -= h;
    }
    return lo;
}

// Reads one dword of each of the 13 64-byte lines of the kernel's explicit arguments
// (the 832-byte ForceArgs) and of the first line of its hidden arguments (grid size,
// which the kernel reads: the address comes from the implicit-argument pointer, so the
// line exists whenever this is called), all issued before one wait in the same asm
// block: the compiler's own reads of the arguments, one dependent round trip each on
// the way to the first barrier, then hit the scalar cache. Returns their XOR for a sink.
__device__ __forceinline__ uint32_t prefetch_args() {
    static_assert(sizeof(ForceArgs) > 12 * 64 && sizeof(ForceArgs) <= 13 * 64,
                  "the prefetch names the 13 lines of ForceArgs: update it with the struct");
    const uint64_t base = (uint64_t)__builtin_amdgcn_kernarg_segment_ptr();
    const uint64_t hidden = (uint64_t)__builtin_amdgcn_implicitarg_ptr();
    uint32_t r[14];
    asm volatile(
        "s_load_dword %0, %14, 0x0\n\t"
        "s_load_dword %1, %14, 0x40\n\t"
        "s_load_dword %2, %14, 0x80\n\t"
        "s_load_dword %3, %14, 0xc0\n\t"
        "s_load_dword %4, %14, 0x100\n\t"
        "s_load_dword %5, %14, 0x140\n\t"
        "s_load_dword %6, %14, 0x180\n\t"
        "s_load_dword %7, %14, 0x1c0\n\t"
        "s_load_dword %8, %14, 0x200\n\t"
        "s_load_dword %9, %14, 0x240\n\t"
        "s_load_dword %10, %14, 0x280\n\t"
        "s_load_dword %11, %14, 0x2c0\n\t"
        "s_load_dword %12, %14, 0x300\n\t"
        "s_load_dword %13, %15, 0x0\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(r[0]), "=&s"(r[1]), "=&s"(r[2]), "=&s"(r[3]), "=&s"(r[4]), "=&s"(r[5]), "=&s"(r[6]), "=&s"(r[7]), "=&s"(r[8]), "=&s"(r[9]), "=&s"(r[10]), "=&s"(r[11]), "=&s"(r[12]), "=&s"(r[13])
        : "s"(base), "s"(hidden));
    uint32_t x = 0;
#pragma unroll
    for (int k = 0; k < 14; ++k) x ^= r[k];
    return x;
}

// FUSED: a run step whose epilogue also takes the particle through the Langevin second
// half of step k and the first half of step k + 1 (a.integrate == 2); the force-free part
// of that runs during the first round, and the force and torque, consumed in the
// epilogue, are not stored. FAR2: a thread evaluates two far row entries of a particle
// (grids too small to fill the GPU: fewer rounds). Register budget: the plain kernel of
// large grids at 4 waves per SIMD (128 VGPRs, no spills, built without machine-level
// loop-invariant code motion: large systems are bound by issue rate across many waves,
// which more waves per SIMD fill better); the other forces-only, energy and counting
// kernels at 3 (168); the fused one, used only while the grid holds <= 2 waves per
// SIMD, and the single-point energy + virial one, unconstrained.
// MULTI: some term's table differs between type pairs (a.xwords != 0): each pair finds its
// parameter sets through their offsets in the block, staged with the extras in dynamic
// LDS. Otherwise the sets sit at fixed offsets of a static copy of DParams, where the
// compiler merges their reads into 64- and 128-bit LDS loads (a set address chosen at run
// time loses that: measured 3 % slower).
template<int P, bool WANT_E, bool WANT_V, bool COUNT, bool FUSED = false, bool FAR2 = false, bool MULTI = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    (FUSED || (WANT_E && WANT_V)) ? 1 : ((FAR2 || WANT_E || WANT_V || COUNT || MULTI) ? 3 : 4))))
void k_forces(ForceArgs a) {
    constexpr uint32_t TPP = 256 / P;
    // A grid of at most two workgroups per compute unit runs in two batches: blockIdx
    // below ncu first (one per compute unit), the rest on compute units that already hold
    // one, where the instruction arbiter favours the older waves. The second batch's waves
    // take a higher priority until their first round ends, so the two batches finish
    // together instead of the second ~19 % behind. (FAR2 instantiations serve the grids
    // this small; the others are left as they are.)
    const bool late = FAR2 && gridDim.x <= 2u * a.ncu && blockIdx.x >= a.ncu;
    if (late) __builtin_amdgcn_s_setprio(1);
    uint32_t arg_lines = 0;
    if constexpr (FAR2) arg_lines = prefetch_args();
    // a run step consumes no energy: its records hold force and torque only, and what
    // only the energies need (the FENE logarithm, ...) is not computed
    constexpr int NBASE = FUSED ? 6 : kRecBase;
    constexpr int NREC = NBASE + (WANT_V ? 6 : 0) + (WANT_E ? 3 : 0);
    constexpr int RV = NBASE, RS = NBASE + (WANT_V ? 6 : 0);  // virial, split energies
    __shared__ double4 s_pos[P];
    __shared__ float4 s_fr[3][P];
    __shared__ size_t s_head[P];
    __shared__ uint32_t s_nnear[P];
    __shared__ uint32_t s_cnt[P];  // rebuild launches: each particle's row length so far
    __shared__ uint32_t s_nn[P];   // row lengths
    __shared__ uint32_t s_off[3][P + 1];  // exclusive prefix sums of class-0 near / class-1 near / far counts
    // FAR2: two rounds per chunk between barriers (their records side by side)
    constexpr uint32_t RPC = FAR2 ? 2u : 1u, CH = 256u * RPC;
    __shared__ float s_rec[NREC][CH];
    // The term parameters, staged in LDS: the inlined term evaluators reach them through
    // addresses the compiler may select per lane (one code path serving the hydrogen-
    // bonding and cross-stacking halves), which from global memory would be a vector
    // load, with its latency, per parameter. Uniform tables: DParams in s_par; MULTI: the
    // whole parameter block (DParams, then a.xwords 16-byte words of per-type-pair sets)
    // at the start of the dynamic LDS. After it, in rebuild launches, 2 rows of max_neigh
    // per particle.
    constexpr uint32_t kParWords = (uint32_t)(sizeof(DParams) / 16);
    __shared__ uint4 s_par[MULTI ? 1 : kParWords];
    extern __shared__ uint4 lds_dyn[];
    const uint32_t pwords = MULTI ? kParWords + a.xwords : 0u;
    uint32_t* lds_rows = reinterpret_cast<uint32_t*>(lds_dyn + pwords);

    const uint32_t t = threadIdx.x;
    const uint32_t slot = t / TPP, lane = t % TPP;
    const uint32_t i = xcd_block() * P + slot;
    const bool active = i < a.N;
    if constexpr (MULTI) {
        for (uint32_t w = t; w < pwords; w += 256) lds_dyn[w] = ((const uint4*)a.P)[w];
    } else {
        for (uint32_t w = t; w < kParWords; w += 256) s_par[w] = ((const uint4*)a.P)[w];
    }
    const DParams& P_ = *reinterpret_cast<const DParams*>(MULTI ? lds_dyn : s_par);
    const uint32_t mask = a.mask;
    const NbrArgs& nb = a.nb;
    constexpr uint32_t fpk = FAR2 ? 2u : 1u;  // far row entries per thread
    const bool rebuild = nb.force_rebuild || (nb.rebuild_flag && *nb.rebuild_flag);

    double4 pi = make_double4(0, 0, 0, 0);
    if (active) pi = a.pos[i];
    // the particle's stored row length, near count, row head and frames, loaded before
    // the rebuild flag is known (a rebuild discards the first three): no load waits for
    // another before the first barrier
    uint32_t ld_nn = 0, ld_near = 0;
    size_t ld_head = 0;
    float4 ld_f0 = make_float4(0, 0, 0, 0), ld_f1 = ld_f0, ld_f2 = ld_f0;
    if (active && lane == 0) {
        ld_nn = a.n_neigh[i];
        ld_near = a.n_near ? a.n_near[i] : 0xffffu;  // rows without classes: all near, class 1
        ld_head = a.head[i];
        ld_f0 = a.fr0[i]; ld_f1 = a.fr1[i]; ld_f2 = a.fr2[i];
    }
    // FUSED: the particle's integrator state, and the part of the merged half-kicks that
    // needs no force (noise, drag), while the first loads are in flight
    // FUSED: each particle's integrator state with the force-free part of the merged
    // half-kicks (noise, drag) applied, computed during the first round by a wavefront
    // whose entries there are light (the class-0 entries lead), kept in LDS for the
    // epilogue
    __shared__ double4 s_st[FUSED ? 4 : 1][FUSED ? P : 1];  // v, p, q, position of the last rebuild
    __shared__ uint32_t s_bin[2][FUSED ? P : 1];            // predicted cell and slot
    __shared__ double s_fene[WANT_E ? 2 * P : 1];           // FP64 FENE energy of each bond slot
    // in two independent parts, on two wavefronts: translation (velocity, cell clearing
    // and the predicted cell slot) and rotation (angular momentum)
    auto pre_rotate = [&](uint32_t ps) {
        const uint32_t ip = xcd_block() * P + ps;
        if (ip >= a.N) return;
        const StepArgs& s = a.st;
        double4 p = s.angmom[ip];
        const double4 q = s.quat[ip];
        lv_merge_pre_rotate(s.lv, ip, *s.step_in, s.seed, q, p);
        s_st[1][ps] = p; s_st[2][ps] = q;
    };
    auto pre_translate = [&](uint32_t ps) {
        const uint32_t ip = xcd_block() * P + ps;
        if (ip >= a.N) return;
        const StepArgs& s = a.st;
        double4 v = s.vel[ip];
        s_st[3][ps] = s.pos_last[ip];
        step_clear(s, s.pos_out[ip]);
        if (s.cells.count_prev) {
            const uint32_t rc = s.cells.rebin[ip];  // a mispredicted cell of two steps back
            if (rc < s.cells.ncell) s.cells.count_prev[rc] = 0u;
        }
        lv_merge_pre_translate(s.lv, ip, *s.step_in, s.seed, v);
        s_st[0][ps] = v;
        // the cell slot for the position the drift gives before the force kick (the kick
        // moves it by dt^2 F / m): its atomic's round trip overlaps the rounds
        uint32_t pc = 0u, pslot = 0u;
        if (s.cells.count) {
            double4 xp = a.pos[ip];
            lv_drift(s.lv, s.L, s.Linv, v, xp);
            pc = cell_index(xp.x, xp.y, xp.z, s.cells.Linv, s.cells.nc);
            pslot = atomicAdd(&s.cells.count[pc], 1u);
        }
        s_bin[0][ps] = pc; s_bin[1][ps] = pslot;
    };

    // ---- (1) rows
    uint32_t nn = 0, nnear = 0, nA = 0;  // row length, near entries, of them class 0
    const uint32_t* srow = nullptr;
    if (rebuild) {
        // neighbour row of i (uniform branch: every thread of the block is here)
        uint32_t* urow = lds_rows + (size_t)slot * 2 * nb.max_neigh;
        uint32_t* row = urow + nb.max_neigh;
        // the cell ids and candidate starts sit in `row` (at least 64 entries) until the
        // rank sort fills it
        uint32_t* c_id = row;
        uint32_t* c_st = row + 32;
        int b0 = -1, b1 = -1;
        V bsi = zero3();
        if (lane == 0) s_cnt[slot] = 0u;
        if (active) {
            b0 = a.bpart[2 * i]; b1 = a.bpart[2 * i + 1];
            if (nb.bb_act2 > 0.f) {
                const float4 f0 = a.fr0[i], f1 = a.fr1[i], f2 = a.fr2[i];
                Frame fi{mk(f0.x, f0.y, f0.z), mk(f1.x, f1.y, f1.z), mk(f2.x, f2.y, f2.z)};
                bsi = site(nb.bb_base, fi);
            }
            cell_starts<TPP>(nb, pi, lane, c_id, c_st);
        }
        __syncthreads();
        if (active) row_candidates<TPP>(nb, pi, bsi, i, b0, b1, lane, c_id, c_st, urow, &s_cnt[slot]);
        __syncthreads();
        const uint32_t total = active ? s_cnt[slot] : 0u;
        const uint32_t m = min(total, nb.max_neigh);
        uint32_t near = 0, nearA = 0;
        if (active) {
            for (uint32_t e = lane; e < m; e += TPP) {
                uint32_t x = urow[e], r = 0;
                // ties (never expected) break by position, so every rank is distinct
                for (uint32_t q = 0; q < m; ++q) r += (urow[q] < x || (urow[q] == x && q < e)) ? 1u : 0u;
                row[r] = x;
                a.nlist[a.head[i] + r] = x;
                near += (x >> kClassShift) < 2u ? 1u : 0u;
                nearA += (x >> kClassShift) == 0u ? 1u : 0u;
            }
        }
#pragma unroll
        for (uint32_t d = TPP / 2; d > 0; d >>= 1) {
            near += __shfl_xor(near, d, TPP);
            nearA += __shfl_xor(nearA, d, TPP);
        }
        if (active && lane == 0) {
            a.n_neigh[i] = m;
            if (a.n_near) a.n_near[i] = near | (nearA << 16);
            nb.pos_last[i] = pi;
            if (total > nb.max_neigh) atomicOr(&nb.flags->overflow, 2u);
            // the largest count sizes the rows: wanted from a forced rebuild (which
            // sizes them with headroom) and after an overflow; per-particle atomics
            // on one address would serialise the ordinary rebuild steps
            if (nb.force_rebuild || total > nb.max_neigh) atomicMax(&nb.flags->max_neigh, total);
            if (i == 0) atomicAdd(&nb.flags->builds, 1ull);
        }
        nn = m;
        nnear = near;
        nA = nearA;
        srow = row;
    } else if (active && lane == 0) {
        nn = ld_nn;
        nnear = min(ld_near & 0xffffu, nn);
        nA = min(ld_near >> 16, nnear);
    }
    if (lane == 0) {
        s_nnear[slot] = active ? nnear : 0u;
        s_off[0][slot + 1] = active ? nA : 0u;
        s_off[1][slot + 1] = active ? nnear - nA : 0u;
        s_off[2][slot + 1] = active ? (nn - nnear + fpk - 1u) / fpk : 0u;  // far slots
        s_nn[slot] = active ? nn : 0u;
        if (active) {
            s_pos[slot] = pi;
            s_fr[0][slot] = ld_f0; s_fr[1][slot] = ld_f1; s_fr[2][slot] = ld_f2;
            s_head[slot] = rebuild ? (size_t)slot * 2 * nb.max_neigh + nb.max_neigh : ld_head;
        }
    }
    __syncthreads();
    if constexpr (FAR2) asm volatile("" ::"s"(arg_lines));
    if constexpr (P <= 16) {
        // exclusive prefix sums over the P particles: lane c P + q of the first wavefront
        // holds particle q's count c, and the three groups of P lanes scan by shuffles
        if (t < 3u * P) {
            const uint32_t c = t / P, q = t % P;
            uint32_t v = s_off[c][q + 1];
#pragma unroll
            for (uint32_t d = 1; d < (uint32_t)P; d <<= 1) {
                const uint32_t u = (uint32_t)__shfl_up((int)v, d, 64);
                if (q >= d) v += u;
            }
            s_off[c][q + 1] = v;
            if (q == 0) s_off[c][0] = 0u;
        }
    } else if (t < 3) {  // exclusive prefix sums over the P particles, one thread per count
        uint32_t v[P];
#pragma unroll
        for (int q = 0; q < P; ++q) v[q] = s_off[t][q + 1];
        uint32_t sum = 0;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            sum += v[q];
            s_off[t][q + 1] = sum;
        }
        s_off[t][0] = 0u;
    }
    __syncthreads();
    // near entries: every particle's class-0 entries, then every particle's class-1
    // entries (the base-base path runs on the wavefronts of the first part)
    const uint32_t NA = s_off[0][P], NN = NA + s_off[1][P], NF = s_off[2][P];
    const uint32_t NB = (mask & (1u << T_BONDED)) ? ((2u * P + 63u) & ~63u) : 0u;
    const uint32_t B0 = (NN + 63u) & ~63u, F0 = B0 + NB, NT = F0 + NF;

    // ---- (2) entries, 256 per round (FAR2: two rounds per barrier); (3) each particle's
    //      lanes sum its records
    float et[8];
    double et_fene = 0.0;  // the FENE energy stays FP64 into the block reduction
    uint32_t cnt[14];
#pragma unroll
    for (int k = 0; k < 8; ++k) et[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 14; ++k) cnt[k] = 0u;
    float acc[NREC];
#pragma unroll
    for (int k = 0; k < NREC; ++k) acc[k] = 0.f;
    const int nt = P_.ntypes;
    // Row entry of thread t in round kk of the chunk at c0. FAR2 chunks hold two rounds
    // with no barrier between them: a wavefront that finishes its first-round entries
    // goes on to its second-round ones while the others still run. In the second round a
    // wavefront takes the entry slot one below its own (wavefront 0, whose first-round
    // entries, the class-0 ones, take longest, the last slot, usually the emptiest; the
    // bond slots go to wavefront 1), so the rounds' chains even out across wavefronts.
    auto entry_of = [&](uint32_t c0, uint32_t kk) -> uint32_t {
        const uint32_t w = t / 64u;
        const uint32_t sw = FAR2 ? ((w + 4u - kk) & 3u) : w;
        return c0 + kk * 256u + sw * 64u + (t & 63u);
    };
    // this thread's bond slot in the round that starts the bond slots: its partner and
    // role loaded before the rounds begin (the bond wavefront's first dependent load)
    const uint32_t rb0 = B0 & ~255u, kb = (rb0 / 256u) % RPC;
    const uint32_t eb = entry_of(rb0 - kb * 256u, kb);
    const bool bpre = NB && eb >= B0 && eb < B0 + 2u * P;
    int b_other = -1, b_info = 0;
    if (bpre) {
        const uint32_t ib = xcd_block() * P + ((eb - B0) >> 1);
        if (ib < a.N) {
            b_other = a.bpart[2 * ib + ((eb - B0) & 1u)];
            b_info = a.binfo[2 * ib + ((eb - B0) & 1u)];
        }
    }
    // the wavefronts after the class-0 entries do the integrator's force-free part,
    // rotation and translation (FAR2: not the first of them, which takes the bond slots in
    // the second round)
    const uint32_t w_rot = min(3u, (NA + 63u) / 64u + (FAR2 ? 1u : 0u));
    const uint32_t w_tra = w_rot < 3u ? w_rot + 1u : 2u;
    if (FUSED && NT == 0) {
        if (t < P) { pre_rotate(t); pre_translate(t); }
        __syncthreads();
    }
    for (uint32_t c0 = 0; c0 < NT; c0 += CH) {
#pragma unroll
      for (uint32_t kk = 0; kk < RPC; ++kk) {
        const uint32_t r0 = c0 + kk * 256u;  // this round's first entry
        if (r0 >= NT) break;
        const uint32_t e = entry_of(c0, kk);
        float rec[NREC];
#pragma unroll
        for (int k = 0; k < NREC; ++k) rec[k] = 0.f;
        uint32_t npair = 0;  // row entries this thread evaluates: one near, or up to fpk far
        uint32_t s = 0, k = 0;
        if (e < NA) {
            s = owner_of<P>(s_off[0], e);
            k = e - s_off[0][s];
            npair = 1;
        } else if (e < NN) {
            const uint32_t x = e - NA;
            s = owner_of<P>(s_off[1], x);
            k = (s_off[0][s + 1] - s_off[0][s]) + (x - s_off[1][s]);
            npair = 1;
        } else if (e >= F0 && e < NT) {
            s = owner_of<P>(s_off[2], e - F0);
            k = s_nnear[s] + fpk * (e - F0 - s_off[2][s]);
            npair = (fpk == 2u && k + 1u < s_nn[s]) ? 2u : 1u;
        }
        // one row entry's evaluation into this thread's record (add: onto what it holds)
        auto entry = [&](uint32_t kk, bool add) {
            const uint32_t j = (srow ? lds_rows[s_head[s] + kk] : a.nlist[s_head[s] + kk]) & kIndexMask;
            const double4 ps = s_pos[s];
            const float4 f0 = s_fr[0][s], f1 = s_fr[1][s], f2 = s_fr[2][s];
            const double4 pj = a.pos[j];
            const float4 g0 = a.fr0[j], g1 = a.fr1[j], g2 = a.fr2[j];
            Frame fi{mk(f0.x, f0.y, f0.z), mk(f1.x, f1.y, f1.z), mk(f2.x, f2.y, f2.z)};
            Frame fj{mk(g0.x, g0.y, g0.z), mk(g1.x, g1.y, g1.z), mk(g2.x, g2.y, g2.z)};
            const int tp = __float_as_int(f1.w) * nt + __float_as_int(g1.w);
            V dr = mk((float)min_image(ps.x - pj.x, a.L[0], a.Linv[0]),
                      (float)min_image(ps.y - pj.y, a.L[1], a.Linv[1]),
                      (float)min_image(ps.z - pj.z, a.L[2], a.Linv[2]));
            const float rsq = dot(dr, dr);
            V Fp = zero3(), Tp = zero3();
            const float ep = pair_terms<COUNT, MULTI>(P_, mask, tp, rsq, -dr, fi, fj, f0.w * g0.w, Fp, Tp, et, cnt);
            float x[NREC];
#pragma unroll
            for (int q = 0; q < NREC; ++q) x[q] = 0.f;
            x[0] = Fp.x; x[1] = Fp.y; x[2] = Fp.z;
            x[3] = Tp.x; x[4] = Tp.y; x[5] = Tp.z;
            if constexpr (NBASE > 6) x[6] = 0.5f * ep;
            if (WANT_V) {
                x[RV + 0] = 0.5f * dr.x * Fp.x; x[RV + 1] = 0.5f * dr.x * Fp.y; x[RV + 2] = 0.5f * dr.x * Fp.z;
                x[RV + 3] = 0.5f * dr.y * Fp.y; x[RV + 4] = 0.5f * dr.y * Fp.z; x[RV + 5] = 0.5f * dr.z * Fp.z;
            }
#pragma unroll
            for (int q = 0; q < NREC; ++q) rec[q] = add ? rec[q] + x[q] : x[q];
        };
        if constexpr (FAR2) {
            // up to two far entries of one particle
            for (uint32_t h = 0; h < npair; ++h) entry(k + h, h > 0);
        } else {
            if (npair) entry(k, false);
        }
        if (npair == 0 && e >= B0 && e < B0 + 2u * P && NB) {
            // bond slot (e - B0) & 1 of particle (e - B0) / 2
            s = (e - B0) >> 1;
            const uint32_t ib = xcd_block() * P + s;
            const bool pre = bpre && r0 == rb0;
            const int other = pre ? b_other : (ib < a.N ? a.bpart[2 * ib + ((e - B0) & 1u)] : -1);
            if (WANT_E && other < 0) s_fene[e - B0] = 0.0;
            if (other >= 0) {
                const int info = pre ? b_info : a.binfo[2 * ib + ((e - B0) & 1u)];
                const int role = info & 1;
                const int bt = info >> 1;
                const uint32_t A = role == 0 ? ib : (uint32_t)other;
                const uint32_t B = role == 0 ? (uint32_t)other : ib;
                const double4 pA = a.pos[A], pB = a.pos[B];
                const float4 x0 = a.fr0[A], x1 = a.fr1[A], x2 = a.fr2[A];
                const float4 y0 = a.fr0[B], y1 = a.fr1[B], y2 = a.fr2[B];
                Frame fA{mk(x0.x, x0.y, x0.z), mk(x1.x, x1.y, x1.z), mk(x2.x, x2.y, x2.z)};
                Frame fB{mk(y0.x, y0.y, y0.z), mk(y1.x, y1.y, y1.z), mk(y2.x, y2.y, y2.z)};
                BondRes r;
                bond_eval(P_.bond[bt], min_image(pB.x - pA.x, a.L[0], a.Linv[0]),
                          min_image(pB.y - pA.y, a.L[1], a.Linv[1]), min_image(pB.z - pA.z, a.L[2], a.Linv[2]),
                          fA, fB, __float_as_int(x1.w), __float_as_int(y1.w), r);
                if (COUNT) { cnt[2 * T_BONDED]++; cnt[2 * T_BONDED + 1] += r.e_stack != 0.f; }
                V F = role == 0 ? -r.FB : r.FB;
                V T = role == 0 ? r.TA : r.TB;
                rec[0] = F.x; rec[1] = F.y; rec[2] = F.z;
                rec[3] = T.x; rec[4] = T.y; rec[5] = T.z;
                // per-term sums count every directed evaluation in full and are halved at
                // the end (each pair and each bond is evaluated from both of its ends)
                const float ef = (float)r.e_fene;
                et_fene += r.e_fene;
                if (WANT_E) s_fene[e - B0] = r.e_fene;
                et[S_BEXC] += r.e_bexc;
                et[S_STACK] += r.e_stack;
                if constexpr (NBASE > 6) rec[6] = 0.5f * (ef + r.e_bexc + r.e_stack);
                if (WANT_V) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) rec[RV + q] = r.vir[q];
                }
                if (WANT_E) { rec[RS + 0] = ef; rec[RS + 1] = r.e_bexc; rec[RS + 2] = r.e_stack; }
            }
        }
#pragma unroll
        for (int q = 0; q < NREC; ++q) s_rec[q][e - c0] = rec[q];
        if (FUSED && r0 == 0 && t % 64u < P) {
            if (t / 64u == w_rot) pre_rotate(t % 64u);
            else if (t / 64u == w_tra) pre_translate(t % 64u);
        }
        if (r0 == 0 && late && RPC > 1) __builtin_amdgcn_s_setprio(0);
      }
        __syncthreads();
        if (c0 == 0 && late && RPC == 1) __builtin_amdgcn_s_setprio(0);
        // this particle's records of the chunk: near entries, bond slots, far entries
        if (active) {
            const uint32_t r0 = c0, r1 = c0 + CH;
            uint32_t lo[4] = {s_off[0][slot], NA + s_off[1][slot], B0 + 2 * slot, F0 + s_off[2][slot]};
            uint32_t hi[4] = {s_off[0][slot + 1], NA + s_off[1][slot + 1], NB ? B0 + 2 * slot + 2 : B0 + 2 * slot,
                              F0 + s_off[2][slot + 1]};
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const uint32_t x0 = max(lo[g], r0), x1 = min(hi[g], r1);
                for (uint32_t x = x0 + lane; x < x1; x += TPP) {
#pragma unroll
                    for (int q = 0; q < NREC; ++q) acc[q] += s_rec[q][x - r0];
                }
            }
        }
        __syncthreads();
    }
    if (COUNT) {
#pragma unroll
        for (int k = 0; k < 14; ++k)
            if (cnt[k]) atomicAdd(&a.counts[k], (unsigned long long)cnt[k]);
    }
    // reduce the TPP lanes of a particle (adjacent lanes of one wavefront)
#pragma unroll
    for (int q = 0; q < NREC; ++q) acc[q] = lane_sum<TPP>(acc[q]);
    if (WANT_E && a.split && active && lane == 0) {
        const double fene = (a.mask & (1u << T_BONDED)) ? s_fene[2 * slot] + s_fene[2 * slot + 1] : 0.0;
        a.split[i] = make_double4(0.5 * fene, 0.5 * acc[RS + 1], 0.5 * acc[RS + 2], 0.0);
    }
    if constexpr (FUSED) {
        // Langevin second half of step k with the force just computed, merged with the
        // first half of step k + 1: both touch only particle i's own state, and the
        // positions and frames of step k + 1 go to the other buffer of each pair, so no
        // workgroup still reading step k sees them. The block's particles take these FP64
        // sequences through LDS, lane p for particle p: the translational parts on the first
        // wavefront and the rotational parts on the second, side by side (on their group
        // leaders they would be issued, one after the other, by every wavefront).
        if (active && lane == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) s_rec[q][slot] = acc[q];
        }
        __syncthreads();
        const uint32_t pe = t & 63u;  // the particle of this lane of the first two wavefronts
        if (t < 128u && pe < P && xcd_block() * P + pe < a.N) {
            const uint32_t ip = xcd_block() * P + pe;
            if (t < 64u) {
                const float4 fo = make_float4(s_rec[0][pe], s_rec[1][pe], s_rec[2][pe], 0.f);
                const double4 x = s_pos[pe];
                merged_translate_pred(a.st, ip, x, s_st[0][pe], fo, rebuild ? x : s_st[3][pe],
                                      CellSlot{s_bin[0][pe], s_bin[1][pe]});
            } else {
                const float4 to = make_float4(s_rec[3][pe], s_rec[4][pe], s_rec[5][pe], 0.f);
                merged_rotate(a.st, ip, s_st[2][pe], s_st[1][pe], to, s_fr[0][pe].w,
                              __int_as_float(type_of(s_pos[pe].w)));
            }
        }
    } else if (active && lane == 0) {
        float ehalf = 0.f;
        if constexpr (NBASE > 6) ehalf = acc[6];
        float4 fo = make_float4(acc[0], acc[1], acc[2], ehalf);
        float4 to = make_float4(acc[3], acc[4], acc[5], 0.f);
        if (a.accumulate) {
            float4 f = a.force[i], tq = a.torque[i];
            fo.x += f.x; fo.y += f.y; fo.z += f.z; fo.w += f.w;
            to.x += tq.x; to.y += tq.y; to.z += tq.z;
        }
        a.force[i] = fo;
        a.torque[i] = to;
        if (WANT_V) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                float v = acc[RV + q];
                if (a.accumulate) v += a.virial[q * a.vpitch + i];
                a.virial[q * a.vpitch + i] = v;
            }
        }
        if (a.integrate) {
            // the Langevin second half of step k alone (the last step of a run)
            const StepArgs& s = a.st;
            double4 v = s.vel[i], p = s.angmom[i], bd, bdt;
            lv_second_half(s.lv, i, *s.step_in, s.seed, fo.x, fo.y, fo.z, to.x, to.y, to.z, s.quat[i], v, p,
                           bd, bdt);
            s.vel[i] = v; s.angmom[i] = p; s.bd[i] = bd; s.bd_tq[i] = bdt;
        }
    }
    if (FUSED && t == 0 && xcd_block() == 0) *a.st.step_out = *a.st.step_in + 1ull;
    if (a.rebuild_clear && t == 0 && xcd_block() == 0) *a.rebuild_clear = 0u;
    if (WANT_E) {
        // per-term totals: FP64 from here on (wave shuffle, then across waves in LDS)
        __shared__ double red[256 / 64][8];
        const int wid = t / 64, wl = t % 64;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double v = q == S_FENE ? et_fene : (double)et[q];
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
            if (wl == 0) red[wid][q] = v;
        }
        __syncthreads();
        if (t < 8) {
            double v = 0.0;
            for (int w = 0; w < (int)(blockDim.x / 64); ++w) v += red[w][t];
            a.epart[(size_t)blockIdx.x * 8 + t] = 0.5 * v;  // full list: each pair seen twice
        }
    }
}

// Sums per-block partial energies into out[8] (one block; fixed order).
__global__ void k_reduce_energy(uint32_t nblocks, const double* __restrict__ part, double* out, int accumulate) {
    __shared__ double s[256];
    for (int slot = 0; slot < 8; ++slot) {
        double v = 0.0;
        for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) v += part[(size_t)b * 8 + slot];
        s[threadIdx.x] = v;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[slot] = (accumulate ? out[slot] : 0.0) + s[0];
        __syncthreads();
    }
}

// Copies the N particle states `from` into `to` (a run chunk's restore).
__global__ void k_copy_state(uint32_t N, RunState from, RunState to) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = g; k < N; k += stride) {
        to.pos[k] = from.pos[k]; to.vel[k] = from.vel[k]; to.quat[k] = from.quat[k];
        to.angmom[k] = from.angmom[k]; to.pos_last[k] = from.pos_last[k];
        to.bd[k] = from.bd[k]; to.bd_tq[k] = from.bd_tq[k];
        to.force[k] = from.force[k]; to.torque[k] = from.torque[k];
    }
}

// What the host reads after a run chunk (pinned host memory, written by k_chunk_end).
struct RunStatus {
    uint32_t overflow, max_cell, max_neigh, done;
    unsigned long long builds;
};

// End of a run chunk: its state is in buffer `odd` of the position and frame pairs;
// bring it to buffer 0 (where every other entry point reads it), record the step count,
// report the overflow record to the host, and leave what the next chunk's step 1
// expects: the ncount cell counts zero, no rebuild request or overflow pending.
__global__ void k_chunk_end(uint32_t N, int odd, double4* pos, float4* fr0, float4* fr1, float4* fr2,
                            uint32_t* counts, size_t ncount, DevFlags* f, unsigned long long step_end,
                            RunStatus* out) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    if (odd) {
        for (size_t k = g; k < N; k += stride) {
            pos[k] = pos[N + k];
            fr0[k] = fr0[N + k]; fr1[k] = fr1[N + k]; fr2[k] = fr2[N + k];
        }
    }
    for (size_t k = g; k < ncount; k += stride) counts[k] = 0u;
    if (g == 0) {
        f->step = step_end;
        out->overflow = f->overflow;
        out->max_cell = f->max_cell;
        out->max_neigh = f->max_neigh;
        out->builds = f->builds;
        out->done = 1u;
        f->rebuild[0] = f->rebuild[1] = f->rebuild[2] = 0u;
        f->overflow = 0u; f->max_cell = 0u; f->max_neigh = 0u;
    }
}

}  // namespace oxd

// =====================================================================================
// context
// =====================================================================================
namespace {

thread_local std::string g_error;

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_OK(call)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) throw Fail(std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

template<class T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        HIP_OK(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

constexpr int kBlock = 256;
constexpr int kTPPMax = 32;
// LDS of the rebuild rows; with the force kernel's own LDS (records, particle data,
// parameters: up to 26 KiB) a workgroup stays within 64 KiB
constexpr size_t kMaxLds = 36 * 1024;
constexpr uint32_t kMaxNeigh = 768;  // 8 particles x 2 rows x 768 x 4 B = 48 KiB of LDS at 32 lanes

}  // namespace

struct oxdna_ctx {
    int device = 0;
    uint32_t ncu = 0;  // compute units of the device
    uint32_t lds_block = 65536, lds_cu = 65536;  // LDS bytes per workgroup, per compute unit
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;

    uint32_t N = 0;
    double L[3] = {0, 0, 0}, Linv[3] = {0, 0, 0};
    DBuf<double4> pos, vel, quat, angmom, pos_last, bd, bd_tq;
    DBuf<double> charge;
    DBuf<float4> fr0, fr1, fr2, force, torque;
    DBuf<uint32_t> rebin;  // 2 per particle (step parity): see CellArgs::rebin
    DBuf<float> virial;
    DBuf<int> bpart, binfo;
    DBuf<uint32_t> n_neigh, n_near, nlist, cell_count, cell_members;
    DBuf<size_t> head;
    DBuf<double> epart, energy;
    DBuf<DevFlags> flags;
    DBuf<uint4> dparams;                 // DParams, then the extras block (16-byte words)

    std::vector<int32_t> bonds_h, btype_h;
    bool have_bonds = false;

    DParams P{};
    std::vector<unsigned char> xset[6];  // per term: its distinct parameter sets beyond the base set
    uint8_t pidx[6][16] = {};            // per term, per type pair: its set index (0: base set)
    bool pair_set[6] = {false, false, false, false, false, false};
    double rcut_max[6] = {0, 0, 0, 0, 0, 0};
    bool bond_set = false;
    bool params_dirty = true;

    double buffer = 0.4;
    double rlist = 0.0;
    int nc[3] = {1, 1, 1};
    uint32_t cell_cap = 0;
    uint32_t max_neigh = 0;
    bool nlist_valid = false;

    bool have_langevin = false;
    double dt = 0.003, kT = 0.1, gamma = 1.0, gamma_r = 1.0;
    uint64_t seed = 0;

    int tpp = 16;                        // force-kernel lanes per particle (autotuned)
    uint32_t tuned_mask = 0xffffffffu;
    uint32_t tuned_n = 0;
    int fixed_tpp = 0;                   // nonzero: lanes per particle pinned by the caller
    uint32_t forces_mask = 0xffffffffu;  // mask of the forces currently in `force`
    hipGraphExec_t graph = nullptr;      // kGraphSteps inner steps
    uint32_t graph_mask = 0;
    uint64_t host_step = 0;
    // the state a run chunk starts from (restored when the chunk must run again)
    DBuf<double4> sv_pos, sv_vel, sv_quat, sv_angmom, sv_pos_last, sv_bd, sv_bd_tq;
    DBuf<float4> sv_force, sv_torque;
    RunStatus* status_h = nullptr;       // pinned host memory written by k_chunk_end
    RunStatus* status_d = nullptr;       // its device address
    // oxdna_time_steps: HIP events around a chunk's graph replays
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double timed_ms = 0.0;
    uint64_t timed_steps = 0;

    // batched replicas (n_rep > 0): particles [rep_off[r], rep_off[r+1]) form replica r
    uint32_t n_rep = 0;
    uint32_t rep_nmax = 0;
    std::vector<uint32_t> rep_off_h;
    std::vector<double> rep_box_h;
    DBuf<uint32_t> rep_off, rep_order;
    DBuf<double> rep_box, rep_energy;
    DBuf<EnsBias> rep_bias;
    bool have_bias = false;
    int ens_nt = 0;                      // workgroup size of the replica kernel (autotuned)
    uint32_t ens_div = 1;                // its tile depth: the LDS share's record capacity / ens_div (autotuned)
    int fixed_ens_nt = 0;                // nonzero: replica workgroup size pinned by the caller ...
    uint32_t fixed_ens_div = 1;          // ... with this tile-depth divisor
    uint32_t ens_tuned_mask = 0xffffffffu;

    ~oxdna_ctx() {
        if (graph) (void)hipGraphExecDestroy(graph);
        sv_pos.release(); sv_vel.release(); sv_quat.release(); sv_angmom.release(); sv_pos_last.release();
        sv_bd.release(); sv_bd_tq.release(); sv_force.release(); sv_torque.release();
        if (status_h) (void)hipHostFree(status_h);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        pos.release(); vel.release(); quat.release(); angmom.release(); pos_last.release();
        bd.release(); bd_tq.release();
        charge.release(); fr0.release(); fr1.release(); fr2.release(); force.release(); torque.release();
        rebin.release();
        virial.release(); bpart.release(); binfo.release(); n_neigh.release(); n_near.release(); nlist.release();
        cell_count.release(); cell_members.release(); head.release(); epart.release(); energy.release();
        flags.release(); dparams.release();
        rep_off.release(); rep_order.release(); rep_box.release(); rep_energy.release(); rep_bias.release();
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }

    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
};

namespace {

int fail(oxdna_ctx* ctx, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_error = msg;
    return -1;
}

template<class F> int guard(oxdna_ctx* ctx, F&& body) {
    try {
        if (ctx) {
            ctx->err.clear();
            HIP_OK(hipSetDevice(ctx->device));
        }
        body();
        return 0;
    } catch (const std::exception& e) {
        return fail(ctx, e.what());
    }
}

uint32_t blocks_for(size_t threads) { return (uint32_t)((threads + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ parameter conversion
void cvt3(float d[3], const double s[3]) { d[0] = (float)s[0]; d[1] = (float)s[1]; d[2] = (float)s[2]; }
// f4 set (a, b, theta0, theta_s, theta_c) plus the open interval of cos(theta) on
// which f4 is nonzero: |theta - theta0| < theta_c  <=>  cos(theta0 + theta_c) < c <
// cos(theta0 - theta_c), either side unbounded where the angle range hits 0 or pi.
F4P cvt_f4(const double s[5]) {
    const double lo_th = s[2] - s[4], hi_th = s[2] + s[4];
    const float clo = hi_th >= M_PI ? -2.0f : (float)std::cos(hi_th);
    const float chi = lo_th <= 0.0 ? 2.0f : (float)std::cos(lo_th);
    return F4P{(float)s[0], (float)s[1], (float)s[2], (float)s[3], (float)s[4], clo, chi};
}
F5P cvt_f5(const double s[4]) { return F5P{(float)s[0], (float)s[1], (float)s[2], (float)s[3]}; }

void set_f1(F1P& d, double a, double r0, double rlow, double rhigh, double rclow, double rchigh, double blow,
            double bhigh) {
    d = F1P{(float)a, (float)r0, (float)rlow, (float)rhigh, (float)rclow, (float)rchigh, (float)blow, (float)bhigh};
}
void set_f2(F2P& d, double k, double r0, double rc, double rlow, double rhigh, double rclow, double rchigh,
            double blow, double bhigh) {
    d = F2P{(float)k, (float)r0, (float)rc, (float)rlow, (float)rhigh, (float)rclow, (float)rchigh, (float)blow,
            (float)bhigh};
}

// One type pair's parameter struct (C ABI, FP64) as the kernels' FP32 set. The
// hydrogen-bonding epsilon is left out of the base-base sets (it has a per-type-pair
// table of its own).
DExcVol cvt_set(const oxdna_excvol_params& t) {
    DExcVol d;
    std::memset(&d, 0, sizeof(d));
    d.eps = (float)t.eps;
    cvt3(d.base, t.base_site); cvt3(d.back, t.back_site);
    for (int k = 0; k < 4; ++k) {
        d.sigma[k] = (float)t.sigma[k]; d.rstar[k] = (float)t.rstar[k];
        d.b[k] = (float)t.b[k]; d.rc[k] = (float)t.rc[k];
    }
    return d;
}
DBaseBase cvt_set(const oxdna_hbond_params& t) {
    DBaseBase d;
    std::memset(&d, 0, sizeof(d));
    cvt3(d.base, t.base_site);
    set_f1(d.f1, t.f1_a, t.f1_r0, t.f1_rlow, t.f1_rhigh, t.f1_rclow, t.f1_rchigh, t.f1_blow, t.f1_bhigh);
    d.shift_factor = (float)t.shift_factor;
    d.hb_t1 = cvt_f4(t.t1); d.hb_t23 = cvt_f4(t.t23); d.hb_t4 = cvt_f4(t.t4); d.hb_t78 = cvt_f4(t.t78);
    return d;
}
DBaseBase cvt_set(const oxdna_crossstacking_params& t) {
    DBaseBase d;
    std::memset(&d, 0, sizeof(d));
    cvt3(d.base, t.base_site);
    set_f2(d.f2, t.f2_k, t.f2_r0, t.f2_rc, t.f2_rlow, t.f2_rhigh, t.f2_rclow, t.f2_rchigh, t.f2_blow, t.f2_bhigh);
    d.cs_t1 = cvt_f4(t.t1); d.cs_t23 = cvt_f4(t.t23); d.cs_t4 = cvt_f4(t.t4); d.cs_t78 = cvt_f4(t.t78);
    d.cs_t4_enabled = t.t4_enabled;
    return d;
}
DBaseBase cvt_set(const oxdna_hbondcross_params& t) {
    DBaseBase d;
    std::memset(&d, 0, sizeof(d));
    cvt3(d.base, t.base_site);
    set_f1(d.f1, t.f1_a, t.f1_r0, t.f1_rlow, t.f1_rhigh, t.f1_rclow, t.f1_rchigh, t.f1_blow, t.f1_bhigh);
    d.shift_factor = (float)t.shift_factor;
    d.hb_t1 = cvt_f4(t.hb_t1); d.hb_t23 = cvt_f4(t.hb_t23); d.hb_t4 = cvt_f4(t.hb_t4); d.hb_t78 = cvt_f4(t.hb_t78);
    set_f2(d.f2, t.f2_k, t.f2_r0, t.f2_rc, t.f2_rlow, t.f2_rhigh, t.f2_rclow, t.f2_rchigh, t.f2_blow, t.f2_bhigh);
    d.cs_t1 = cvt_f4(t.cs_t1); d.cs_t23 = cvt_f4(t.cs_t23); d.cs_t4 = cvt_f4(t.cs_t4); d.cs_t78 = cvt_f4(t.cs_t78);
    d.cs_t4_enabled = t.cs_t4_enabled;
    return d;
}
DCoax cvt_set(const oxdna_coaxstacking_params& t) {
    DCoax d;
    std::memset(&d, 0, sizeof(d));
    cvt3(d.stack, t.stack_site); cvt3(d.backref, t.stack_back_ref);
    d.gamma = (float)t.gamma;
    set_f2(d.f2, t.f2_k, t.f2_r0, t.f2_rc, t.f2_rlow, t.f2_rhigh, t.f2_rclow, t.f2_rchigh, t.f2_blow, t.f2_bhigh);
    d.t1 = cvt_f4(t.t1); d.t4 = cvt_f4(t.t4); d.t56 = cvt_f4(t.t56);
    d.phi3 = cvt_f5(t.phi3); d.phi4 = cvt_f5(t.phi4);
    d.t1_mode = t.t1_mode; d.phi3_enabled = t.phi3_enabled; d.rna_coax = t.rna_coax;
    d.t1_sa = (float)t.t1_sa; d.t1_sb = (float)t.t1_sb;
    if (d.t1_mode != 0 && d.t1_mode != 1) throw Fail("coaxial stacking: t1_mode must be 0 or 1");
    // theta1's second piece: mode 0 f4(2 pi - theta1), nonzero for theta1 > 2 pi - t0 - tc
    // (theta1 <= pi); mode 1 SA (theta1 - SB)^2, nonzero for theta1 > SB
    const double t1_lo = d.t1_mode == 0 ? 2.0 * M_PI - t.t1[2] - t.t1[4] : t.t1_sb;  // (a, b, t0, ts, tc)
    d.t1_cx = t1_lo >= M_PI ? -2.0f : (t1_lo <= 0.0 ? 2.0f : (float)std::cos(t1_lo));
    return d;
}
DDebye cvt_set(const oxdna_debye_params& t) {
    DDebye d;
    std::memset(&d, 0, sizeof(d));
    cvt3(d.back, t.back_site);
    d.minus_kappa = (float)t.minus_kappa; d.prefactor = (float)t.prefactor;
    d.b_smooth = (float)t.b_smooth; d.rc = (float)t.rc; d.rhigh = (float)t.rhigh;
    return d;
}

// Distinct parameter sets of one term's ntypes x ntypes table: type pair 0's set is the
// base set, every other distinct set is appended to `extra` (bytes), and idx[tp]
// receives each type pair's set index (type pairs beyond the table: 0).
template<class S, class D>
void assign_sets(std::vector<unsigned char>& extra, uint8_t idx[16], int nt, const S* tab, D& base) {
    std::vector<D> sets;
    for (int k = 0; k < 16; ++k) {
        if (k >= nt * nt) {
            idx[k] = 0;
            continue;
        }
        const D d = cvt_set(tab[k]);
        size_t j = 0;
        while (j < sets.size() && std::memcmp(&sets[j], &d, sizeof(D)) != 0) ++j;
        if (j == sets.size()) sets.push_back(d);
        idx[k] = (uint8_t)j;
    }
    base = sets[0];
    extra.assign(reinterpret_cast<const unsigned char*>(sets.data() + 1),
                 reinterpret_cast<const unsigned char*>(sets.data() + sets.size()));
}

// Converts one term's table into P (base set and, for the hydrogen-bonding terms, the
// per-type-pair epsilon), `extra` (the term's further distinct sets, bytes; empty when
// the table is uniform) and idx (each type pair's set index).
void load_pair_params(DParams& P, std::vector<unsigned char>& extra, uint8_t idx[16], int term, int nt,
                      const void* params) {
    switch (term) {
    case OXDNA_TERM_EXCVOL:
        assign_sets(extra, idx, nt, (const oxdna_excvol_params*)params, P.ev);
        break;
    case OXDNA_TERM_HBOND: {
        auto* t = (const oxdna_hbond_params*)params;
        assign_sets(extra, idx, nt, t, P.hb);
        for (int k = 0; k < 16; ++k) P.hb.hb_eps[k] = k < nt * nt ? (float)t[k].eps : 0.f;
        break;
    }
    case OXDNA_TERM_CROSS:
        assign_sets(extra, idx, nt, (const oxdna_crossstacking_params*)params, P.cs);
        break;
    case OXDNA_TERM_HBONDCROSS: {
        auto* t = (const oxdna_hbondcross_params*)params;
        assign_sets(extra, idx, nt, t, P.hbx);
        for (int k = 0; k < 16; ++k) P.hbx.hb_eps[k] = k < nt * nt ? (float)t[k].eps : 0.f;
        break;
    }
    case OXDNA_TERM_COAX:
        assign_sets(extra, idx, nt, (const oxdna_coaxstacking_params*)params, P.cx);
        break;
    case OXDNA_TERM_DEBYE:
        assign_sets(extra, idx, nt, (const oxdna_debye_params*)params, P.dh);
        break;
    default:
        throw Fail("unknown nonbonded term id");
    }
}

// Largest extras block (every term's further distinct sets together): with the
// kernels' other LDS a workgroup stays well within the CU's LDS.
constexpr size_t kMaxExtraBytes = 16 * 1024;

// The parameter block the kernels read: DParams with the extras block right after it
// (16-byte words). Sets P.poff (every type pair's set offsets) and P.xwords from the
// terms' extras and set indices idx[term][tp].
std::vector<uint4> param_blob(DParams& P, const std::vector<unsigned char>* extra, const uint8_t (*idx)[16]) {
    static const size_t base_off[6] = {offsetof(DParams, ev), offsetof(DParams, hb), offsetof(DParams, cs),
                                       offsetof(DParams, cx), offsetof(DParams, dh), offsetof(DParams, hbx)};
    static const size_t set_size[6] = {sizeof(DExcVol), sizeof(DBaseBase), sizeof(DBaseBase),
                                       sizeof(DCoax), sizeof(DDebye), sizeof(DBaseBase)};
    size_t xoff[6], total = 0;
    for (int t = 0; t < 6; ++t) {
        xoff[t] = total;
        total += extra[t].size();
    }
    if (total > kMaxExtraBytes)
        throw Fail("per-type-pair parameter sets: " + std::to_string(total) + " bytes of distinct sets beyond the "
                   "first of each term; at most " + std::to_string(kMaxExtraBytes));
    for (int k = 0; k < 16; ++k) {
        for (int t = 0; t < 6; ++t) {
            const size_t j = idx[t][k];
            P.poff[k][t] = (uint16_t)(j == 0 ? base_off[t] : sizeof(DParams) + xoff[t] + (j - 1) * set_size[t]);
        }
        P.poff[k][6] = P.poff[k][7] = 0;
    }
    P.xwords = (uint32_t)((total + 15) / 16);
    std::vector<uint4> blob(sizeof(DParams) / 16 + P.xwords);
    std::memcpy(blob.data(), &P, sizeof(DParams));
    unsigned char* x = reinterpret_cast<unsigned char*>(blob.data() + sizeof(DParams) / 16);
    std::memset(x, 0, (size_t)P.xwords * 16);
    for (int t = 0; t < 6; ++t)
        if (!extra[t].empty()) std::memcpy(x + xoff[t], extra[t].data(), extra[t].size());
    return blob;
}

void load_bond_params(DBonded& d, const oxdna_bonded_params& s) {
    std::memset(&d, 0, sizeof(d));
    d.eps_backbone = s.eps_backbone; d.r0_backbone = s.r0_backbone;
    d.delta_backbone = s.delta_backbone; d.delta2 = s.delta2;
    cvt3(d.back, s.back_site); cvt3(d.base, s.base_site);
    d.bx.eps = (float)s.excl_eps;
    for (int k = 0; k < 3; ++k) {
        d.bx.sigma[k] = (float)s.bx_sigma[k]; d.bx.rstar[k] = (float)s.bx_rstar[k];
        d.bx.b[k] = (float)s.bx_b[k]; d.bx.rc[k] = (float)s.bx_rc[k];
    }
    d.stacking_enabled = s.stacking_enabled;
    d.rna_stacking = s.rna_stacking;
    cvt3(d.stack, s.stack_site); cvt3(d.backref, s.stack_back_ref);
    d.gamma = (float)s.st_gamma;
    set_f1(d.f1, s.st_a, s.st_r0, s.st_rlow, s.st_rhigh, s.st_rclow, s.st_rchigh, s.st_blow, s.st_bhigh);
    d.shift_factor = (float)s.st_shift_factor;
    for (int k = 0; k < 16; ++k) d.st_eps[k] = (float)s.st_eps[k];
    d.t4 = cvt_f4(s.t4); d.t5 = cvt_f4(s.t5); d.p1 = cvt_f5(s.p1);
    cvt3(d.stack3, s.stack_3_site); cvt3(d.stack5, s.stack_5_site);
    cvt3(d.bbv3, s.bbvector_3); cvt3(d.bbv5, s.bbvector_5);
    d.tB1 = cvt_f4(s.tB1); d.tB2 = cvt_f4(s.tB2);
}

double pair_rmax(const oxdna_ctx* c, uint32_t mask) {
    double r = 0.0;
    for (int t = 0; t < 6; ++t)
        if (((mask >> t) & 1u) && c->pair_set[t]) r = std::max(r, c->rcut_max[t]);
    return r;
}

uint32_t effective_mask(const oxdna_ctx* c, uint32_t mask) {
    uint32_t m = 0;
    for (int t = 0; t < 6; ++t)
        if (((mask >> t) & 1u) && c->pair_set[t]) m |= 1u << t;
    if (((mask >> OXDNA_TERM_BONDED) & 1u) && c->bond_set && c->have_bonds) m |= 1u << OXDNA_TERM_BONDED;
    return m;
}

void upload_params(oxdna_ctx* c) {
    if (!c->params_dirty) return;
    const std::vector<uint4> blob = param_blob(c->P, c->xset, c->pidx);
    c->dparams.alloc(blob.size());
    HIP_OK(hipMemcpyAsync(c->dparams.p, blob.data(), sizeof(uint4) * blob.size(), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));  // the host blob goes out of scope
    c->params_dirty = false;
}

// ------------------------------------------------------------------ force launches
// One templated launcher for every instantiation: TPP lanes per particle (8, 16, 32).
void check_force_args(const ForceArgs& a);
template<bool E, bool V, bool CNT, bool FUSED = false, bool FAR2 = false>
void launch_k(int tpp, dim3 g, dim3 b, size_t rows_lds, hipStream_t st, const ForceArgs& a) {
    check_force_args(a);
    if (a.xwords) {
        // some table differs between type pairs: the general instantiation (8 or 32 lanes
        // per particle, no fused integrator, one far entry per thread); its dynamic LDS
        // starts with the parameter block
        const size_t lds = sizeof(DParams) + 16 * (size_t)a.xwords + rows_lds;
        if constexpr (FUSED || FAR2) {
            throw Fail("internal: per-type-pair parameter sets on a fused or two-far-entry launch");
        } else if (tpp == 8) {
            hipLaunchKernelGGL((k_forces<32, E, V, CNT, false, false, true>), g, b, lds, st, a);
        } else if (tpp == 32) {
            hipLaunchKernelGGL((k_forces<8, E, V, CNT, false, false, true>), g, b, lds, st, a);
        } else {
            throw Fail("internal: per-type-pair parameter sets need 8 or 32 lanes per particle");
        }
        return;
    }
    const size_t lds = rows_lds;  // the parameter block is static LDS here
    // tpp lanes per particle: 256 / tpp particles per workgroup
    switch (tpp) {
    case 4: hipLaunchKernelGGL((k_forces<64, E, V, CNT, FUSED, FAR2>), g, b, lds, st, a); break;
    case 8: hipLaunchKernelGGL((k_forces<32, E, V, CNT, FUSED, FAR2>), g, b, lds, st, a); break;
    case 16: hipLaunchKernelGGL((k_forces<16, E, V, CNT, FUSED, FAR2>), g, b, lds, st, a); break;
    default: hipLaunchKernelGGL((k_forces<8, E, V, CNT, FUSED, FAR2>), g, b, lds, st, a); break;
    }
}
void check_force_args(const ForceArgs& a) {
    if (a.N == 0) return;
    if (!a.pos || !a.fr0 || !a.fr1 || !a.fr2 || !a.P || !a.force || !a.torque || !a.n_neigh || !a.nlist || !a.head)
        throw Fail("internal: force kernel launched with an unset array");
    if ((a.mask & (1u << T_BONDED)) && (!a.bpart || !a.binfo)) throw Fail("internal: bonded launch without bond table");
    if (a.nb.flags && (!a.nb.count || !a.nb.members || !a.nb.pos_last || !a.bpart))
        throw Fail("internal: row rebuild without a cell list");
    if (a.nb.rebuild_flag && !a.nb.flags) throw Fail("internal: rebuild request without a flags block");
    const StepArgs& t = a.st;
    if (a.integrate && (!t.step_in || !t.quat || !t.vel || !t.angmom || !t.bd || !t.bd_tq))
        throw Fail("internal: fused Langevin half-kick without its arrays");
    if (a.integrate == 2 && (!t.step_out || !t.pos_out || !t.pos_last || !t.force || !t.torque || !t.fr0 ||
                             !t.fr1 || !t.fr2 || !t.flags || !t.rebuild_set || t.pos_out == a.pos))
        throw Fail("internal: fused first half-kick without its arrays");
    if (a.integrate == 2 && t.cells.count && (t.cells.count == a.nb.count || t.cells.members == a.nb.members))
        throw Fail("internal: fused first half-kick bins into the cells the launch reads");
}
void launch_k_any(int tpp, bool e, bool v, dim3 g, dim3 b, size_t lds, hipStream_t st, const ForceArgs& a) {
    check_force_args(a);
    if (a.integrate == 2) {
        if (e || v || a.accumulate) throw Fail("internal: fused run step with energies, virial or accumulation");
        // a run step is fused only on grids too small to fill the GPU: far entries in pairs
        launch_k<false, false, false, true, true>(tpp, g, b, lds, st, a);
    } else if (e && v) launch_k<true, true, false>(tpp, g, b, lds, st, a);
    else if (e) launch_k<true, false, false>(tpp, g, b, lds, st, a);
    else if (v) launch_k<false, true, false>(tpp, g, b, lds, st, a);
    else if (a.far_pairs) launch_k<false, false, false, false, true>(tpp, g, b, lds, st, a);
    else launch_k<false, false, false>(tpp, g, b, lds, st, a);
}

// ------------------------------------------------------------------ neighbour list management
constexpr size_t kMaxCells = size_t(1) << 24;

// LDS of the in-kernel row rebuild: two rows (unsorted, sorted) per particle of a block.
size_t lds_bytes(const oxdna_ctx* c, int tpp) { return (size_t)(kBlock / tpp) * 2 * c->max_neigh * sizeof(uint32_t); }

size_t ncell_total(const oxdna_ctx* c) { return (size_t)c->nc[0] * c->nc[1] * c->nc[2]; }

// Positions and frames come in two buffers of N: run step k's live in buffer k mod 2;
// between runs the state is in buffer 0.
double4* pos_buf(oxdna_ctx* c, uint64_t k) { return c->pos.p + (size_t)(k & 1) * c->N; }
float4* fr_buf(DBuf<float4>& b, const oxdna_ctx* c, uint64_t k) { return b.p + (size_t)(k & 1) * c->N; }

void zero_counts(oxdna_ctx* c) {
    const size_t n = 3 * ncell_total(c);
    if (n && c->cell_count.p)
        hipLaunchKernelGGL(k_zero_u32, dim3(std::min<size_t>(blocks_for(n), 4096)), dim3(kBlock), 0, c->stream,
                           c->cell_count.p, n);
}

void setup_cells(oxdna_ctx* c) {
    // neighbour radius: the largest cut-off of every attached nonbonded term + skin
    double rmax = pair_rmax(c, 0x3f);
    c->rlist = rmax > 0 ? rmax + c->buffer : 0.0;
    size_t total = 1;
    for (int a = 0; a < 3; ++a) {
        int n = c->rlist > 0 ? (int)std::floor(c->L[a] / c->rlist) : 1;
        c->nc[a] = std::max(1, std::min(n, 1024));
    }
    // bound the cell count (cells may only grow, never drop below r_list)
    while ((total = (size_t)c->nc[0] * c->nc[1] * c->nc[2]) > kMaxCells)
        for (int a = 0; a < 3; ++a) c->nc[a] = std::max(1, c->nc[a] / 2);
    if (c->cell_cap == 0) {
        double per_cell = (double)c->N / (double)total;
        c->cell_cap = (uint32_t)std::max(16.0, std::ceil(per_cell * 4.0 + 8.0));
    }
    c->cell_count.alloc(3 * total);        // counts of steps k mod 3
    c->cell_members.alloc(2 * total * c->cell_cap);  // members of steps k mod 2
    zero_counts(c);
    if (c->max_neigh == 0) c->max_neigh = 64;
    if (c->max_neigh > kMaxNeigh) throw Fail("neighbour rows would exceed " + std::to_string(kMaxNeigh) + " entries");
    c->nlist.alloc((size_t)c->N * c->max_neigh);
    std::vector<size_t> h(c->N);
    for (uint32_t i = 0; i < c->N; ++i) h[i] = (size_t)i * c->max_neigh;
    c->head.alloc(c->N);
    HIP_OK(hipMemcpyAsync(c->head.p, h.data(), sizeof(size_t) * c->N, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (lds_bytes(c, c->tpp) > kMaxLds) c->tpp = 32;
}

// Cells of run step t (t = 0: the state a run starts from, and single-point
// evaluations): the counts it bins into (t mod 3; those of step t - 2, index
// (t + 1) mod 3, are cleared as its particles leave them) and its members (t mod 2).
CellArgs cell_args(oxdna_ctx* c, bool with_cells, uint64_t t) {
    CellArgs a;
    a.N = c->N;
    a.pos = pos_buf(c, t);
    for (int k = 0; k < 3; ++k) { a.Linv[k] = c->Linv[k]; a.nc[k] = c->nc[k]; }
    a.cap = c->cell_cap;
    const size_t nct = ncell_total(c);
    a.count = with_cells ? c->cell_count.p + (size_t)(t % 3) * nct : nullptr;
    a.count_prev = with_cells ? c->cell_count.p + (size_t)((t + 1) % 3) * nct : nullptr;
    a.members = c->cell_members.p + (size_t)(t & 1) * nct * c->cell_cap;
    a.flags = c->flags.p;
    a.ncell = (uint32_t)nct;
    a.rebin = c->rebin.p + (size_t)(t & 1) * c->N;
    return a;
}

void read_flags(oxdna_ctx* c, DevFlags& f) {
    HIP_OK(hipMemcpyAsync(&f, c->flags.p, sizeof(DevFlags), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
}

void clear_overflow(oxdna_ctx* c) {
    DevFlags f;
    read_flags(c, f);
    f.overflow = 0; f.max_cell = 0; f.max_neigh = 0; f.rebuild[0] = f.rebuild[1] = f.rebuild[2] = 0;
    HIP_OK(hipMemcpyAsync(c->flags.p, &f, sizeof(DevFlags), hipMemcpyHostToDevice, c->stream));
}

void ensure_ready(oxdna_ctx* c) {
    if (c->N == 0) throw Fail("no particles: call oxdna_set_state first");
    if (c->params_dirty) {
        c->drop_graph();
        setup_cells(c);
        c->nlist_valid = false;
    }
    upload_params(c);
    // per-type-pair sets run on the general kernel instantiations: 8 or 32 lanes
    if (c->P.xwords && c->tpp != 8 && c->tpp != 32) {
        c->tpp = lds_bytes(c, 8) > kMaxLds ? 32 : 8;
        c->drop_graph();
    }
}

StepArgs step_args(oxdna_ctx* c, bool with_cells, uint64_t t);
bool fuse_second_half(const oxdna_ctx* c);

// Forces of run step k (k = 0: the stored state). rebuild: 0 never (rows are current),
// 1 always, 2 when step k's skin test asked for it (run steps, which also clear step
// k + 2's request); integrate: 0 none, 1 the Langevin second half of step k, 2 that and
// the first half of step k + 1, in the epilogue.
ForceArgs force_args(oxdna_ctx* c, uint32_t mask, int accumulate, int rebuild, uint64_t k = 0,
                     int integrate = 0) {
    ForceArgs a{};
    a.N = c->N;
    a.pos = pos_buf(c, k);
    a.fr0 = fr_buf(c->fr0, c, k); a.fr1 = fr_buf(c->fr1, c, k); a.fr2 = fr_buf(c->fr2, c, k);
    a.n_neigh = c->n_neigh.p;
    a.n_near = c->n_near.p;
    a.nlist = c->nlist.p;
    a.head = c->head.p;
    a.far_pairs = fuse_second_half(c) ? 1 : 0;
    a.ncu = c->ncu;
    a.bpart = c->bpart.p;
    a.binfo = c->binfo.p;
    a.P = reinterpret_cast<const DParams*>(c->dparams.p);
    for (int k = 0; k < 3; ++k) { a.L[k] = c->L[k]; a.Linv[k] = c->Linv[k]; }
    a.mask = mask;
    a.xwords = c->P.xwords;
    a.force = c->force.p;
    a.torque = c->torque.p;
    a.virial = c->virial.p;
    a.vpitch = c->N;
    a.epart = c->epart.p;
    a.accumulate = accumulate;
    a.split = nullptr;
    a.counts = nullptr;
    NbrArgs& nb = a.nb;
    nb.pos = a.pos;
    for (int d = 0; d < 3; ++d) { nb.L[d] = c->L[d]; nb.Linv[d] = c->Linv[d]; nb.nc[d] = c->nc[d]; }
    nb.cap = c->cell_cap;
    const CellArgs cells = cell_args(c, true, k);
    nb.count = cells.count;
    nb.members = cells.members;
    nb.rlist2 = c->rlist * c->rlist;
    // row classes (skin included): within reach of the base-base and coaxial terms; of
    // excluded volume; of Debye-Hueckel only
    double rbb = 0.0, rnear = 0.0;
    for (int t = 0; t < 6; ++t) {
        if (!((mask >> t) & 1u) || t == T_DEBYE) continue;
        rnear = std::max(rnear, c->rcut_max[t]);
        if (t != T_EXCVOL) rbb = std::max(rbb, c->rcut_max[t]);
    }
    nb.rnear2 = ((mask >> T_DEBYE) & 1u) ? (rnear + c->buffer) * (rnear + c->buffer) : nb.rlist2;
    nb.rbb2 = ((mask >> T_EXCVOL) & 1u) ? (rbb + c->buffer) * (rbb + c->buffer) : nb.rnear2;
    // near pairs within the base-base radial reach at the rebuild go first (class 0)
    nb.fr0 = a.fr0; nb.fr1 = a.fr1; nb.fr2 = a.fr2;
    nb.bb_act2 = 0.f;
    const DBaseBase* bb = ((mask >> T_HBONDCROSS) & 1u) ? &c->P.hbx
                          : ((mask >> T_HBOND) & 1u) ? &c->P.hb : ((mask >> T_CROSS) & 1u) ? &c->P.cs : nullptr;
    if (bb) {
        float reach = 0.f;
        if ((mask >> T_HBONDCROSS) & 1u) reach = std::max(bb->f1.rchigh, bb->f2.rchigh);
        if ((mask >> T_HBOND) & 1u) reach = std::max(reach, c->P.hb.f1.rchigh);
        if ((mask >> T_CROSS) & 1u) reach = std::max(reach, c->P.cs.f2.rchigh);
        for (int d = 0; d < 3; ++d) nb.bb_base[d] = bb->base[d];
        nb.bb_act2 = reach * reach;
    }
    nb.max_neigh = c->max_neigh;
    // an in-kernel rebuild keeps the cell ids and candidate starts in the sorted row's
    // LDS space until the sort: 60 entries (rows start at 64 and only grow)
    if (rebuild != 0 && c->max_neigh < 64) throw Fail("neighbour rows below 64 entries");
    nb.pos_last = c->pos_last.p;
    nb.flags = (rebuild != 0 && (mask & 0x3fu)) ? c->flags.p : nullptr;
    const bool pairs = (mask & 0x3fu) != 0;
    nb.rebuild_flag = (rebuild == 2 && pairs) ? &c->flags.p->rebuild[k % 3] : nullptr;
    nb.force_rebuild = (rebuild == 1 && pairs) ? 1 : 0;
    a.rebuild_clear = (rebuild == 2 && pairs) ? &c->flags.p->rebuild[(k + 2) % 3] : nullptr;
    a.integrate = integrate;
    if (integrate) a.st = step_args(c, pairs, k + 1);
    return a;
}

uint32_t force_blocks(const oxdna_ctx* c, int tpp) { return blocks_for((size_t)c->N * tpp); }

void launch_forces(oxdna_ctx* c, uint32_t mask, int accumulate, bool want_e, bool want_v, int rebuild,
                   uint64_t k = 0, int integrate = 0) {
    ForceArgs a = force_args(c, mask, accumulate, rebuild, k, integrate);
    dim3 grid(force_blocks(c, c->tpp)), block(kBlock);
    size_t lds = a.nb.flags ? lds_bytes(c, c->tpp) : 0;
    launch_k_any(c->tpp, want_e, want_v, grid, block, lds, c->stream, a);
    if (want_e)
        hipLaunchKernelGGL(k_reduce_energy, dim3(1), dim3(256), 0, c->stream, grid.x, c->epart.p, c->energy.p,
                           accumulate);
}

// Evaluates `m` at the current state, rebuilding every neighbour row first; grows
// the cell / row capacities (with headroom for the rows to grow during a run) until
// nothing overflows.
void compute_rebuild(oxdna_ctx* c, uint32_t m, int fused) {
    bool pairs = (m & 0x3fu) != 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (pairs) {
            clear_overflow(c);
            zero_counts(c);  // start from empty cells
            hipLaunchKernelGGL(k_cell_assign, dim3(blocks_for(c->N)), dim3(kBlock), 0, c->stream,
                               cell_args(c, true, 0));
        }
        if (fused || m == 0) {
            launch_forces(c, m, 0, true, true, pairs ? 1 : 0);
        } else {
            int first = 1;
            for (int t = 0; t < OXDNA_N_TERMS; ++t) {
                if (!((m >> t) & 1u)) continue;
                launch_forces(c, 1u << t, first ? 0 : 1, true, true, (first && pairs) ? 1 : 0);
                first = 0;
            }
            if (first) launch_forces(c, 0, 0, true, true, 0);
        }
        if (!pairs) {
            c->nlist_valid = true;
            return;
        }
        zero_counts(c);  // every buffer starts the next run empty
        DevFlags f;
        read_flags(c, f);
        bool grow = false;
        uint32_t want_cap = (uint32_t)std::ceil(f.max_cell * 1.5) + 4;
        if (want_cap > c->cell_cap) {
            c->cell_cap = want_cap;
            grow = true;
        }
        uint32_t want = std::max<uint32_t>(16, (uint32_t)std::ceil(f.max_neigh * 1.5) + 8);
        if (want > c->max_neigh) {
            c->max_neigh = want;
            grow = true;
        }
        if (!grow && f.overflow == 0) {
            c->nlist_valid = true;
            return;
        }
        c->drop_graph();
        setup_cells(c);
    }
    throw Fail("neighbour list could not be sized");
}

void compute(oxdna_ctx* c, uint32_t mask, int fused) {
    ensure_ready(c);
    hipLaunchKernelGGL(k_frames, dim3(blocks_for(c->N)), dim3(kBlock), 0, c->stream, c->N, c->quat.p, c->charge.p,
                       c->pos.p, c->fr0.p, c->fr1.p, c->fr2.p);
    uint32_t m = effective_mask(c, mask);
    compute_rebuild(c, m, fused);
    c->forces_mask = m;
}

// First half of run step t (t >= 1): from buffer (t - 1) mod 2 to buffer t mod 2.
StepArgs step_args(oxdna_ctx* c, bool with_cells, uint64_t t) {
    StepArgs a;
    a.N = c->N;
    a.bd = c->bd.p; a.bd_tq = c->bd_tq.p;
    a.pos_in = pos_buf(c, t - 1);
    a.pos_out = pos_buf(c, t);
    a.vel = c->vel.p; a.quat = c->quat.p; a.angmom = c->angmom.p;
    a.force = c->force.p; a.torque = c->torque.p; a.charge = c->charge.p;
    a.fr0 = fr_buf(c->fr0, c, t); a.fr1 = fr_buf(c->fr1, c, t); a.fr2 = fr_buf(c->fr2, c, t);
    a.pos_last = c->pos_last.p;
    a.lv = make_langevin(c->dt, c->kT, c->gamma, c->gamma_r, 1.0, 1.0);
    for (int k = 0; k < 3; ++k) { a.L[k] = c->L[k]; a.Linv[k] = c->Linv[k]; }
    a.skin_half2 = 0.25 * c->buffer * c->buffer;
    a.seed = c->seed;
    a.flags = c->flags.p;
    a.rebuild_set = &c->flags.p->rebuild[t % 3];
    a.step_in = &c->flags.p->stepc[(t - 1) & 1];
    a.step_out = &c->flags.p->stepc[t & 1];
    a.cells = cell_args(c, with_cells, t);
    return a;
}

// The Langevin second half-kick rides in the force kernel's epilogue (with the next
// step's first half) while the force grid leaves the GPU latency-bound (few waves per
// SIMD: a launch saved is time saved); on larger systems its one-active-lane-per-
// particle epilogue would cost more issue slots than a kernel with a thread per particle.
bool fuse_second_half(const oxdna_ctx* c) {
    if (c->P.xwords) return false;  // per-type-pair sets: the general kernel, no fused integrator
    const size_t waves = ((size_t)c->N * c->tpp + 63) / 64;
    return waves <= 2048;  // resident at once: 2 waves per SIMD (the fused kernel's occupancy) on 256 CUs
}

// A chunk of a run is: the first half of step 1 (k_step1), inner steps 1 ... n - 1 (the
// forces of step k, its second half and the first half of step k + 1: one fused force
// launch, or a force launch and k_step21 on large systems), and a last step n (forces
// and second half only), which leaves the state complete.
void enqueue_inner(oxdna_ctx* c, uint32_t mask, uint64_t k0, uint32_t len) {
    const bool pairs = (mask & 0x3fu) != 0;
    const bool fuse = fuse_second_half(c);
    for (uint64_t k = k0 + 1; k <= k0 + len; ++k) {
        launch_forces(c, mask, 0, false, false, pairs ? 2 : 0, k, fuse ? 2 : 0);
        if (!fuse)
            hipLaunchKernelGGL(k_step21, dim3(blocks_for(c->N)), dim3(kBlock), 0, c->stream, step_args(c, pairs, k + 1));
    }
}

void enqueue_last(oxdna_ctx* c, uint32_t mask, uint64_t k) {
    const bool pairs = (mask & 0x3fu) != 0;
    const bool fuse = fuse_second_half(c);
    launch_forces(c, mask, 0, false, false, pairs ? 2 : 0, k, fuse ? 1 : 0);
    if (!fuse) hipLaunchKernelGGL(k_step2, dim3(blocks_for(c->N)), dim3(kBlock), 0, c->stream, step_args(c, pairs, k + 1));
}

// Inner steps per graph: a multiple of 6, so that its buffer indices (step mod 2 and
// mod 3) are the same at every replay. A chunk replays the graph, then launches the
// remaining (< 96) inner steps directly (measured: for short chunks direct launches
// beat graphs of 6 steps, whose start and end add gaps to the stream).
constexpr uint32_t kGraphSteps = 96;

hipGraphExec_t capture_inner(oxdna_ctx* c, uint32_t mask, uint32_t len) {
    hipGraph_t g;
    hipGraphExec_t x = nullptr;
    HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_inner(c, mask, 0, len);
    HIP_OK(hipStreamEndCapture(c->stream, &g));
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(e);
    return x;
}

void build_graph(oxdna_ctx* c, uint32_t mask) {
    c->drop_graph();
    c->graph = capture_inner(c, mask, kGraphSteps);
    c->graph_mask = mask;
}

// Autotuner: lanes per particle (8, 16, 32) of the force kernel, timed at the
// current state (rows current, no rebuild) with HIP events; keeps the fastest.
void autotune(oxdna_ctx* c, uint32_t m) {
    if (c->P.xwords) {
        // per-type-pair sets: the general kernel runs 8 lanes per particle (32 when the rows
        // are too long for that many particles per workgroup)
        const int want = lds_bytes(c, 8) > kMaxLds ? 32 : 8;
        if (c->tpp != want) {
            c->tpp = want;
            c->drop_graph();
        }
        return;
    }
    if (c->fixed_tpp) {
        // the pinned value, unless the rows grew too long for its LDS (then 32 lanes)
        const int want = lds_bytes(c, c->fixed_tpp) > kMaxLds ? 32 : c->fixed_tpp;
        if (c->tpp != want) {
            c->tpp = want;
            c->drop_graph();
        }
        return;
    }
    if (c->tuned_mask == m && c->tuned_n == c->N) return;
    const int cand[4] = {4, 8, 16, 32};
    float best = 1e30f;
    int best_tpp = c->tpp;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    for (int tpp : cand) {
        if (lds_bytes(c, tpp) > kMaxLds) continue;
        c->tpp = tpp;
        for (int w = 0; w < 3; ++w) launch_forces(c, m, 0, false, false, 0);
        for (int trial = 0; trial < 3; ++trial) {  // best of three: timing noise picks no loser
            HIP_OK(hipEventRecord(e0, c->stream));
            for (int r = 0; r < 20; ++r) launch_forces(c, m, 0, false, false, 0);
            HIP_OK(hipEventRecord(e1, c->stream));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; best_tpp = tpp; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tpp = best_tpp;
    c->tuned_mask = m;
    c->tuned_n = c->N;
    c->drop_graph();
    // the timing launches overwrote the forces with identical values; nothing to redo
}

RunState live_state(oxdna_ctx* c) {
    return RunState{c->pos.p, c->vel.p, c->quat.p, c->angmom.p, c->pos_last.p,
                    c->bd.p, c->bd_tq.p, c->force.p, c->torque.p};
}

RunState saved_state(oxdna_ctx* c) {
    c->sv_pos.alloc(c->N); c->sv_vel.alloc(c->N); c->sv_quat.alloc(c->N); c->sv_angmom.alloc(c->N);
    c->sv_pos_last.alloc(c->N); c->sv_bd.alloc(c->N); c->sv_bd_tq.alloc(c->N);
    c->sv_force.alloc(c->N); c->sv_torque.alloc(c->N);
    return RunState{c->sv_pos.p, c->sv_vel.p, c->sv_quat.p, c->sv_angmom.p, c->sv_pos_last.p,
                    c->sv_bd.p, c->sv_bd_tq.p, c->sv_force.p, c->sv_torque.p};
}

uint32_t grid_stride_blocks(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>(blocks_for(n), 2048)); }

// Runs `nsteps` Langevin steps in chunks of up to 42 graph lengths. Between runs (and
// after every single-point evaluation) the cell counts are zero and no rebuild request
// or overflow is pending, so a chunk is: k_step1 (which also saves the chunk's starting
// state), the inner steps, the last step, and k_chunk_end (state to buffer 0, status
// into pinned host memory, cells and flags left clean), then one stream
// synchronisation. A chunk whose cells or rows overflowed is run again from the saved
// start with larger capacities.
void run(oxdna_ctx* c, uint32_t nsteps, uint32_t mask, int use_graph) {
    if (!c->have_langevin) throw Fail("no integrator: call oxdna_set_langevin first");
    ensure_ready(c);
    uint32_t m = effective_mask(c, mask);
    if (c->forces_mask != m || !c->nlist_valid) compute(c, m, 1);
    autotune(c, m);
    // the graphs are captured and instantiated once, at the first run of a setup (its
    // cost, ~1 ms, goes there and not into the first chunk long enough to replay them)
    if (use_graph && (!c->graph || c->graph_mask != m)) build_graph(c, m);
    const uint32_t chunk_max = 42 * kGraphSteps + 1;
    const bool pairs = (m & 0x3fu) != 0;
    uint32_t done = 0;
    while (done < nsteps) {
        const uint32_t chunk = std::min(chunk_max, nsteps - done);
        // the chunk's steps are numbered 1 ... chunk from its starting state (step 0)
        const unsigned long long step0 = c->host_step + done;
        for (int attempt = 0;; ++attempt) {
            const size_t ncount = pairs ? 3 * ncell_total(c) : 0;
            hipLaunchKernelGGL(k_step1, dim3((uint32_t)((c->N + kStep1Block / 2 - 1) / (kStep1Block / 2))),
                               dim3(kStep1Block), 0, c->stream, step_args(c, pairs, 1),
                               saved_state(c), step0);
            uint32_t k = 0;  // inner steps enqueued
            const bool graphs = use_graph && chunk - 1 >= kGraphSteps;
            const bool timed = c->timing && graphs;
            if (graphs) {
                if (!c->graph || c->graph_mask != m) build_graph(c, m);
                if (timed) HIP_OK(hipEventRecord(c->ev0, c->stream));
                for (; k + kGraphSteps <= chunk - 1; k += kGraphSteps) HIP_OK(hipGraphLaunch(c->graph, c->stream));
                if (timed) HIP_OK(hipEventRecord(c->ev1, c->stream));
            }
            // the rest without the graph: the same launches, so both paths are bit-identical
            enqueue_inner(c, m, k, chunk - 1 - k);
            enqueue_last(c, m, chunk);
            c->status_h->done = 0u;
            hipLaunchKernelGGL(k_chunk_end, dim3(grid_stride_blocks(std::max<size_t>(c->N, ncount))), dim3(kBlock), 0,
                               c->stream, c->N, (int)(chunk & 1u), c->pos.p, c->fr0.p, c->fr1.p, c->fr2.p,
                               c->cell_count.p, ncount, c->flags.p, step0 + chunk, c->status_d);
            HIP_OK(hipGetLastError());
            HIP_OK(hipStreamSynchronize(c->stream));
            const volatile RunStatus* st = c->status_h;
            if (st->done != 1u) throw Fail("internal: run chunk status not written");
            if (timed) {
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
                c->timed_ms += ms;
                c->timed_steps += k;
            }
            if (st->overflow == 0u) break;
            if (attempt >= 4) throw Fail("neighbour list kept overflowing during the run");
            // grow, restore the chunk's starting state and run it again (setup_cells and
            // compute leave the cells empty and the flags clean)
            const uint32_t max_neigh = st->max_neigh, max_cell = st->max_cell;
            hipLaunchKernelGGL(k_copy_state, dim3(grid_stride_blocks(c->N)), dim3(kBlock), 0, c->stream, c->N,
                               saved_state(c), live_state(c));
            c->max_neigh = std::max(c->max_neigh, (uint32_t)std::ceil(max_neigh * 1.5) + 8);
            c->cell_cap = std::max(c->cell_cap, (uint32_t)std::ceil(max_cell * 1.5) + 4);
            c->drop_graph();
            setup_cells(c);
            compute(c, m, 1);
        }
        done += chunk;
    }
    c->host_step += nsteps;
}

// Neighbour-list skin autotuner: from the current state, runs `warm` untimed and then
// `steps` timed Langevin steps at each candidate buffer (rows rebuilt at that skin, cell
// and row capacities sized afresh, lanes per particle tuned again), measures steps/s on
// the host clock around whole run calls (rebuild launches, chunk ends and all), and
// keeps the fastest. The state (positions, orientations, momenta, the stored
// half-kick terms, the step count) is put back after every trial and at the end, so
// the caller's trajectory continues from where it stood, at the chosen skin.
double tune_buffer(oxdna_ctx* c, uint32_t mask, const std::vector<double>& cand, uint32_t steps, uint32_t warm,
                   double* rates) {
    if (c->n_rep) throw Fail("oxdna_tune_nlist_buffer: one-system contexts only (replicas have no neighbour list)");
    if (!c->have_langevin) throw Fail("oxdna_tune_nlist_buffer: no integrator: call oxdna_set_langevin first");
    if (cand.empty()) throw Fail("oxdna_tune_nlist_buffer: no candidate buffers");
    for (double b : cand)
        if (!(b > 0.0) || !(b <= 4.0)) throw Fail("oxdna_tune_nlist_buffer: candidate buffers must lie in (0, 4]");
    if (steps == 0) throw Fail("oxdna_tune_nlist_buffer: steps must be > 0");
    ensure_ready(c);
    const uint32_t m = effective_mask(c, mask);
    if (!(m & 0x3fu)) throw Fail("oxdna_tune_nlist_buffer: no nonbonded term selected (no neighbour list to tune)");
    if (c->forces_mask != m || !c->nlist_valid) compute(c, m, 1);
    // the starting state, kept apart from the run's own per-chunk save (released however
    // the tuner ends)
    struct Kept {
        DBuf<double4> pos, vel, quat, ang, last, bd, bdtq;
        DBuf<float4> f, t;
        ~Kept() {
            pos.release(); vel.release(); quat.release(); ang.release(); last.release();
            bd.release(); bdtq.release(); f.release(); t.release();
        }
    } kept;
    kept.pos.alloc(c->N); kept.vel.alloc(c->N); kept.quat.alloc(c->N); kept.ang.alloc(c->N);
    kept.last.alloc(c->N); kept.bd.alloc(c->N); kept.bdtq.alloc(c->N); kept.f.alloc(c->N); kept.t.alloc(c->N);
    const RunState keep{kept.pos.p, kept.vel.p, kept.quat.p, kept.ang.p, kept.last.p,
                        kept.bd.p, kept.bdtq.p, kept.f.p, kept.t.p};
    hipLaunchKernelGGL(k_copy_state, dim3(grid_stride_blocks(c->N)), dim3(kBlock), 0, c->stream, c->N,
                       live_state(c), keep);
    HIP_OK(hipGetLastError());
    const uint64_t step_kept = c->host_step;
    const double buffer_kept = c->buffer;
    DevFlags f0;
    read_flags(c, f0);
    auto set_skin = [&](double b) {
        hipLaunchKernelGGL(k_copy_state, dim3(grid_stride_blocks(c->N)), dim3(kBlock), 0, c->stream, c->N, keep,
                           live_state(c));
        HIP_OK(hipGetLastError());
        c->host_step = step_kept;
        c->buffer = b;
        c->params_dirty = true;
        c->nlist_valid = false;
        c->max_neigh = 0;   // capacities sized afresh for this skin
        c->cell_cap = 0;
        c->tuned_mask = 0xffffffffu;
        compute(c, m, 1);
    };
    // the step and build counts as they stood (the trials' steps are not the caller's)
    auto restore_counts = [&]() {
        DevFlags f;
        read_flags(c, f);
        f.step = f0.step;
        f.builds = f0.builds;
        HIP_OK(hipMemcpyAsync(c->flags.p, &f, sizeof(DevFlags), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    };
    double best = -1.0, best_b = cand[0];
    try {
        for (size_t k = 0; k < cand.size(); ++k) {
            set_skin(cand[k]);
            if (warm) run(c, warm, m, 1);
            HIP_OK(hipStreamSynchronize(c->stream));
            const auto t0 = std::chrono::steady_clock::now();
            run(c, steps, m, 1);
            HIP_OK(hipStreamSynchronize(c->stream));
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const double rate = steps / std::max(sec, 1e-9);
            if (rates) rates[k] = rate;
            if (rate > best) { best = rate; best_b = cand[k]; }
        }
    } catch (...) {
        // a failed trial leaves the caller's state, skin and counts as they were
        try {
            set_skin(buffer_kept);
            restore_counts();
        } catch (...) {
        }
        throw;
    }
    set_skin(best_b);
    autotune(c, m);
    build_graph(c, m);  // the caller's next run replays it instead of capturing it
    restore_counts();
    return best_b;
}

// ------------------------------------------------------------------ batched replicas
// Largest replica (particles), one workgroup each: its LDS layout with the minimum record
// capacity (ncap + 64 pair records) fills at most the 160 KiB of one gfx950 workgroup.
constexpr uint32_t kEnsMaxN = 496;
constexpr uint32_t kEnsChunk = 2048;   // Langevin steps per persistent launch

// Sums the per-replica energies into the context's 8 term totals (one block, fixed order).
__global__ void k_rep_sum(uint32_t n_rep, const double* __restrict__ rep, double* out) {
    __shared__ double s[256];
    for (int slot = 0; slot < 8; ++slot) {
        double v = 0.0;
        for (uint32_t r = threadIdx.x; r < n_rep; r += blockDim.x) v += rep[(size_t)r * ENS_SLOTS + slot];
        s[threadIdx.x] = v;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[slot] = s[0];
        __syncthreads();
    }
}

void ensure_ready_ens(oxdna_ctx* c, uint32_t m) {
    if (c->N == 0) throw Fail("no particles: call oxdna_set_state first");
    upload_params(c);
    double rmax = pair_rmax(c, m);
    for (uint32_t r = 0; r < c->n_rep; ++r)
        for (int k = 0; k < 3; ++k)
            if (!(c->rep_box_h[3 * r + k] > 2.0 * rmax))
                throw Fail("replica " + std::to_string(r) + ": box length must exceed twice the largest cut-off");
}

EnsArgs ens_args(oxdna_ctx* c, uint32_t mask, uint32_t nsteps, uint32_t nevals, int want_e) {
    EnsArgs a;
    a.n_rep = c->n_rep;
    a.offsets = c->rep_off.p;
    a.order = c->rep_order.p;
    a.boxes = c->rep_box.p;
    a.pos = c->pos.p; a.vel = c->vel.p; a.quat = c->quat.p; a.angmom = c->angmom.p;
    a.bd = c->bd.p; a.bd_tq = c->bd_tq.p;
    a.charge = c->charge.p;
    a.bpart = c->bpart.p; a.binfo = c->binfo.p;
    a.P = reinterpret_cast<const DParams*>(c->dparams.p);
    a.mask = mask;
    a.force = c->force.p; a.torque = c->torque.p;
    a.rep_energy = c->rep_energy.p;
    a.bias = c->have_bias ? c->rep_bias.p : nullptr;
    a.lv = make_langevin(c->dt, c->kT, c->gamma, c->gamma_r, 1.0, 1.0);
    a.seed = c->seed;
    a.step0 = c->host_step;
    a.nsteps = nsteps;
    a.nevals = nevals;
    a.want_e = want_e;
    a.counts = nullptr;
    a.ncap = 0;  // set by launch_ens
    a.cap = 0;
    // near: inside any cut-off but Debye-Hueckel's; far: inside Debye-Hueckel's only
    float rn = 0.f, rf = 0.f;
    for (int t = 0; t < 6; ++t) {
        if (!((mask >> t) & 1u)) continue;
        for (int k = 0; k < 16; ++k) {
            float r2 = c->P.rcutsq[t][k];
            if (t == T_DEBYE) rf = std::max(rf, r2);
            else rn = std::max(rn, r2);
        }
    }
    a.rnear2 = rn;
    a.rfar2 = std::max(rf, rn);
    return a;
}

// Smallest replica workgroup (power of two, 64..512 threads) that holds the largest
// replica: one thread per particle.
int ens_nmax(const oxdna_ctx* c) {
    int nmax = 64;
    while ((uint32_t)nmax < c->rep_nmax) nmax *= 2;
    return nmax;
}

// LDS shape of a replica launch on nt-thread workgroups: particle capacity (the largest
// replica, rounded up to 8) and pair records per pass (the tile depth), as many as fit
// the LDS share of one workgroup when ceil(replicas / compute units) workgroups share a
// compute unit (at most 8 wavefronts per compute unit: the kernel's registers allow 2
// per SIMD), divided by `div` (1, 2, 4: shallower tiles, more passes, less LDS). At
// least ncap + 64 records, so a pass holds any particle's entries and some more.
struct EnsShape {
    uint32_t ncap, cap, lds;
};

uint32_t ens_ncap(const oxdna_ctx* c) { return std::max<uint32_t>(8u, (c->rep_nmax + 7u) & ~7u); }

// The general replica instantiation (512 threads, full depth, the parameter block read in
// global memory) runs when the tables differ between type pairs or when the largest
// replica's layout with the block in LDS does not fit one workgroup at the minimum depth
// (489-496 particles); every other launch keeps the block in LDS.
bool ens_general(const oxdna_ctx* c) {
    if (c->P.xwords) return true;
    const uint32_t ncap = ens_ncap(c);
    return ens_layout(ncap, ncap + 64u, 512u, true).bytes > c->lds_block;
}

EnsShape ens_shape(const oxdna_ctx* c, int nt, uint32_t div = 1) {
    EnsShape s{};
    s.ncap = ens_ncap(c);
    const bool staged = !ens_general(c);
    const uint32_t ncu = std::max<uint32_t>(1u, c->ncu);
    const uint32_t waves = (uint32_t)nt / 64u;
    const uint32_t need = s.ncap + 64u;
    const uint32_t pairs = s.ncap * (s.ncap - 1u) / 2u;
    const uint32_t base = ens_layout(s.ncap, 0u, (uint32_t)nt, staged).bytes;
    uint32_t k = std::min<uint32_t>((c->n_rep + ncu - 1u) / ncu, std::max<uint32_t>(1u, 8u / waves));
    for (k = std::max<uint32_t>(k, 1u); k >= 1u; --k) {
        const uint32_t budget = std::min<uint32_t>(c->lds_block, c->lds_cu / k);
        if (base + 48u * need > budget) continue;
        s.cap = std::min<uint32_t>((budget - base) / 48u, std::max(need, pairs));
        s.cap = std::max(need, s.cap / std::max<uint32_t>(div, 1u));
        s.lds = ens_layout(s.ncap, s.cap, (uint32_t)nt, staged).bytes;
        return s;
    }
    throw Fail("replicas of " + std::to_string(c->rep_nmax) + " particles do not fit the LDS of one workgroup (" +
               std::to_string(c->lds_block) + " bytes)");
}

// Launches the replica kernel with workgroup size nt (count: the counting variant). The
// general instantiation, and the count, run 512-thread workgroups at full depth.
void launch_ens(oxdna_ctx* c, int nt, bool count, EnsArgs a) {
    const bool general = ens_general(c);
    if (general || count) nt = 512;
    if (nt < ens_nmax(c) || a.n_rep == 0 || !a.offsets || !a.order || !a.P || c->rep_nmax > (uint32_t)nt)
        throw Fail("internal: replica kernel launch does not fit the replicas");
    const EnsShape sh = ens_shape(c, nt, (general || count) ? 1u : c->ens_div);
    a.ncap = sh.ncap;
    a.cap = sh.cap;
    HIP_OK(launch_ensemble(nt, count, general, a.n_rep, sh.lds, c->stream, a));
}

int ens_default_nt(const oxdna_ctx* c) { return std::max(256, ens_nmax(c)); }

// Autotuner: workgroup size of the replica kernel (from the replica capacity up to
// 512 threads, every power of two) and its tile depth (pair records per pass: the LDS
// share's full depth, a half or a quarter of it, each distinct depth once), timed as
// forces-only launches (the state is not changed) with HIP events; keeps the fastest.
void autotune_ens(oxdna_ctx* c, uint32_t m) {
    if (ens_general(c)) {  // the general instantiation has one workgroup size
        c->ens_nt = 512;
        c->ens_div = 1;
        c->ens_tuned_mask = m;
        return;
    }
    if (c->ens_tuned_mask == m && c->ens_nt != 0) return;
    if (c->fixed_ens_nt) {
        if (c->fixed_ens_nt < ens_nmax(c))
            throw Fail("oxdna_set_replica_tiling: a workgroup of " + std::to_string(c->fixed_ens_nt) +
                       " threads is smaller than the largest replica needs (" + std::to_string(ens_nmax(c)) + ")");
        c->ens_nt = c->fixed_ens_nt;
        c->ens_div = c->fixed_ens_div;
        c->ens_tuned_mask = m;
        return;
    }
    float best = 1e30f;
    int best_nt = ens_default_nt(c);
    uint32_t best_div = 1;
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    for (int nt = ens_nmax(c); nt <= 512; nt *= 2) {
        uint32_t last_cap = 0;
        for (uint32_t div = 1; div <= 4; div *= 2) {
            const uint32_t cap = ens_shape(c, nt, div).cap;
            if (cap == last_cap) continue;  // the depth is at its floor already
            last_cap = cap;
            c->ens_div = div;
            EnsArgs a = ens_args(c, m, 0, 8, 0);
            launch_ens(c, nt, false, a);
            for (int trial = 0; trial < 3; ++trial) {
                HIP_OK(hipEventRecord(e0, c->stream));
                launch_ens(c, nt, false, a);
                HIP_OK(hipEventRecord(e1, c->stream));
                HIP_OK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; best_nt = nt; best_div = div; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->ens_nt = best_nt;
    c->ens_div = best_div;
    c->ens_tuned_mask = m;
}

void compute_ens(oxdna_ctx* c, uint32_t mask) {
    uint32_t m = effective_mask(c, mask);
    ensure_ready_ens(c, m);
    if (c->fixed_ens_nt) autotune_ens(c, m);  // takes the pinned shape, times nothing
    int nt = ens_general(c) ? 512 : (c->ens_nt ? c->ens_nt : ens_default_nt(c));
    launch_ens(c, nt, false, ens_args(c, m, 0, 1, 1));
    hipLaunchKernelGGL(k_rep_sum, dim3(1), dim3(256), 0, c->stream, c->n_rep, c->rep_energy.p, c->energy.p);
    HIP_OK(hipMemsetAsync(c->virial.p, 0, sizeof(float) * 6 * (size_t)c->N, c->stream));
    c->forces_mask = m;
}

void run_ens(oxdna_ctx* c, uint32_t nsteps, uint32_t mask) {
    if (!c->have_langevin) throw Fail("no integrator: call oxdna_set_langevin first");
    uint32_t m = effective_mask(c, mask);
    ensure_ready_ens(c, m);
    if (c->forces_mask != m) compute_ens(c, m);
    autotune_ens(c, m);
    uint32_t done = 0;
    while (done < nsteps) {
        uint32_t chunk = std::min(kEnsChunk, nsteps - done);
        launch_ens(c, c->ens_nt, false, ens_args(c, m, chunk, 0, 0));
        c->host_step += chunk;
        done += chunk;
    }
    HIP_OK(hipGetLastError());
}

}  // namespace

// =====================================================================================
// C ABI
// =====================================================================================
extern "C" {

const char* oxdna_last_error(const oxdna_ctx* ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }

int oxdna_abi_version(void) { return OXDNA_ABI_VERSION; }

int oxdna_device_name(int device, char* buf, size_t buflen) {
    return guard(nullptr, [&] {
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, device));
        std::snprintf(buf, buflen, "%s", prop.gcnArchName);
    });
}

int oxdna_ctx_create(int device, void* stream, oxdna_ctx** out) {
    if (!out) return fail(nullptr, "oxdna_ctx_create: out is NULL");
    *out = nullptr;
    oxdna_ctx* c = new oxdna_ctx();
    c->device = device;
    int rc = guard(nullptr, [&] {
        int count = 0;
        HIP_OK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw Fail("oxdna_ctx_create: no such HIP device");
        HIP_OK(hipSetDevice(device));
        int ncu = 0;
        HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        c->ncu = (uint32_t)ncu;
        int lds_block = 0, lds_cu = 0;
        HIP_OK(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        HIP_OK(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
        c->lds_block = (uint32_t)lds_block;
        c->lds_cu = (uint32_t)std::max(lds_cu, lds_block);
        if (stream) c->stream = (hipStream_t)stream;
        else {
            HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
        c->flags.alloc(1);
        DevFlags f{};
        HIP_OK(hipMemcpy(c->flags.p, &f, sizeof(f), hipMemcpyHostToDevice));
        // the run status: written by the GPU at the end of each chunk, read by the host
        // after the stream synchronisation (coherent: no cache flush between them)
        HIP_OK(hipHostMalloc((void**)&c->status_h, sizeof(RunStatus), hipHostMallocMapped | hipHostMallocCoherent));
        std::memset(c->status_h, 0, sizeof(RunStatus));
        HIP_OK(hipHostGetDevicePointer((void**)&c->status_d, c->status_h, 0));
        c->energy.alloc(8);
        HIP_OK(hipMemset(c->energy.p, 0, 8 * sizeof(double)));
    });
    if (rc != 0) {
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

int oxdna_ctx_destroy(oxdna_ctx* ctx) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return 0;
}

int oxdna_set_state(oxdna_ctx* c, unsigned int n, const double* pos, const double* quat, const int32_t* type,
                    const double* charge, const oxdna_box* box) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (n == 0) throw Fail("oxdna_set_state: n must be > 0");
        if (n > kIndexMask) throw Fail("oxdna_set_state: n must be below 2^30 (neighbour rows keep 30-bit indices)");
        if (!pos || !quat || !type || !box) throw Fail("oxdna_set_state: NULL array");
        for (int a = 0; a < 3; ++a)
            if (!(box->L[a] > 0.0)) throw Fail("oxdna_set_state: box lengths must be positive");
        bool resized = n != c->N;
        c->N = n;
        for (int a = 0; a < 3; ++a) { c->L[a] = box->L[a]; c->Linv[a] = 1.0 / box->L[a]; }
        std::vector<double4> p(n), q(n);
        std::vector<double> ch(n);
        for (uint32_t i = 0; i < n; ++i) {
            if (type[i] < 0 || type[i] >= 4) throw Fail("oxdna_set_state: type ids must be in [0, 4)");
            double w;
            uint64_t bits = (uint64_t)(uint32_t)type[i];
            std::memcpy(&w, &bits, sizeof(w));
            p[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], w);
            q[i] = make_double4(quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]);
            ch[i] = charge ? charge[i] : 1.0;
        }
        if (resized) {
            c->pos.release(); c->quat.release(); c->vel.release(); c->angmom.release(); c->pos_last.release();
            c->bd.release(); c->bd_tq.release();
            c->charge.release(); c->fr0.release(); c->fr1.release(); c->fr2.release(); c->force.release();
            c->rebin.release();
            c->torque.release(); c->virial.release(); c->n_neigh.release(); c->n_near.release(); c->nlist.release();
            c->head.release();
            c->epart.release();
            c->have_bonds = false;
            c->n_rep = 0;  // the replica layout no longer fits
            c->have_bias = false;
        }
        c->pos.alloc(2 * (size_t)n); c->quat.alloc(n); c->vel.alloc(n); c->angmom.alloc(n); c->pos_last.alloc(n);
        c->bd.alloc(n); c->bd_tq.alloc(n);
        c->charge.alloc(n); c->fr0.alloc(2 * (size_t)n); c->fr1.alloc(2 * (size_t)n); c->fr2.alloc(2 * (size_t)n);
        c->rebin.alloc(2 * (size_t)n);
        HIP_OK(hipMemsetAsync(c->rebin.p, 0xff, sizeof(uint32_t) * 2 * n, c->stream));  // kNoCell
        c->force.alloc(n); c->torque.alloc(n); c->virial.alloc(6 * (size_t)n); c->n_neigh.alloc(n);
        c->n_near.alloc(n);
        c->epart.alloc((size_t)blocks_for((size_t)n * kTPPMax) * 8);
        HIP_OK(hipMemcpyAsync(c->pos.p, p.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->stream));
        // the second buffer too: a run reads its stale positions (to clear cells)
        HIP_OK(hipMemcpyAsync(c->pos.p + n, p.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->pos_last.p, p.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->quat.p, q.data(), sizeof(double4) * n, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->charge.p, ch.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        if (resized) {
            HIP_OK(hipMemsetAsync(c->vel.p, 0, sizeof(double4) * n, c->stream));
            HIP_OK(hipMemsetAsync(c->angmom.p, 0, sizeof(double4) * n, c->stream));
            HIP_OK(hipMemsetAsync(c->bd.p, 0, sizeof(double4) * n, c->stream));
            HIP_OK(hipMemsetAsync(c->bd_tq.p, 0, sizeof(double4) * n, c->stream));
            std::vector<int> none(2 * (size_t)n, -1);
            c->bpart.alloc(2 * (size_t)n); c->binfo.alloc(2 * (size_t)n);
            HIP_OK(hipMemcpyAsync(c->bpart.p, none.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice, c->stream));
            HIP_OK(hipMemcpyAsync(c->binfo.p, none.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice, c->stream));
            c->cell_cap = 0;
            c->max_neigh = 0;
            c->params_dirty = true;
        }
        HIP_OK(hipStreamSynchronize(c->stream));
        c->drop_graph();
        c->nlist_valid = false;
        c->forces_mask = 0xffffffffu;
    });
}

int oxdna_set_momenta(oxdna_ctx* c, const double* vel, const double* angmom) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (c->N == 0) throw Fail("oxdna_set_momenta: set the state first");
        std::vector<double4> v(c->N), p(c->N);
        for (uint32_t i = 0; i < c->N; ++i) {
            v[i] = vel ? make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0) : make_double4(0, 0, 0, 0);
            p[i] = angmom ? make_double4(angmom[4 * i], angmom[4 * i + 1], angmom[4 * i + 2], angmom[4 * i + 3])
                          : make_double4(0, 0, 0, 0);
        }
        HIP_OK(hipMemcpyAsync(c->vel.p, v.data(), sizeof(double4) * c->N, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->angmom.p, p.data(), sizeof(double4) * c->N, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int oxdna_set_bonds(oxdna_ctx* c, unsigned int nb, const int32_t* ab, const int32_t* btype) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (c->N == 0) throw Fail("oxdna_set_bonds: set the state first");
        std::vector<int> part(2 * (size_t)c->N, -1), info(2 * (size_t)c->N, -1);
        for (uint32_t b = 0; b < nb; ++b) {
            int A = ab[2 * b], B = ab[2 * b + 1];
            int t = btype ? btype[b] : 0;
            if (A < 0 || B < 0 || A >= (int)c->N || B >= (int)c->N || A == B)
                throw Fail("oxdna_set_bonds: bond " + std::to_string(b) + " has an invalid particle index");
            if (t < 0 || t >= MAX_BOND_TYPES) throw Fail("oxdna_set_bonds: bond type out of range");
            int ends[2] = {A, B};
            for (int r = 0; r < 2; ++r) {
                int i = ends[r];
                int slot = part[2 * i] < 0 ? 0 : (part[2 * i + 1] < 0 ? 1 : -1);
                if (slot < 0)
                    throw Fail("oxdna_set_bonds: particle " + std::to_string(i) +
                               " has more than two backbone bonds");
                part[2 * i + slot] = ends[1 - r];
                info[2 * i + slot] = r | (t << 1);
            }
        }
        c->bpart.alloc(2 * (size_t)c->N);
        c->binfo.alloc(2 * (size_t)c->N);
        HIP_OK(hipMemcpyAsync(c->bpart.p, part.data(), sizeof(int) * 2 * c->N, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(c->binfo.p, info.data(), sizeof(int) * 2 * c->N, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->have_bonds = nb > 0;
        c->nlist_valid = false;
        c->forces_mask = 0xffffffffu;
        c->drop_graph();
    });
}

int oxdna_set_pair_params(oxdna_ctx* c, int term, int ntypes, const void* params, const double* rcut) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (term < 0 || term >= OXDNA_TERM_BONDED) throw Fail("oxdna_set_pair_params: not a nonbonded term id");
        if (!params) {
            c->pair_set[term] = false;
            c->rcut_max[term] = 0.0;
            c->xset[term].clear();
            for (int k = 0; k < 16; ++k) c->pidx[term][k] = 0;
        } else {
            if (ntypes < 1 || ntypes > 4)
                throw Fail("oxdna_set_pair_params: ntypes must be 1..4 (the oxDNA type set A, G, C, T/U)");
            if (!rcut) throw Fail("oxdna_set_pair_params: rcut table is NULL");
            for (int k = 0; k < ntypes * ntypes; ++k)
                if (!(rcut[k] >= 0.0)) throw Fail("oxdna_set_pair_params: negative or NaN cut-off");
            DParams P = c->P;
            std::vector<unsigned char> xset[6];
            uint8_t pidx[6][16];
            std::memcpy(pidx, c->pidx, sizeof(pidx));
            for (int t = 0; t < 6; ++t)
                if (t != term) xset[t] = c->xset[t];
            load_pair_params(P, xset[term], pidx[term], term, ntypes, params);
            (void)param_blob(P, xset, pidx);  // checks the size of the extras block
            c->P = P;
            c->xset[term] = xset[term];
            std::memcpy(c->pidx, pidx, sizeof(pidx));
            double rmax = 0.0;
            for (int k = 0; k < 16; ++k) {
                double r = k < ntypes * ntypes ? rcut[k] : 0.0;
                c->P.rcutsq[term][k] = (float)(r * r);
                rmax = std::max(rmax, r);
            }
            c->P.ntypes = ntypes;
            c->pair_set[term] = true;
            c->rcut_max[term] = rmax;
        }
        c->P.mask = 0;
        for (int t = 0; t < 6; ++t) c->P.mask |= c->pair_set[t] ? (1u << t) : 0u;
        c->params_dirty = true;
        c->forces_mask = 0xffffffffu;
    });
}

int oxdna_set_bond_params(oxdna_ctx* c, unsigned int nbt, const oxdna_bonded_params* params) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (!params || nbt == 0) {
            c->bond_set = false;
        } else {
            if (nbt > MAX_BOND_TYPES) throw Fail("oxdna_set_bond_params: at most 4 bond types");
            for (uint32_t k = 0; k < nbt; ++k) {
                if (!(params[k].delta_backbone > 0.0)) throw Fail("oxdna_set_bond_params: delta must be > 0");
                load_bond_params(c->P.bond[k], params[k]);
            }
            c->P.n_bond_types = (int)nbt;
            c->bond_set = true;
        }
        c->params_dirty = true;
        c->forces_mask = 0xffffffffu;
    });
}

int oxdna_set_nlist(oxdna_ctx* c, double buffer) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (!(buffer >= 0.0)) throw Fail("oxdna_set_nlist: buffer must be >= 0");
        c->buffer = buffer;
        c->params_dirty = true;
        c->nlist_valid = false;
    });
}

int oxdna_compute(oxdna_ctx* c, unsigned int mask, int fused) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (c->n_rep) compute_ens(c, mask);
        else compute(c, mask, fused);
    });
}

int oxdna_get_term_energies(oxdna_ctx* c, double out[8]) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        HIP_OK(hipMemcpyAsync(out, c->energy.p, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int oxdna_get_forces(oxdna_ctx* c, double* force4, double* torque3, double* virial6) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        uint32_t n = c->N;
        std::vector<float4> f(n), t(n);
        std::vector<float> v(6 * (size_t)n);
        HIP_OK(hipMemcpyAsync(f.data(), c->force.p, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(t.data(), c->torque.p, sizeof(float4) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(v.data(), c->virial.p, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < n; ++i) {
            if (force4) { force4[4 * i] = f[i].x; force4[4 * i + 1] = f[i].y; force4[4 * i + 2] = f[i].z; force4[4 * i + 3] = f[i].w; }
            if (torque3) { torque3[3 * i] = t[i].x; torque3[3 * i + 1] = t[i].y; torque3[3 * i + 2] = t[i].z; }
            if (virial6)
                for (int s = 0; s < 6; ++s) virial6[6 * i + s] = v[(size_t)s * n + i];
        }
    });
}

int oxdna_get_nlist_pairs(oxdna_ctx* c, int32_t* pairs, size_t max_pairs, size_t* npairs) {
    if (!c) return fail(nullptr, "NULL context");
    int rc = 0;
    int g = guard(c, [&] {
        if (c->n_rep) throw Fail("oxdna_get_nlist_pairs: replicas use no neighbour list");
        if (!c->nlist_valid) compute(c, 0x3fu, 1);
        uint32_t n = c->N;
        std::vector<uint32_t> nn(n), rows((size_t)n * c->max_neigh);
        HIP_OK(hipMemcpyAsync(nn.data(), c->n_neigh.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(rows.data(), c->nlist.p, sizeof(uint32_t) * rows.size(), hipMemcpyDeviceToHost,
                              c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        size_t cnt = 0;
        std::vector<uint32_t> js;
        for (uint32_t i = 0; i < n; ++i) {
            js.clear();
            for (uint32_t k = 0; k < nn[i]; ++k) {
                uint32_t j = rows[(size_t)i * c->max_neigh + k] & kIndexMask;  // drop the class bits
                if (j > i) js.push_back(j);
            }
            std::sort(js.begin(), js.end());  // rows hold near pairs first
            for (uint32_t j : js) {
                if (cnt < max_pairs) { pairs[2 * cnt] = (int32_t)i; pairs[2 * cnt + 1] = (int32_t)j; }
                ++cnt;
            }
        }
        if (npairs) *npairs = cnt;
        if (cnt > max_pairs) rc = fail(c, "oxdna_get_nlist_pairs: buffer too small");
    });
    return g != 0 ? g : rc;
}

int oxdna_set_langevin(oxdna_ctx* c, double dt, double kT, double gamma, double gamma_r, uint64_t seed) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (!(dt > 0.0) || !(kT >= 0.0) || !(gamma >= 0.0) || !(gamma_r >= 0.0))
            throw Fail("oxdna_set_langevin: need dt > 0, kT >= 0, gamma >= 0, gamma_r >= 0");
        c->dt = dt; c->kT = kT; c->gamma = gamma; c->gamma_r = gamma_r; c->seed = seed;
        c->have_langevin = true;
        c->drop_graph();
    });
}

int oxdna_run(oxdna_ctx* c, unsigned int nsteps, unsigned int mask, int use_graph) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (nsteps == 0) return;
        if (c->n_rep) run_ens(c, nsteps, mask);
        else run(c, nsteps, mask, use_graph);
    });
}

int oxdna_get_state(oxdna_ctx* c, double* pos, double* quat, double* vel, double* angmom) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        uint32_t n = c->N;
        std::vector<double4> p(n), q(n), v(n), a(n);
        HIP_OK(hipMemcpyAsync(p.data(), c->pos.p, sizeof(double4) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(q.data(), c->quat.p, sizeof(double4) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(v.data(), c->vel.p, sizeof(double4) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipMemcpyAsync(a.data(), c->angmom.p, sizeof(double4) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        for (uint32_t i = 0; i < n; ++i) {
            if (pos) { pos[3 * i] = p[i].x; pos[3 * i + 1] = p[i].y; pos[3 * i + 2] = p[i].z; }
            if (quat) { quat[4 * i] = q[i].x; quat[4 * i + 1] = q[i].y; quat[4 * i + 2] = q[i].z; quat[4 * i + 3] = q[i].w; }
            if (vel) { vel[3 * i] = v[i].x; vel[3 * i + 1] = v[i].y; vel[3 * i + 2] = v[i].z; }
            if (angmom) { angmom[4 * i] = a[i].x; angmom[4 * i + 1] = a[i].y; angmom[4 * i + 2] = a[i].z; angmom[4 * i + 3] = a[i].w; }
        }
    });
}


int oxdna_set_replicas(oxdna_ctx* c, unsigned int n_replicas, const unsigned int* offsets, const double* boxes) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (n_replicas == 0) {
            c->n_rep = 0;
            c->have_bias = false;
            c->forces_mask = 0xffffffffu;
            return;
        }
        if (c->N == 0) throw Fail("oxdna_set_replicas: set the state first");
        if (!offsets || !boxes) throw Fail("oxdna_set_replicas: NULL array");
        if (offsets[0] != 0 || offsets[n_replicas] != c->N)
            throw Fail("oxdna_set_replicas: offsets must run from 0 to the particle count");
        uint32_t nmax = 0;
        for (uint32_t r = 0; r < n_replicas; ++r) {
            if (offsets[r + 1] <= offsets[r]) throw Fail("oxdna_set_replicas: every replica needs a particle");
            nmax = std::max<uint32_t>(nmax, offsets[r + 1] - offsets[r]);
            for (int k = 0; k < 3; ++k)
                if (!(boxes[3 * r + k] > 0.0)) throw Fail("oxdna_set_replicas: box lengths must be positive");
        }
        if (nmax > kEnsMaxN)
            throw Fail("oxdna_set_replicas: a replica has " + std::to_string(nmax) + " particles; at most " +
                       std::to_string(kEnsMaxN) + " (use one context per larger system)");
        // bonds must stay inside their replica
        std::vector<int> part(2 * (size_t)c->N);
        HIP_OK(hipMemcpyAsync(part.data(), c->bpart.p, sizeof(int) * part.size(), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        for (uint32_t r = 0; r < n_replicas; ++r)
            for (uint32_t i = offsets[r]; i < offsets[r + 1]; ++i)
                for (int s = 0; s < 2; ++s) {
                    int o = part[2 * (size_t)i + s];
                    if (o >= 0 && ((uint32_t)o < offsets[r] || (uint32_t)o >= offsets[r + 1]))
                        throw Fail("oxdna_set_replicas: bond " + std::to_string(i) + "-" + std::to_string(o) +
                                   " crosses replicas");
                }
        c->rep_off_h.assign(offsets, offsets + n_replicas + 1);
        c->rep_box_h.assign(boxes, boxes + 3 * (size_t)n_replicas);
        // largest replicas first: the workgroups that take longest start first
        std::vector<uint32_t> order(n_replicas);
        for (uint32_t r = 0; r < n_replicas; ++r) order[r] = r;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
        });
        c->rep_off.release(); c->rep_order.release(); c->rep_box.release(); c->rep_energy.release();
        c->rep_bias.release();
        c->rep_off.alloc(n_replicas + 1);
        c->rep_order.alloc(n_replicas);
        c->rep_box.alloc(3 * (size_t)n_replicas);
        c->rep_energy.alloc(ENS_SLOTS * (size_t)n_replicas);
        HIP_OK(hipMemcpyAsync(c->rep_off.p, offsets, sizeof(uint32_t) * (n_replicas + 1), hipMemcpyHostToDevice,
                              c->stream));
        HIP_OK(hipMemcpyAsync(c->rep_order.p, order.data(), sizeof(uint32_t) * n_replicas, hipMemcpyHostToDevice,
                              c->stream));
        HIP_OK(hipMemcpyAsync(c->rep_box.p, boxes, sizeof(double) * 3 * n_replicas, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemsetAsync(c->rep_energy.p, 0, sizeof(double) * ENS_SLOTS * n_replicas, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->n_rep = n_replicas;
        c->rep_nmax = nmax;
        c->have_bias = false;
        c->ens_nt = 0;
        c->ens_div = 1;
        c->ens_tuned_mask = 0xffffffffu;
        c->forces_mask = 0xffffffffu;
        c->drop_graph();
    });
}

int oxdna_set_replica_bias(oxdna_ctx* c, const int32_t* ends, const double* k, const double* d0) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (c->n_rep == 0) throw Fail("oxdna_set_replica_bias: set the replicas first");
        if (!ends || !k || !d0) {
            c->have_bias = false;
            c->forces_mask = 0xffffffffu;
            return;
        }
        std::vector<EnsBias> b(c->n_rep);
        for (uint32_t r = 0; r < c->n_rep; ++r) {
            int n = (int)(c->rep_off_h[r + 1] - c->rep_off_h[r]);
            if (ends[2 * r] < 0 || ends[2 * r] >= n || ends[2 * r + 1] < 0 || ends[2 * r + 1] >= n ||
                ends[2 * r] == ends[2 * r + 1])
                throw Fail("oxdna_set_replica_bias: replica " + std::to_string(r) + " has invalid end indices");
            if (!(k[r] >= 0.0) || !std::isfinite(d0[r]))
                throw Fail("oxdna_set_replica_bias: need k >= 0 and a finite d0");
            b[r] = EnsBias{k[r], d0[r], ends[2 * r], ends[2 * r + 1]};
        }
        c->rep_bias.alloc(c->n_rep);
        HIP_OK(hipMemcpyAsync(c->rep_bias.p, b.data(), sizeof(EnsBias) * c->n_rep, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        c->have_bias = true;
        c->forces_mask = 0xffffffffu;
    });
}

int oxdna_get_replica_energies(oxdna_ctx* c, double* out) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (c->n_rep == 0) throw Fail("oxdna_get_replica_energies: no replicas");
        if (!out) throw Fail("oxdna_get_replica_energies: out is NULL");
        HIP_OK(hipMemcpyAsync(out, c->rep_energy.p, sizeof(double) * ENS_SLOTS * c->n_rep, hipMemcpyDeviceToHost,
                              c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    });
}

int oxdna_time_forces(oxdna_ctx* c, unsigned int mask, unsigned int nrep, double* avg_ms) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (nrep == 0) throw Fail("oxdna_time_forces: nrep must be > 0");
        if (c->n_rep) {
            // replicas: one forces-only launch evaluating nrep times
            uint32_t m = effective_mask(c, mask);
            ensure_ready_ens(c, m);
            autotune_ens(c, m);
            hipEvent_t e0, e1;
            HIP_OK(hipEventCreate(&e0));
            HIP_OK(hipEventCreate(&e1));
            EnsArgs a = ens_args(c, m, 0, nrep, 0);
            launch_ens(c, c->ens_nt, false, ens_args(c, m, 0, 1, 0));
            HIP_OK(hipEventRecord(e0, c->stream));
            launch_ens(c, c->ens_nt, false, a);
            HIP_OK(hipEventRecord(e1, c->stream));
            HIP_OK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            *avg_ms = (double)ms / nrep;
            return;
        }
        ensure_ready(c);
        uint32_t m = effective_mask(c, mask);
        if (c->forces_mask != m || !c->nlist_valid) compute(c, m, 1);
        // one event pair around nrep back-to-back launches on the context's stream: the
        // average then holds no per-launch event overhead, as inside the step graph
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        launch_forces(c, m, 0, false, false, false);  // warm caches and clocks
        HIP_OK(hipEventRecord(e0, c->stream));
        for (uint32_t r = 0; r < nrep; ++r) launch_forces(c, m, 0, false, false, false);
        HIP_OK(hipEventRecord(e1, c->stream));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        double tot = ms;
        *avg_ms = tot / nrep;
    });
}

int oxdna_time_steps(oxdna_ctx* c, unsigned int mask, unsigned int nsteps, double* avg_ms) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (!avg_ms) throw Fail("oxdna_time_steps: avg_ms is NULL");
        if (c->n_rep) throw Fail("oxdna_time_steps: one-system contexts only (replicas: oxdna_time_forces)");
        if (nsteps < 1 + kGraphSteps)
            throw Fail("oxdna_time_steps: nsteps must be at least " + std::to_string(1 + kGraphSteps));
        if (!c->ev0) HIP_OK(hipEventCreate(&c->ev0));
        if (!c->ev1) HIP_OK(hipEventCreate(&c->ev1));
        c->timed_ms = 0.0;
        c->timed_steps = 0;
        c->timing = true;
        try {
            run(c, nsteps, mask, 1);
        } catch (...) {
            c->timing = false;
            throw;
        }
        c->timing = false;
        if (c->timed_steps == 0) throw Fail("oxdna_time_steps: no graph-launched step was timed");
        *avg_ms = c->timed_ms / (double)c->timed_steps;
    });
}

int oxdna_count_evaluations(oxdna_ctx* c, unsigned int mask, uint64_t out[14]) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (c->n_rep) {
            // replicas: every unordered pair and every bond once
            uint32_t m = effective_mask(c, mask);
            ensure_ready_ens(c, m);
            DBuf<unsigned long long> d;
            d.alloc(14);
            HIP_OK(hipMemsetAsync(d.p, 0, 14 * sizeof(unsigned long long), c->stream));
            EnsArgs a = ens_args(c, m, 0, 1, 0);
            a.counts = d.p;
            launch_ens(c, 512, true, a);
            unsigned long long h[14];
            HIP_OK(hipMemcpyAsync(h, d.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
            d.release();
            for (int k = 0; k < 14; ++k) out[k] = h[k];
            return;
        }
        ensure_ready(c);
        uint32_t m = effective_mask(c, mask);
        if (c->forces_mask != m || !c->nlist_valid) compute(c, m, 1);
        DBuf<unsigned long long> d;
        d.alloc(14);
        HIP_OK(hipMemsetAsync(d.p, 0, 14 * sizeof(unsigned long long), c->stream));
        ForceArgs a = force_args(c, m, 0, 0);
        a.counts = d.p;
        launch_k<false, false, true>(c->tpp, dim3(force_blocks(c, c->tpp)), dim3(kBlock), 0, c->stream, a);
        unsigned long long h[14];
        HIP_OK(hipMemcpyAsync(h, d.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        d.release();
        for (int k = 0; k < 14; ++k) out[k] = h[k];
    });
}

int oxdna_synchronize(oxdna_ctx* c) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] { HIP_OK(hipStreamSynchronize(c->stream)); });
}

int oxdna_get_stats(oxdna_ctx* c, uint64_t* n_builds, uint32_t* max_neigh, uint64_t* step) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        DevFlags f;
        read_flags(c, f);
        if (n_builds) *n_builds = f.builds;
        if (max_neigh) *max_neigh = c->max_neigh;
        if (step) *step = c->n_rep ? c->host_step : f.step;
    });
}

int oxdna_set_lanes_per_particle(oxdna_ctx* c, int lanes) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (lanes != 0 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32)
            throw Fail("oxdna_set_lanes_per_particle: 0 (autotune), 4, 8, 16 or 32");
        if (lanes != 0 && lds_bytes(c, lanes) > kMaxLds)
            throw Fail("oxdna_set_lanes_per_particle: neighbour rows too long for that many particles per block");
        c->fixed_tpp = lanes;
        if (lanes) c->tpp = lanes;
        c->tuned_mask = 0xffffffffu;
        c->drop_graph();
    });
}

int oxdna_set_replica_tiling(oxdna_ctx* c, int workgroup, int depth_divisor) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (workgroup != 0 && workgroup != 64 && workgroup != 128 && workgroup != 256 && workgroup != 512)
            throw Fail("oxdna_set_replica_tiling: workgroup 0 (autotune), 64, 128, 256 or 512");
        if (workgroup != 0 && depth_divisor != 1 && depth_divisor != 2 && depth_divisor != 4)
            throw Fail("oxdna_set_replica_tiling: depth_divisor 1, 2 or 4");
        c->fixed_ens_nt = workgroup;
        c->fixed_ens_div = workgroup ? (uint32_t)depth_divisor : 1u;
        c->ens_nt = 0;
        c->ens_div = 1;
        c->ens_tuned_mask = 0xffffffffu;
        c->drop_graph();
    });
}

int oxdna_get_tuning(oxdna_ctx* c, int* lanes_per_particle, int* replica_workgroup) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (lanes_per_particle) *lanes_per_particle = c->tpp;
        if (replica_workgroup) *replica_workgroup = c->ens_nt;
    });
}

int oxdna_get_tuning_ex(oxdna_ctx* c, oxdna_tuning* out) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        if (!out) throw Fail("oxdna_get_tuning_ex: out is NULL");
        out->lanes_per_particle = c->tpp;
        out->particles_per_workgroup = kBlock / c->tpp;
        out->replica_workgroup = c->ens_nt;
        out->replica_records_per_pass = (c->n_rep && c->ens_nt) ? (int)ens_shape(c, c->ens_nt, c->ens_div).cap : 0;
        out->nlist_buffer = c->buffer;
    });
}

int oxdna_tune_nlist_buffer(oxdna_ctx* c, unsigned int mask, const double* buffers, unsigned int n_buffers,
                            unsigned int steps, double* steps_per_s, double* chosen) {
    if (!c) return fail(nullptr, "NULL context");
    return guard(c, [&] {
        std::vector<double> cand;
        if (buffers) cand.assign(buffers, buffers + n_buffers);
        else
            for (int k = 1; k <= 8; ++k) cand.push_back(0.1 * k);
        const double b = tune_buffer(c, mask, cand, steps, std::min<uint32_t>(std::max<uint32_t>(steps / 4, 1), 400),
                                     steps_per_s);
        if (chosen) *chosen = b;
    });
}

}  // extern "C"

// ------------------------------------------------------------------ stateless kernel drivers
namespace {
// Device workspace of one stateless driver call (parameter block, frames, energy
// partials, bond tables): one stream-ordered allocation on the caller's stream from a
// memory pool the library keeps per device (its memory stays reserved between calls,
// so an allocation after the first costs no device allocation), returned to the pool
// on the same stream when the call's launches are enqueued. Calls share no device
// memory: concurrent calls on one device, from other streams or threads, are safe.
hipMemPool_t g_pool[64] = {};
std::mutex g_pool_mu;

hipMemPool_t driver_pool(int dev) {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (!g_pool[dev]) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        HIP_OK(hipMemPoolCreate(&pool, &props));
        uint64_t keep = ~0ull;  // never trim the pool at synchronisation
        HIP_OK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
        g_pool[dev] = pool;
    }
    return g_pool[dev];
}

struct Workspace {
    unsigned char* p = nullptr;
    hipStream_t st = nullptr;
    size_t used = 0;
    Workspace(size_t bytes, hipStream_t stream) : st(stream) {
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        if (dev < 0 || dev >= 64) throw Fail("device index out of range");
        HIP_OK(hipMallocFromPoolAsync((void**)&p, std::max<size_t>(bytes, 256), driver_pool(dev), st));
    }
    ~Workspace() {
        if (p) (void)hipFreeAsync(p, st);
    }
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    // the next 256-byte-aligned piece of n elements of T
    template<class T> T* take(size_t n) {
        T* q = reinterpret_cast<T*>(p + used);
        used += (n * sizeof(T) + 255) & ~size_t(255);
        return q;
    }
};

constexpr size_t ws_bytes(size_t n, size_t elem) { return (n * elem + 255) & ~size_t(255); }

void put_params(const std::vector<uint4>& blob, uint4* dst, hipStream_t st);

// Lanes per particle of the stateless drivers' force launches.
constexpr int kDriverTpp = 8;

// Checks the arguments of a stateless pair-driver call and builds its parameter block
// (DParams, then the extras block of its per-type-pair sets).
std::vector<uint4> pair_driver_params(const char* who, int term, unsigned int N, float* d_force, float* d_torque,
                                      const double* d_pos, const double* d_orientation, const double* d_charge,
                                      const oxdna_box* box, const unsigned int* d_n_neigh, const unsigned int* d_nlist,
                                      const size_t* d_head_list, int ntypes, const void* params,
                                      const double* rcutsq) {
    const std::string w(who);
    if (term < 0 || term >= OXDNA_TERM_BONDED) throw Fail(w + ": not a nonbonded term id");
    if (ntypes < 1 || ntypes > 4) throw Fail(w + ": ntypes must be 1..4 (the oxDNA type set A, G, C, T/U)");
    if (!params || !rcutsq || !box) throw Fail(w + ": NULL parameter table, cut-offs or box");
    DParams P{};
    std::vector<unsigned char> xset[6];
    uint8_t pidx[6][16] = {};
    load_pair_params(P, xset[term], pidx[term], term, ntypes, params);
    for (int k = 0; k < ntypes * ntypes; ++k) {
        if (!(rcutsq[k] >= 0.0)) throw Fail(w + ": negative or NaN squared cut-off");
        P.rcutsq[term][k] = (float)rcutsq[k];
    }
    P.ntypes = ntypes;
    P.mask = 1u << term;
    std::vector<uint4> blob = param_blob(P, xset, pidx);
    if (N == 0) return blob;
    if (!d_force || !d_torque || !d_pos || !d_orientation || !d_n_neigh || !d_nlist || !d_head_list)
        throw Fail(w + ": NULL device array");
    if (term == OXDNA_TERM_DEBYE && !d_charge) throw Fail(w + ": Debye-Hueckel needs charges");
    if (N > kIndexMask) throw Fail(w + ": N must be below 2^30 (neighbour indices keep 30 bits)");
    return blob;
}

// Frames, the force launch over the caller's rows and the energy reduction of one
// stateless pair-driver call, on the given device workspace (dP: the parameter block
// on the device, xwords its extras; fr0..fr2: N float4 each; epart: 8 doubles per force
// workgroup).
void pair_driver_launch(int term, float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                        unsigned int N, const double* d_pos, const double* d_orientation, const double* d_charge,
                        const oxdna_box* box, const unsigned int* d_n_neigh, const unsigned int* d_nlist,
                        const size_t* d_head_list, const DParams* dP, uint32_t xwords, float4* fr0, float4* fr1,
                        float4* fr2, double* epart, double* d_term_energy, hipStream_t st) {
    hipLaunchKernelGGL(k_frames, dim3(blocks_for(N)), dim3(kBlock), 0, st, N, (const double4*)d_orientation, d_charge,
                       (const double4*)d_pos, fr0, fr1, fr2);
    const uint32_t blocks = blocks_for((size_t)N * kDriverTpp);
    ForceArgs a{};
    a.N = N;
    a.pos = (const double4*)d_pos;
    a.fr0 = fr0; a.fr1 = fr1; a.fr2 = fr2;
    // the caller's rows are only read (no rebuild without a flags block)
    a.n_neigh = const_cast<uint32_t*>(d_n_neigh);
    a.nlist = const_cast<uint32_t*>(d_nlist);
    a.head = d_head_list;
    a.P = dP;
    for (int k = 0; k < 3; ++k) { a.L[k] = box->L[k]; a.Linv[k] = 1.0 / box->L[k]; }
    a.mask = 1u << term;
    a.xwords = xwords;
    a.force = (float4*)d_force; a.torque = (float4*)d_torque;
    a.virial = d_virial; a.vpitch = virial_pitch;
    a.epart = epart;
    const bool we = d_term_energy != nullptr, wv = d_virial != nullptr;
    launch_k_any(kDriverTpp, we, wv, dim3(blocks), dim3(kBlock), 0, st, a);
    if (we) hipLaunchKernelGGL(k_reduce_energy, dim3(1), dim3(256), 0, st, blocks, epart, d_term_energy, 0);
    HIP_OK(hipGetLastError());
}

// ------------------------------------------------------------------ no-allocation driver
// oxdna_pair_forces_noalloc runs on one fixed device array instead of scratch the
// library allocates: the parameter block, three frame arrays of kNoAllocMaxN float4 and
// the per-workgroup energy partials, 3.3 MiB in all, which each device that runs the
// library holds until the process ends. Every call on a device shares its copy (the
// last call's contents stay): calls that overlap on one device race on it.
constexpr uint32_t kNoAllocMaxN = 65536;
constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kNaParams = 0;
constexpr size_t kNaFrames = align256(sizeof(DParams) + kMaxExtraBytes);
constexpr size_t kNaEpart = kNaFrames + align256(3 * (size_t)kNoAllocMaxN * sizeof(float4));
constexpr size_t kNaBytes =
    kNaEpart + align256(((size_t)kNoAllocMaxN * kDriverTpp + kBlock - 1) / kBlock * 8 * sizeof(double));

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// The device address of `scratch` on each device (each holds its own copy), resolved
// once per device with that device current. Host code uses these, never the array's
// name (a host-side stand-in).
unsigned char* g_scratch_ptr[64] = {};
std::mutex g_scratch_mu;

unsigned char* noalloc_scratch() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw Fail("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    return g_scratch_ptr[dev];
}

// The parameter block goes to the device as kernel arguments (2 KiB per launch: two
// for DParams, up to eight more for per-type-pair sets), so a call makes no host-to-
// device copy from pageable memory.
struct ParamChunk {
    uint4 w[128];
};
static_assert(sizeof(DParams) % sizeof(uint4) == 0, "DParams is a whole number of 16-byte words");

__global__ void k_put_params(ParamChunk h, uint4* dst, uint32_t nwords) {
    for (uint32_t k = threadIdx.x; k < nwords; k += blockDim.x) dst[k] = h.w[k];
}

void put_params(const std::vector<uint4>& blob, uint4* dst, hipStream_t st) {
    if (blob.size() * sizeof(uint4) > sizeof(DParams) + kMaxExtraBytes) throw Fail("internal: parameter block too large");
    const uint32_t words = (uint32_t)blob.size();
    for (uint32_t h0 = 0; h0 < words; h0 += 128) {
        ParamChunk h{};
        const uint32_t n = std::min<uint32_t>(128u, words - h0);
        std::memcpy(h.w, blob.data() + h0, n * sizeof(uint4));
        hipLaunchKernelGGL(k_put_params, dim3(1), dim3(128), 0, st, h, dst + h0, n);
    }
}

__global__ void k_bond_table(uint32_t N, const uint32_t* blist, const uint32_t* bpos, uint32_t pitch,
                             const uint32_t* n_bonds, int* bpart, int* binfo, uint32_t* err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint32_t nb = n_bonds[i];
    if (nb > 2) atomicOr(err, 1u);
    for (uint32_t b = 0; b < 2; ++b) {
        if (b < nb) {
            size_t k = (size_t)b * pitch + i;
            bpart[2 * i + b] = (int)blist[2 * k];
            binfo[2 * i + b] = (int)((bpos[k] & 1u) | (blist[2 * k + 1] << 1));
        } else {
            bpart[2 * i + b] = -1;
            binfo[2 * i + b] = -1;
        }
    }
}
}  // namespace

extern "C" {

int oxdna_pair_forces(int term, float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                      unsigned int N, const double* d_pos, const double* d_orientation, const double* d_charge,
                      const oxdna_box* box, const unsigned int* d_n_neigh, const unsigned int* d_nlist,
                      const size_t* d_head_list, int ntypes, const void* params, const double* rcutsq,
                      double* d_term_energy, void* stream) {
    return guard(nullptr, [&] {
        std::vector<uint4> blob = pair_driver_params("oxdna_pair_forces", term, N, d_force, d_torque, d_pos,
                                                     d_orientation, d_charge, box, d_n_neigh, d_nlist, d_head_list,
                                                     ntypes, params, rcutsq);
        if (N == 0) return;
        const uint32_t xwords = reinterpret_cast<const DParams*>(blob.data())->xwords;
        hipStream_t st = (hipStream_t)stream;
        const size_t nep = (size_t)blocks_for((size_t)N * kDriverTpp) * 8;
        Workspace ws(ws_bytes(blob.size(), sizeof(uint4)) + 3 * ws_bytes(N, sizeof(float4)) +
                         ws_bytes(nep, sizeof(double)), st);
        uint4* dP = ws.take<uint4>(blob.size());
        float4* fr0 = ws.take<float4>(N);
        float4* fr1 = ws.take<float4>(N);
        float4* fr2 = ws.take<float4>(N);
        double* epart = ws.take<double>(nep);
        put_params(blob, dP, st);
        pair_driver_launch(term, d_force, d_torque, d_virial, virial_pitch, N, d_pos, d_orientation, d_charge, box,
                           d_n_neigh, d_nlist, d_head_list, (const DParams*)dP, xwords, fr0, fr1, fr2, epart,
                           d_term_energy, st);
    });
}

int oxdna_pair_forces_noalloc(int term, float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                              unsigned int N, const double* d_pos, const double* d_orientation,
                              const double* d_charge, const oxdna_box* box, const unsigned int* d_n_neigh,
                              const unsigned int* d_nlist, const size_t* d_head_list, int ntypes,
                              const void* params, const double* rcutsq, double* d_term_energy, void* stream) {
    return guard(nullptr, [&] {
        const std::vector<uint4> blob =
            pair_driver_params("oxdna_pair_forces_noalloc", term, N, d_force, d_torque, d_pos, d_orientation, d_charge,
                               box, d_n_neigh, d_nlist, d_head_list, ntypes, params, rcutsq);
        if (N > kNoAllocMaxN)
            throw Fail("oxdna_pair_forces_noalloc: " + std::to_string(N) + " particles; the fixed scratch holds at most " +
                       std::to_string(kNoAllocMaxN) + " (use oxdna_pair_forces)");
        unsigned char* base = noalloc_scratch();  // resolves the array's address on this device once
        if (N == 0) return;
        hipStream_t st = (hipStream_t)stream;
        uint4* dP = (uint4*)(base + kNaParams);
        put_params(blob, dP, st);
        float4* fr = (float4*)(base + kNaFrames);
        double* epart = (double*)(base + kNaEpart);
        pair_driver_launch(term, d_force, d_torque, d_virial, virial_pitch, N, d_pos, d_orientation, d_charge, box,
                           d_n_neigh, d_nlist, d_head_list, (const DParams*)dP,
                           reinterpret_cast<const DParams*>(blob.data())->xwords, fr, fr + kNoAllocMaxN,
                           fr + 2 * (size_t)kNoAllocMaxN, epart, d_term_energy, st);
    });
}

int oxdna_bonded_forces(float* d_force, float* d_torque, float* d_virial, size_t virial_pitch,
                        double* d_split_energy, unsigned int N, const double* d_pos, const double* d_orientation,
                        const oxdna_box* box, const unsigned int* d_blist, const unsigned int* d_bpos,
                        unsigned int pitch, const unsigned int* d_n_bonds, const oxdna_bonded_params* params,
                        unsigned int n_bond_types, void* stream) {
    return guard(nullptr, [&] {
        if (!params || n_bond_types == 0 || n_bond_types > MAX_BOND_TYPES)
            throw Fail("oxdna_bonded_forces: need 1..4 bond parameter sets");
        if (!box) throw Fail("oxdna_bonded_forces: NULL box");
        if (N == 0) return;
        if (!d_force || !d_torque || !d_pos || !d_orientation || !d_blist || !d_bpos || !d_n_bonds)
            throw Fail("oxdna_bonded_forces: NULL device array");
        if (N > kIndexMask) throw Fail("oxdna_bonded_forces: N must be below 2^30");
        hipStream_t st = (hipStream_t)stream;
        DParams P{};
        for (uint32_t k = 0; k < n_bond_types; ++k) load_bond_params(P.bond[k], params[k]);
        P.n_bond_types = (int)n_bond_types;
        P.ntypes = 4;
        P.mask = 1u << OXDNA_TERM_BONDED;
        std::vector<uint4> blob(sizeof(DParams) / sizeof(uint4));
        std::memcpy(blob.data(), &P, sizeof(DParams));
        const int tpp = 8;
        const size_t nep = (size_t)blocks_for((size_t)N * tpp) * 8;
        // parameters, frames, bond tables, the error word with the empty rows' zeros after
        // it (N + 1 words), zero row heads, energy partials
        Workspace ws(ws_bytes(blob.size(), sizeof(uint4)) + 3 * ws_bytes(N, sizeof(float4)) +
                         2 * ws_bytes(2 * (size_t)N, sizeof(int)) + ws_bytes((size_t)N + 1, sizeof(uint32_t)) +
                         ws_bytes(N, sizeof(size_t)) + ws_bytes(nep, sizeof(double)), st);
        uint4* dP = ws.take<uint4>(blob.size());
        float4* fr0 = ws.take<float4>(N);
        float4* fr1 = ws.take<float4>(N);
        float4* fr2 = ws.take<float4>(N);
        int* bpart = ws.take<int>(2 * (size_t)N);
        int* binfo = ws.take<int>(2 * (size_t)N);
        uint32_t* zeros = ws.take<uint32_t>((size_t)N + 1);  // [0]: error word, then N empty row lengths
        size_t* zhead = ws.take<size_t>(N);
        double* epart = ws.take<double>(nep);
        put_params(blob, dP, st);
        hipLaunchKernelGGL(k_frames, dim3(blocks_for(N)), dim3(kBlock), 0, st, N, (const double4*)d_orientation,
                           (const double*)nullptr, (const double4*)d_pos, fr0, fr1, fr2);
        // fill kernels, not memset nodes: a caller may capture the call into a graph
        hipLaunchKernelGGL(k_zero_u32, dim3(std::min<size_t>(blocks_for((size_t)N + 1), 4096)), dim3(kBlock), 0, st,
                           zeros, (size_t)N + 1);
        hipLaunchKernelGGL(k_zero_u32, dim3(std::min<size_t>(blocks_for(2 * (size_t)N), 4096)), dim3(kBlock), 0, st,
                           (uint32_t*)zhead, 2 * (size_t)N);
        hipLaunchKernelGGL(k_bond_table, dim3(blocks_for(N)), dim3(kBlock), 0, st, N, d_blist, d_bpos, pitch,
                           d_n_bonds, bpart, binfo, zeros);
        ForceArgs a{};
        a.N = N;
        a.pos = (const double4*)d_pos;
        a.fr0 = fr0; a.fr1 = fr1; a.fr2 = fr2;
        a.bpart = bpart; a.binfo = binfo;
        a.P = (const DParams*)dP;
        for (int k = 0; k < 3; ++k) { a.L[k] = box->L[k]; a.Linv[k] = 1.0 / box->L[k]; }
        a.mask = 1u << OXDNA_TERM_BONDED;
        a.force = (float4*)d_force; a.torque = (float4*)d_torque;
        a.virial = d_virial; a.vpitch = virial_pitch;
        a.split = (double4*)d_split_energy;
        // no nonbonded neighbours: every row is empty
        a.n_neigh = zeros + 1;
        a.nlist = zeros + 1;
        a.head = zhead;
        a.epart = epart;
        // split energies need the energy instantiation
        launch_k_any(tpp, d_split_energy != nullptr, d_virial != nullptr, dim3(blocks_for((size_t)N * tpp)),
                     dim3(kBlock), 0, st, a);
        uint32_t herr = 0;
        HIP_OK(hipMemcpyAsync(&herr, zeros, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (herr) throw Fail("oxdna_bonded_forces: a particle has more than two backbone bonds");
    });
}

}  // extern "C"
