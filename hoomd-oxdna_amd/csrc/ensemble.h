// ensemble.h — batched replicas: many independent small systems (hairpins, duplexes,
// umbrella windows) advanced together, one workgroup per replica, by one persistent
// launch that runs many Langevin steps.
//
// A replica's positions (FP64), body frames, bond table, pair tables and pair records
// live in LDS for the whole launch; thread t owns particle t (its integrator state stays
// in the state arrays, touched only by that thread, so it stays in cache). Every pair
// of the replica is enumerated once, as (i, i + d mod n) with d = 1 .. n / 2 (for even
// n the last row only for i < n / 2), so particle i is the first end of the pairs in
// its row i and the second end of (i - d, i). Each step:
//
//   test    every pair's centre distance (FP64 minimum image, as the evaluation) sorts
//           it as near (inside a cut-off of some term but Debye-Hueckel's), far (inside
//           Debye-Hueckel's only) or neither; row i's results are bits d - 1 of two
//           bit words per 32 rows (near, far) of particle i;
//   scan    each particle's near / far counts are scanned in particle order: near
//           entries fill the pair records from the front, far ones from the back, each
//           particle's entries together in d order; with one pass, every thread then
//           writes the list for one (particle, 8 rows) item at a time;
//   eval    whole wavefronts evaluate near entries, then far ones (Debye-Hueckel only),
//           then the bonds and the bias; a pair's record holds the force and torque on
//           its first end and the torque on its second (F_j = -F_i, and
//           tau_j = -tau_i + (r_j - r_i) x F_i for a rotation-invariant pair energy);
//   gather  particle t sums its first-end records (contiguous), then its second-end
//           records (for replicas of up to 128 particles from a table of incoming
//           record indices that the fill writes, else from the bit words of particle
//           t - d and its scan offset), its bond records and the bias, always in the
//           same order, so a run is bit-for-bit repeatable; then it integrates.
//
// No atomics: every sum has a fixed order. When a step's entries exceed the record
// capacity the pairs are split into passes over ranges of first ends; each pass fills,
// evaluates and gathers its own records.
#pragma once

#include <type_traits>

#include "integrator.h"
#include "kernels.h"

namespace oxd {

// Harmonic bias k (|r_e - r_s| - d0)^2 on the centres of two particles of a replica
// (indices local to the replica; k = 0 disables it).
struct EnsBias {
    double k, d0;
    int s, e;
};

enum { S_BIAS = 8, ENS_SLOTS = 9 };

struct EnsArgs {
    uint32_t n_rep;
    const uint32_t* offsets;  // n_rep + 1 particle offsets
    const uint32_t* order;    // workgroup b runs replica order[b]
    const double* boxes;      // 3 box lengths per replica
    double4* pos;             // x, y, z, type bits (.w)
    double4* vel;
    double4* quat;
    double4* angmom;
    double4* bd;
    double4* bd_tq;
    const double* charge;
    const int* bpart;         // 2 per particle: global index of the partner or -1
    const int* binfo;         // 2 per particle: bit0 role (0: A), bits 1.. bond type
    const DParams* P;
    uint32_t mask;
    float4* force;            // forces of the final state (.w: half energy when want_e)
    float4* torque;
    double* rep_energy;       // ENS_SLOTS per replica (want_e)
    const EnsBias* bias;      // one per replica, or NULL
    Langevin lv;
    unsigned long long seed;  // replica r draws with key seed + r and particle counter t
    unsigned long long step0; // integration steps done before this launch
    uint32_t nsteps;          // Langevin steps of this launch (0: forces only)
    uint32_t nevals;          // forces-only launches: evaluate this many times (timing)
    int want_e;               // energies of the final evaluation
    float rnear2;             // pairs closer than this (squared) get every term ...
    float rfar2;              // ... pairs closer than this only Debye-Hueckel (0: no such term)
    unsigned long long* counts;  // COUNT instantiation: per term (inside cut-off, nonzero)
    uint32_t ncap;            // particle capacity of the LDS layout (>= every replica)
    uint32_t cap;             // pair records per pass (>= 2 (ncap / 2) + 64)
};

// Bit words per particle: one per 32 rows d.
__host__ __device__ inline uint32_t ens_words(uint32_t ncap) {
    const uint32_t d = ncap / 2;
    return d > 32 ? (d + 31) / 32 : 1u;
}

// Byte offsets of the dynamic LDS arrays of one replica (ncap particles, cap records).
struct EnsLayout {
    uint32_t par, xy, z, f0, f1, f2, bond, brf, brt, bits, pt, gl, inc, acc, pass, red, ctl, rf, rt, rj, bytes;
};

// Replicas of up to this many particles keep a table of incoming record indices (one
// 16-bit entry per pair, 16 KB at 128); larger ones find them from the bit words.
constexpr uint32_t kEnsIncMax = 128;

// staged: the parameter block (DParams) is copied to the start of the layout.
__host__ __device__ inline EnsLayout ens_layout(uint32_t ncap, uint32_t cap, uint32_t nt, bool staged) {
    EnsLayout l{};
    uint32_t o = 0;
    auto take = [&o](uint32_t bytes) {
        const uint32_t at = o;
        o += (bytes + 15u) & ~15u;
        return at;
    };
    l.par = take(staged ? (uint32_t)sizeof(DParams) : 0u);  // DParams, read by every pair
    l.xy = take(16 * ncap);                     // double2 x, y
    l.z = take(8 * ncap);                       // double z
    l.f0 = take(16 * ncap);                     // float4 a1 + charge
    l.f1 = take(16 * ncap);                     // float4 a2 + type bits
    l.f2 = take(16 * ncap);                     // float4 a3
    l.bond = take(16 * ncap);                   // int4 {partner 0, partner 1, info 0, info 1}
    l.brf = take(32 * ncap);                    // float4 [2][ncap]: bond force + half energy
    l.brt = take(32 * ncap);                    // float4 [2][ncap]: bond torque
    l.bits = take(8 * ens_words(ncap) * ncap);  // uint2 [W][ncap]: near, far bits of rows 32 w + 1 ..
    l.pt = take(8 * ncap);                      // uint2 [ncap]: {near | far << 16 offset, pass}
    l.gl = take(8 * ncap);                      // uint2 [ncap]: first near / far entry of the step
    l.inc = take(ncap <= kEnsIncMax ? 2 * (ncap / 2) * ncap : 0);  // u16 [d - 1][j]: record of (j - d, j)
    l.acc = take(32 * ncap);                    // float4 [2][ncap]: force + energy, torque of earlier passes
    l.pass = take(8 * (ncap + 2));              // uint2 [passes + 1]: near / far scan bases
    l.red = take(8 * ENS_SLOTS * (nt / 64));    // per-wavefront energy sums
    l.ctl = take(64);                           // pass count, eval iteration counter, bias record
    // pair records, and one record of zeros at index cap
    l.rf = take(16 * (cap + 1));                // float4: F on the first end + pair energy
    l.rt = take(16 * (cap + 1));                // float4: torque on the first end (.w: list entry)
    l.rj = take(16 * (cap + 1));                // float4: torque on the second end
    l.bytes = o;
    return l;
}

// Launches k_ensemble<nt, count, multi> with one workgroup per replica and lds bytes
// of dynamic LDS (count and multi: nt = 512 only); hipErrorInvalidValue for a shape
// without an instantiation. Defined in ensemble.hip.
hipError_t launch_ensemble(int nt, bool count, bool multi, uint32_t n_rep, uint32_t lds, hipStream_t st,
                           const EnsArgs& a);

// Data-parallel-primitive (DPP) lane moves within a wavefront: no LDS round trip.
template<int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}
template<int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Wavefront inclusive prefix sum, lanes in order (rows of 16 by shifts 1, 2, 4, 8, then
// the row totals carried by the row broadcasts of lanes 15 and 31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += dpp_u32<0x111, 0xf>(v);  // row_shr:1
    v += dpp_u32<0x112, 0xf>(v);  // row_shr:2
    v += dpp_u32<0x114, 0xf>(v);  // row_shr:4
    v += dpp_u32<0x118, 0xf>(v);  // row_shr:8
    v += dpp_u32<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_u32<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Sum over the wavefront, in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
}

// Sum over aligned groups of G = 1, 2, 4 or 8 lanes, in every lane of the group: a
// butterfly (pairs, then pairs of pairs, then the two halves of 8), so every lane of a
// group adds the same values in the same order.
__device__ __forceinline__ float group_sum(float v, uint32_t G) {
    if (G >= 2) v += dpp_f32<0xb1>(v);   // quad_perm [1, 0, 3, 2]
    if (G >= 4) v += dpp_f32<0x4e>(v);   // quad_perm [2, 3, 0, 1]
    if (G >= 8) v += dpp_f32<0x141>(v);  // row_half_mirror
    return v;
}

// MULTI: the general instantiation, for tables that differ between type pairs and for
// replicas too large for a layout with the parameter block: it reads the block in global
// memory, each pair finding its sets through their offsets in it. Otherwise the block is
// copied to LDS and the base sets sit at fixed offsets.
template<int NT, bool COUNT, bool MULTI>
__global__ __launch_bounds__(NT) void k_ensemble(EnsArgs a) {
    static_assert(NT % 64 == 0 && NT <= 512, "whole wavefronts");
    constexpr int NW = NT / 64;
    // A grid of exactly one workgroup per replica, checked against the grid size: reading
    // the grid size also gives the kernel the hidden-argument block that every other
    // kernel of the library has (without it the argument segment ends at EnsArgs, and a
    // tool that reads a dispatch's hidden arguments reads past its end).
    if (gridDim.x != a.n_rep) return;
    extern __shared__ uint4 ens_lds[];
    unsigned char* lds = reinterpret_cast<unsigned char*>(ens_lds);
    const uint32_t ncap = a.ncap, cap = a.cap;
    const EnsLayout lo = ens_layout(ncap, cap, NT, !MULTI);
    double2* Sxy = reinterpret_cast<double2*>(lds + lo.xy);
    double* Sz = reinterpret_cast<double*>(lds + lo.z);
    float4* Sf0 = reinterpret_cast<float4*>(lds + lo.f0);
    float4* Sf1 = reinterpret_cast<float4*>(lds + lo.f1);
    float4* Sf2 = reinterpret_cast<float4*>(lds + lo.f2);
    int4* Sbond = reinterpret_cast<int4*>(lds + lo.bond);
    float4* Sbrf = reinterpret_cast<float4*>(lds + lo.brf);
    float4* Sbrt = reinterpret_cast<float4*>(lds + lo.brt);
    uint2* Sbits = reinterpret_cast<uint2*>(lds + lo.bits);
    uint2* Spt = reinterpret_cast<uint2*>(lds + lo.pt);
    uint2* Sgl = reinterpret_cast<uint2*>(lds + lo.gl);
    uint16_t* Sinc = reinterpret_cast<uint16_t*>(lds + lo.inc);
    float4* Sacc = reinterpret_cast<float4*>(lds + lo.acc);
    uint2* Spass = reinterpret_cast<uint2*>(lds + lo.pass);
    double* Sred = reinterpret_cast<double*>(lds + lo.red);
    uint32_t* Sctl = reinterpret_cast<uint32_t*>(lds + lo.ctl);  // [0] passes
    float4* Sbias = reinterpret_cast<float4*>(lds + lo.ctl + 16);  // force on e, energy
    float4* Rf = reinterpret_cast<float4*>(lds + lo.rf);
    float4* Rt = reinterpret_cast<float4*>(lds + lo.rt);
    float4* Rj = reinterpret_cast<float4*>(lds + lo.rj);
    const bool has_inc = ncap <= kEnsIncMax;

    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint32_t r = a.order[blockIdx.x];
    const uint32_t off = a.offsets[r];
    const uint32_t n = a.offsets[r + 1] - off;  // <= min(NT, ncap), checked by the host
    if (n == 0) {  // the whole workgroup leaves: no barrier is skipped
        if (a.want_e && t < ENS_SLOTS) a.rep_energy[(size_t)r * ENS_SLOTS + t] = 0.0;
        return;
    }
    const uint32_t D = n / 2;                   // rows d = 1 .. D
    const uint32_t Wn = D > 32 ? (D + 31) / 32 : 1u;
    // particle p belongs to the group of G lanes p G .. p G + G - 1 (G a power of two, at
    // most 8, the groups within the first 256 threads, so the group leaders that
    // integrate lie in at most four wavefronts: one per SIMD); the leader owns p
    uint32_t G = 1;
    while (G < 8u && 2u * G * n <= min(256u, (uint32_t)NT)) G *= 2u;
    const uint32_t p = t / G, q = t % G;
    const bool grp = p < n;
    const bool leader = grp && q == 0;
    const double L[3] = {a.boxes[3 * r], a.boxes[3 * r + 1], a.boxes[3 * r + 2]};
    const double Linv[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    // uniform tables: the parameter block in LDS, so a cut-off or parameter read is an
    // LDS read, not a global round trip in every branch of the pair terms; MULTI reads
    // it (and the per-type-pair sets after it) in global memory
    if constexpr (!MULTI) {
        uint4* Spar = reinterpret_cast<uint4*>(lds + lo.par);
        for (uint32_t w = t; w < (uint32_t)(sizeof(DParams) / 16); w += NT)
            Spar[w] = reinterpret_cast<const uint4*>(a.P)[w];
        __syncthreads();
    }
    const DParams& P = MULTI ? *a.P : *reinterpret_cast<const DParams*>(lds + lo.par);
    const uint32_t mask = a.mask;
    const bool bonded = (mask & (1u << T_BONDED)) != 0u;
    const int nt = P.ntypes;
    const unsigned long long key = a.seed + r;
    EnsBias bias{0.0, 0.0, 0, 0};
    if (a.bias) bias = a.bias[r];
    const uint32_t has_bias = (bias.k != 0.0 && n > 1) ? 1u : 0u;
    const uint32_t g = off + p;

    // ---- load the replica; the leader keeps its particle's charge and type bits
    if (t == 0) {
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        Rf[cap] = z4; Rt[cap] = z4; Rj[cap] = z4;
    }
    float charge = 1.f, tbits = 0.f;
    if (leader) {
        const double4 x = a.pos[g];
        Sxy[p] = make_double2(x.x, x.y);
        Sz[p] = x.z;
        const double4 qv = a.quat[g];
        charge = a.charge ? (float)a.charge[g] : 1.f;
        tbits = __int_as_float(type_of(x.w));
        float4 f0, f1, f2;
        quat_axes(qv.x, qv.y, qv.z, qv.w, f0, f1, f2);
        f0.w = charge; f1.w = tbits; f2.w = 0.f;
        Sf0[p] = f0; Sf1[p] = f1; Sf2[p] = f2;
        int4 b;
        const int o0 = a.bpart[2 * g], o1 = a.bpart[2 * g + 1];
        b.x = o0 >= 0 ? o0 - (int)off : -1;
        b.y = o1 >= 0 ? o1 - (int)off : -1;
        b.z = a.binfo[2 * g];
        b.w = a.binfo[2 * g + 1];
        Sbond[p] = b;
        // first half of the launch's first step, with the stored force, torque and
        // Langevin force / torque of the previous second half
        if (a.nsteps > 0) {
            double4 xs = make_double4(x.x, x.y, x.z, 0.0);
            double4 v = a.vel[g], qs = qv, pm = a.angmom[g];
            const float4 fo = a.force[g], to = a.torque[g];
            lv_first_half(a.lv, L, Linv, fo.x, fo.y, fo.z, to.x, to.y, to.z, a.bd[g], a.bd_tq[g], xs, v, qs, pm);
            Sxy[p] = make_double2(xs.x, xs.y);
            Sz[p] = xs.z;
            a.vel[g] = v; a.quat[g] = qs; a.angmom[g] = pm;
            quat_axes(qs.x, qs.y, qs.z, qs.w, f0, f1, f2);
            f0.w = charge; f1.w = tbits; f2.w = 0.f;
            Sf0[p] = f0; Sf1[p] = f1; Sf2[p] = f2;
        }
    }
    __syncthreads();

    constexpr uint32_t kPending = 0x80000000u;  // pass tag of a particle whose pass has not begun
    const uint32_t total = a.nsteps > 0 ? a.nsteps : max(a.nevals, 1u);
    const uint32_t nby = D > 0 ? (D + 7u) / 8u : 1u;  // bytes of row bits per particle
    const uint32_t items = n * nby;
    const uint32_t bstart = (items + 63u) & ~63u;     // the bias item starts a wavefront
    const uint32_t nbond = bonded ? n : 0u;
    const uint32_t nstep_items = bstart + has_bias;
    float et[ENS_SLOTS];
    double et_fene = 0.0;
    uint32_t cnt[14];
#pragma unroll
    for (int s = 0; s < 14; ++s) cnt[s] = 0u;

    // List entries and incoming-record entries of particle p in a pass of several (lanes
    // of its group): p's near entries start at record en, its far ones at cap - 1 - ef
    // going down; lane q writes those of rows 32 Wn q / G + 1 .. 32 Wn (q + 1) / G.
    auto fill = [&](uint32_t en, uint32_t ef) {
        const uint32_t chunk = 32u * Wn / G, lo = q * chunk, hi = lo + chunk;
        for (uint32_t w = 0; w < Wn; ++w) {
            uint2 x = Sbits[w * ncap + p];
            const uint32_t b0 = 32u * w;
            const uint32_t below = lo <= b0 ? 0u : (lo >= b0 + 32u ? 0xffffffffu : (1u << (lo - b0)) - 1u);
            const uint32_t upto = hi >= b0 + 32u ? 0xffffffffu : (hi <= b0 ? 0u : (1u << (hi - b0)) - 1u);
            en += __popc(x.x & below);
            ef += __popc(x.y & below);
            x.x &= upto & ~below;
            x.y &= upto & ~below;
            while (x.x) {
                const uint32_t d = b0 + __ffs(x.x);
                x.x &= x.x - 1u;
                if (has_inc) Sinc[(d - 1u) * ncap + (p + d < n ? p + d : p + d - n)] = (uint16_t)en;
                Rt[en++].w = __uint_as_float(p | (d << 16));
            }
            while (x.y) {
                const uint32_t d = b0 + __ffs(x.y);
                x.y &= x.y - 1u;
                const uint32_t e = cap - 1u - ef++;
                if (has_inc) Sinc[(d - 1u) * ncap + (p + d < n ? p + d : p + d - n)] = (uint16_t)e;
                Rt[e].w = __uint_as_float(p | (d << 16));
            }
        }
    };

    for (uint32_t step = 0; step < total; ++step) {
        const bool last = step + 1 == total;
        const bool want_e = last && a.want_e;
#pragma unroll
        for (int s = 0; s < ENS_SLOTS; ++s) et[s] = 0.f;
        et_fene = 0.0;
        if (t == 0) Sctl[1] = 0u;  // the eval's iteration counter (read after two barriers)

        // ---- test and bias. Test item (i, byte b) sets bits 8 (b % 4) .. of word b / 4 of
        //      particle i (rows d = 8 b + 1 .. 8 b + 8) and resets the incoming-record
        //      entries of those pairs; the last byte item of a particle zeroes the rest of
        //      its words.
        for (uint32_t u = t; u < nstep_items; u += NT) {
            if (u < items) {
                const uint32_t i = u % n, b = u / n;
                const double2 xi = Sxy[i];
                const double zi = Sz[i];
                const int4 bo = Sbond[i];
                uint32_t jv[8];
                bool ok[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint32_t d = 8u * b + k + 1u;
                    const bool valid = d <= D && !(2u * d == n && i >= d);
                    uint32_t j = i + d;
                    if (j >= n) j -= n;
                    // every entry of rows 1 .. D is reset, the never-filled second ends
                    // p < n / 2 of row n / 2 included, so the gather reads them unguarded
                    if (d <= D && has_inc) Sinc[(d - 1u) * ncap + j] = 0xffffu;
                    if (!valid) j = i;
                    jv[k] = j;
                    ok[k] = valid && (int)j != bo.x && (int)j != bo.y;
                }
                double2 xj[8];
                double zj[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) { xj[k] = Sxy[jv[k]]; zj[k] = Sz[jv[k]]; }
                uint32_t nb = 0u, fb = 0u;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    // the same arithmetic as the evaluation below, so both see one rsq
                    const V dr = mk((float)min_image(xi.x - xj[k].x, L[0], Linv[0]),
                                    (float)min_image(xi.y - xj[k].y, L[1], Linv[1]),
                                    (float)min_image(zi - zj[k], L[2], Linv[2]));
                    const float rsq = dot(dr, dr);
                    if (ok[k] && rsq < a.rnear2) nb |= 1u << k;
                    else if (ok[k] && rsq < a.rfar2) fb |= 1u << k;
                }
                unsigned char* w = reinterpret_cast<unsigned char*>(&Sbits[(b >> 2) * ncap + i]);
                w[b & 3u] = (unsigned char)nb;
                w[4u + (b & 3u)] = (unsigned char)fb;
                if (b + 1u == nby) {
                    for (uint32_t bb = b + 1u; bb < 4u * Wn; ++bb) {
                        unsigned char* z = reinterpret_cast<unsigned char*>(&Sbits[(bb >> 2) * ncap + i]);
                        z[bb & 3u] = 0u;
                        z[4u + (bb & 3u)] = 0u;
                    }
                }
            } else if (u == bstart && has_bias) {
                // harmonic bias between the centres of particles s and e
                const uint32_t bs = (uint32_t)bias.s, be = (uint32_t)bias.e;
                const double2 xs = Sxy[bs], xe = Sxy[be];
                const double dx = min_image(xe.x - xs.x, L[0], Linv[0]);
                const double dy = min_image(xe.y - xs.y, L[1], Linv[1]);
                const double dz = min_image(Sz[be] - Sz[bs], L[2], Linv[2]);
                const double dd = sqrt(dx * dx + dy * dy + dz * dz);
                const double uu = dd - bias.d0;
                const double eb = bias.k * uu * uu;
                const double c = dd > 0.0 ? -2.0 * bias.k * uu / dd : 0.0;  // force on e = c (r_e - r_s)
                Sbias[0] = make_float4((float)(c * dx), (float)(c * dy), (float)(c * dz), (float)eb);
                et[S_BIAS] += (float)eb;
            }
        }
        __syncthreads();

        // ---- scan (wavefronts holding leaders): each particle's first near / far entry
        //      in particle order, the split into passes
        if (wv * 64u < n * G) {
            uint32_t cn = 0u, cf = 0u;
            if (leader)
                for (uint32_t w = 0; w < Wn; ++w) {
                    const uint2 x = Sbits[w * ncap + p];
                    cn += __popc(x.x);
                    cf += __popc(x.y);
                }
            // near / far totals of the particles of earlier wavefronts and of all of them
            // (lane l adds particles l, l + 64, ...)
            const uint32_t pfirst = wv * 64u / G;  // first particle of this wavefront
            uint32_t pn = 0u, pf = 0u, tn = 0u, tf = 0u;
            for (uint32_t i = lane; i < n; i += 64u) {
                uint32_t c1 = 0u, c2 = 0u;
                for (uint32_t w = 0; w < Wn; ++w) {
                    const uint2 x = Sbits[w * ncap + i];
                    c1 += __popc(x.x);
                    c2 += __popc(x.y);
                }
                tn += c1; tf += c2;
                if (i < pfirst) { pn += c1; pf += c2; }
            }
            pn = wave_sum(pn); pf = wave_sum(pf);
            tn = wave_sum(tn); tf = wave_sum(tf);
            const uint32_t gn = pn + wave_incl_scan(cn) - cn;  // first near entry of particle p
            const uint32_t gf = pf + wave_incl_scan(cf) - cf;
            uint32_t mypass = 0u;
            if (tn + tf <= cap) {
                // one pass: the offsets are the scan; the list is written after the
                // barrier, by every thread
                if (t == 0) {
                    Sctl[0] = 1u;
                    Spass[0] = make_uint2(0u, 0u);
                    Spass[1] = make_uint2(tn, tf);
                }
                if (leader) Spt[p] = make_uint2(gn | (gf << 16), 0u);
            } else {
                // passes over ranges of first ends: particle p goes to pass
                // (its combined prefix) / H; a particle holds at most D entries, so no
                // pass holds more than H + D = cap entries and none is skipped
                const uint32_t H = cap - D;
                const uint32_t sc = gn + gf;
                mypass = sc / H;
                bool first = p == 0;
                if (leader && p > 0) {
                    uint32_t c = 0u;
                    for (uint32_t w = 0; w < Wn; ++w) {
                        const uint2 x = Sbits[w * ncap + p - 1];
                        c += __popc(x.x) + __popc(x.y);
                    }
                    first = (sc - c) / H != mypass;
                }
                if (leader && first) Spass[mypass] = make_uint2(gn, gf);
                if (leader && p == n - 1) {
                    Sctl[0] = mypass + 1u;
                    Spass[mypass + 1u] = make_uint2(tn, tf);
                }
                if (leader) {
                    Sgl[p] = make_uint2(gn, gf);
                    Spt[p] = make_uint2(0u, kPending | mypass);
                }
            }
        }
        __syncthreads();
        const uint32_t npass = Sctl[0];
        if (npass == 1) {
            // ---- fill (one pass): list entries and incoming-record entries, one item
            //      (particle i, rows 8 b + 1 .. 8 b + 8) per thread in turn; an entry's
            //      record is i's offset plus the entries of i's earlier rows, the order
            //      of the per-particle fill below
            for (uint32_t u = t; u < items; u += NT) {
                const uint32_t i = u % n, b = u / n, w = b >> 2, sh = 8u * (b & 3u);
                const uint2 x = Sbits[w * ncap + i];
                uint32_t nb = (x.x >> sh) & 0xffu, fb = (x.y >> sh) & 0xffu;
                if (!(nb | fb)) continue;
                const uint2 pt = Spt[i];
                uint32_t en = pt.x & 0xffffu, ef = pt.x >> 16;
                for (uint32_t w2 = 0; w2 < w; ++w2) {
                    const uint2 y = Sbits[w2 * ncap + i];
                    en += __popc(y.x);
                    ef += __popc(y.y);
                }
                const uint32_t below = (1u << sh) - 1u;
                en += __popc(x.x & below);
                ef += __popc(x.y & below);
                while (nb) {
                    const uint32_t d = 8u * b + (uint32_t)__ffs(nb);
                    nb &= nb - 1u;
                    if (has_inc) Sinc[(d - 1u) * ncap + (i + d < n ? i + d : i + d - n)] = (uint16_t)en;
                    Rt[en++].w = __uint_as_float(i | (d << 16));
                }
                while (fb) {
                    const uint32_t d = 8u * b + (uint32_t)__ffs(fb);
                    fb &= fb - 1u;
                    const uint32_t e = cap - 1u - ef++;
                    if (has_inc) Sinc[(d - 1u) * ncap + (i + d < n ? i + d : i + d - n)] = (uint16_t)e;
                    Rt[e].w = __uint_as_float(i | (d << 16));
                }
            }
            __syncthreads();
        }

        for (uint32_t pass = 0; pass < npass; ++pass) {
            const uint2 pb = Spass[pass], pe = Spass[pass + 1];
            // ---- fill (several passes; the groups of this pass's particles): offsets within
            //      the pass, list entries, incoming-record entries
            if (npass > 1) {
                const bool mine = grp && Spt[p].y == (kPending | pass);
                if (mine) {
                    const uint2 gl = Sgl[p];
                    const uint32_t en = gl.x - pb.x, ef = gl.y - pb.y;
                    fill(en, ef);
                    if (leader) Spt[p] = make_uint2(en | (ef << 16), pass);
                }
                if (t == 0) Sctl[1] = 0u;  // every wavefront left the last pass's eval
                __syncthreads();
            }

            // ---- eval: (first pass) the bonds, then near entries, then far ones, in
            //      whole-wavefront iterations. A launch that reports energies deals them to
            //      the wavefronts from the last one (the first ones lead the groups), so each
            //      thread's energy sums have a fixed order; otherwise each wavefront takes
            //      its next iteration from a counter in LDS, so those that draw cheap ones
            //      (far entries) take more. A record's place does not depend on who writes
            //      it, so the trajectory is the same either way.
            {
                const uint32_t nn = pe.x - pb.x, nf = pe.y - pb.y;
                const uint32_t In = (nn + 63u) / 64u, If = (nf + 63u) / 64u;
                const uint32_t Ib = pass == 0 ? (nbond + 63u) / 64u : 0u;
                const uint32_t Itot = Ib + In + If;
                auto take = [&]() -> uint32_t {
                    uint32_t x = 0u;
                    if (lane == 0u) x = atomicAdd(&Sctl[1], 1u);
                    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
                };
                // the bonds first: a bond's FP64 chain waits mostly on latency, which the
                // other wavefront of its SIMD then fills with pairs
                for (uint32_t it = want_e ? NW - 1u - wv : take(); it < Itot; it = want_e ? it + NW : take()) {
                  if (it < Ib) {
                    // bonds of particle A where it is the A end; each end's share goes
                    // to the record of the slot that holds the bond
                    const uint32_t A = it * 64u + lane;
                    if (A < nbond) {
                    const int4 ba = Sbond[A];
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        const int Bi = s ? ba.y : ba.x;
                        const int info = s ? ba.w : ba.z;
                        if (Bi < 0 || (info & 1) != 0) continue;
                        const uint32_t B = (uint32_t)Bi;
                        const float4 x0 = Sf0[A], x1 = Sf1[A], x2 = Sf2[A];
                        const float4 y0 = Sf0[B], y1 = Sf1[B], y2 = Sf2[B];
                        const Frame fA{mk(x0.x, x0.y, x0.z), mk(x1.x, x1.y, x1.z), mk(x2.x, x2.y, x2.z)};
                        const Frame fB{mk(y0.x, y0.y, y0.z), mk(y1.x, y1.y, y1.z), mk(y2.x, y2.y, y2.z)};
                        const double2 xa = Sxy[A], xb = Sxy[B];
                        BondRes br;
                        bond_eval(P.bond[info >> 1], min_image(xb.x - xa.x, L[0], Linv[0]),
                                  min_image(xb.y - xa.y, L[1], Linv[1]), min_image(Sz[B] - Sz[A], L[2], Linv[2]),
                                  fA, fB, __float_as_int(x1.w), __float_as_int(y1.w), br, want_e);
                        if (COUNT) { cnt[2 * T_BONDED]++; cnt[2 * T_BONDED + 1] += br.e_stack != 0.f; }
                        et_fene += br.e_fene;
                        et[S_BEXC] += br.e_bexc;
                        et[S_STACK] += br.e_stack;
                        const float eh = 0.5f * ((float)br.e_fene + br.e_bexc + br.e_stack);
                        // the slot of B that holds this bond: a B-role slot pointing back to A
                        // (the second one for A's second bond to B)
                        const int4 bb = Sbond[B];
                        const bool second = s == 1 && ba.x == Bi && (ba.z & 1) == 0;
                        const bool b0 = bb.x == (int)A && (bb.z & 1) != 0;
                        const bool b1 = bb.y == (int)A && (bb.w & 1) != 0;
                        const uint32_t sB = (b0 && !(second && b1)) ? 0u : 1u;
                        Sbrf[s * ncap + A] = make_float4(-br.FB.x, -br.FB.y, -br.FB.z, eh);
                        Sbrt[s * ncap + A] = make_float4(br.TA.x, br.TA.y, br.TA.z, 0.f);
                        Sbrf[sB * ncap + B] = make_float4(br.FB.x, br.FB.y, br.FB.z, eh);
                        Sbrt[sB * ncap + B] = make_float4(br.TB.x, br.TB.y, br.TB.z, 0.f);
                    }
                    }
                  } else {
                    const bool near = it < Ib + In;
                    const uint32_t k = (near ? it - Ib : it - Ib - In) * 64u + lane;
                    if (k < (near ? nn : nf)) {
                    const uint32_t e = near ? k : cap - 1u - k;
                    const uint32_t packed = __float_as_uint(Rt[e].w);
                    const uint32_t i = packed & 0xffffu, d = packed >> 16;
                    uint32_t j = i + d;
                    if (j >= n) j -= n;
                    const double2 xi = Sxy[i], xj = Sxy[j];
                    const V dr = mk((float)min_image(xi.x - xj.x, L[0], Linv[0]),
                                    (float)min_image(xi.y - xj.y, L[1], Linv[1]),
                                    (float)min_image(Sz[i] - Sz[j], L[2], Linv[2]));
                    const float rsq = dot(dr, dr);
                    const float4 g0 = Sf0[i], g1 = Sf1[i], g2 = Sf2[i];
                    const float4 h0 = Sf0[j], h1 = Sf1[j], h2 = Sf2[j];
                    const int tp = __float_as_int(g1.w) * nt + __float_as_int(h1.w);
                    const Frame fi{mk(g0.x, g0.y, g0.z), mk(g1.x, g1.y, g1.z), mk(g2.x, g2.y, g2.z)};
                    const Frame fj{mk(h0.x, h0.y, h0.z), mk(h1.x, h1.y, h1.z), mk(h2.x, h2.y, h2.z)};
                    const V Rc = -dr;
                    V Fp = zero3(), Tp = zero3();
                    const uint32_t m = near ? mask : (mask & (1u << T_DEBYE));
                    const float ep = pair_terms<COUNT, MULTI>(P, m, tp, rsq, Rc, fi, fj, g0.w * h0.w, Fp, Tp, et, cnt);
                    const V Tj = cross(Rc, Fp) - Tp;
                    Rf[e] = make_float4(Fp.x, Fp.y, Fp.z, ep);
                    Rt[e] = make_float4(Tp.x, Tp.y, Tp.z, 0.f);
                    Rj[e] = make_float4(Tj.x, Tj.y, Tj.z, 0.f);
                    }
                  }
                }
            }
            __syncthreads();

            // ---- gather: lane q of particle p's group adds p's first-end records
            //      q, q + G, ..., then its second-end records of rows d = q + 1, q + 1 + G,
            //      ..., eight at a time (a missing record reads the zero record); the group
            //      sums its lanes; the leader adds the bonds and the bias (first pass) and
            //      the sums of earlier passes
            V F = zero3(), T = zero3();
            float E = 0.f;
            // the leader's integrator state, loaded while the group gathers (the last pass)
            double4 v = make_double4(0.0, 0.0, 0.0, 0.0), qv = v, pm = v;
            if (leader && a.nsteps > 0 && pass + 1 == npass) {
                v = a.vel[g]; qv = a.quat[g]; pm = a.angmom[g];
            }
            if (grp) {
                const uint2 pt = Spt[p];
                if (pt.y == pass) {
                    uint32_t cn = 0u, cf = 0u;
                    for (uint32_t w = 0; w < Wn; ++w) {
                        const uint2 x = Sbits[w * ncap + p];
                        cn += __popc(x.x);
                        cf += __popc(x.y);
                    }
                    const uint32_t on = pt.x & 0xffffu, of = pt.x >> 16;
                    for (uint32_t k = q; k < cn + cf; k += 8u * G) {
                        uint32_t e[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            const uint32_t kk = k + c * G;
                            e[c] = kk < cn ? on + kk : (kk < cn + cf ? cap - 1u - (of + kk - cn) : cap);
                        }
                        float4 f[8], m[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c) { f[c] = Rf[e[c]]; m[c] = Rt[e[c]]; }
                        // all sixteen reads in flight before the first sum
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            F += mk(f[c].x, f[c].y, f[c].z);
                            T += mk(m[c].x, m[c].y, m[c].z);
                            E += 0.5f * f[c].w;
                        }
                    }
                }
                // an entry is this pass's record when its first end p - d is in this pass
                // (entries of earlier passes are stale, of later ones unset); one pass
                // skips that test and its reads
                auto second_ends = [&](auto several) {
                    for (uint32_t d0 = 1u + q; d0 <= D; d0 += 8u * G) {
                        uint32_t e[8], ps[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            const uint32_t d = min(d0 + c * G, D);
                            e[c] = Sinc[(d - 1u) * ncap + p];
                            if constexpr (decltype(several)::value) ps[c] = Spt[p >= d ? p - d : p + n - d].y;
                        }
                        // the compiler would wait for each read before issuing the next
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            bool none = d0 + c * G > D || e[c] == 0xffffu;
                            if constexpr (decltype(several)::value) none = none || ps[c] != pass;
                            if (none) e[c] = cap;
                        }
                        float4 f[8], m[8];
#pragma unroll
                        for (int c = 0; c < 8; ++c) { f[c] = Rf[e[c]]; m[c] = Rj[e[c]]; }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int c = 0; c < 8; ++c) {
                            F -= mk(f[c].x, f[c].y, f[c].z);
                            T += mk(m[c].x, m[c].y, m[c].z);
                            E += 0.5f * f[c].w;
                        }
                    }
                };
                if (has_inc && npass == 1) {
                    second_ends(std::false_type{});
                } else if (has_inc) {
                    second_ends(std::true_type{});
                } else {
                    // (replicas above the table's size) second-end records found from the
                    // bit words of particle p - d
                    for (uint32_t d = 1u + q; d <= D; d += G) {
                        uint32_t i = p + n - d;
                        if (i >= n) i -= n;
                        if (2u * d == n && i >= d) continue;  // row n / 2 holds i < n / 2 only
                        const uint32_t w = (d - 1u) >> 5, bit = (d - 1u) & 31u;
                        const uint2 x = Sbits[w * ncap + i];
                        const uint32_t isn = (x.x >> bit) & 1u, isf = (x.y >> bit) & 1u;
                        if (!(isn | isf)) continue;
                        const uint2 pi = Spt[i];
                        if (pi.y != pass) continue;
                        uint32_t rank = __popc((isn ? x.x : x.y) & ((1u << bit) - 1u));
                        for (uint32_t w2 = 0; w2 < w; ++w2) {
                            const uint2 y = Sbits[w2 * ncap + i];
                            rank += __popc(isn ? y.x : y.y);
                        }
                        const uint32_t e = isn ? (pi.x & 0xffffu) + rank : cap - 1u - ((pi.x >> 16) + rank);
                        const float4 f = Rf[e], m = Rj[e];
                        F -= mk(f.x, f.y, f.z);
                        T += mk(m.x, m.y, m.z);
                        E += 0.5f * f.w;
                    }
                }
            }
            F = mk(group_sum(F.x, G), group_sum(F.y, G), group_sum(F.z, G));
            T = mk(group_sum(T.x, G), group_sum(T.y, G), group_sum(T.z, G));
            E = group_sum(E, G);
            if (leader) {
                if (pass == 0) {
                    if (bonded) {
                        const int4 b = Sbond[p];
                        if (b.x >= 0) {
                            const float4 f = Sbrf[p], m = Sbrt[p];
                            F += mk(f.x, f.y, f.z);
                            T += mk(m.x, m.y, m.z);
                            E += f.w;
                        }
                        if (b.y >= 0) {
                            const float4 f = Sbrf[ncap + p], m = Sbrt[ncap + p];
                            F += mk(f.x, f.y, f.z);
                            T += mk(m.x, m.y, m.z);
                            E += f.w;
                        }
                    }
                    if (has_bias && (p == (uint32_t)bias.e || p == (uint32_t)bias.s)) {
                        const float4 fb = Sbias[0];
                        if (p == (uint32_t)bias.e) F += mk(fb.x, fb.y, fb.z);
                        if (p == (uint32_t)bias.s) F -= mk(fb.x, fb.y, fb.z);
                        E += 0.5f * fb.w;
                    }
                } else {
                    const float4 f0 = Sacc[p], t0 = Sacc[ncap + p];
                    F = mk(f0.x, f0.y, f0.z) + F;
                    T = mk(t0.x, t0.y, t0.z) + T;
                    E = f0.w + E;
                }
                if (pass + 1 < npass) {
                    Sacc[p] = make_float4(F.x, F.y, F.z, E);
                    Sacc[ncap + p] = make_float4(T.x, T.y, T.z, 0.f);
                } else {
                    a.force[g] = make_float4(F.x, F.y, F.z, want_e ? E : 0.f);
                    a.torque[g] = make_float4(T.x, T.y, T.z, 0.f);
                    // ---- integration: the second half of this step merged with the first
                    //      half of the next, or the closing second half
                    if (a.nsteps > 0) {
                        const unsigned long long sk = a.step0 + step + 1;
                        if (!last) {
                            double4 x = make_double4(Sxy[p].x, Sxy[p].y, Sz[p], 0.0);
                            lv_merge_pre(a.lv, p, sk, key, qv, v, pm);
                            lv_merge_kick(a.lv, F.x, F.y, F.z, T.x, T.y, T.z, qv, v, pm);
                            lv_drift(a.lv, L, Linv, v, x);
                            lv_rotate(a.lv, qv, pm);
                            Sxy[p] = make_double2(x.x, x.y);
                            Sz[p] = x.z;
                            float4 f0, f1, f2;
                            quat_axes(qv.x, qv.y, qv.z, qv.w, f0, f1, f2);
                            f0.w = charge; f1.w = tbits; f2.w = 0.f;
                            Sf0[p] = f0; Sf1[p] = f1; Sf2[p] = f2;
                            a.vel[g] = v; a.quat[g] = qv; a.angmom[g] = pm;
                        } else {
                            double4 bd, bdt;
                            lv_second_half(a.lv, p, sk, key, F.x, F.y, F.z, T.x, T.y, T.z, qv, v, pm, bd, bdt);
                            a.vel[g] = v; a.angmom[g] = pm; a.bd[g] = bd; a.bd_tq[g] = bdt;
                        }
                    }
                }
            }
            // a next pass's fill rewrites only list words (.w of the torque records, which
            // this gather does not use), incoming entries (read only with one pass) and
            // offsets of its own pass's particles (tagged, so a gather of this pass never
            // takes them); its evaluation starts after a barrier
        }
        if (!last) __syncthreads();
    }

    // ---- write the positions back
    if (leader && a.nsteps > 0) a.pos[g] = make_double4(Sxy[p].x, Sxy[p].y, Sz[p], a.pos[g].w);
    if (COUNT) {
#pragma unroll
        for (int s = 0; s < 14; ++s)
            if (cnt[s]) atomicAdd(&a.counts[s], (unsigned long long)cnt[s]);
    }
    if (a.want_e) {
        // per-replica totals in FP64, in a fixed order: wave butterfly, then the
        // wavefronts in order
#pragma unroll
        for (int s = 0; s < ENS_SLOTS; ++s) {
            double v = s == S_FENE ? et_fene : (double)et[s];
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
            if (lane == 0) Sred[wv * ENS_SLOTS + s] = v;
        }
        __syncthreads();
        if (t < ENS_SLOTS) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) v += Sred[w * ENS_SLOTS + t];
            a.rep_energy[(size_t)r * ENS_SLOTS + t] = v;
        }
    }
}

}  // namespace oxd
