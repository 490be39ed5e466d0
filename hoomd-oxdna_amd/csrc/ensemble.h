// ensemble.h — batched replicas: many independent small systems (hairpins, duplexes,
// umbrella windows) advanced together, one workgroup per replica, by one persistent
// launch that runs many Langevin steps.
//
// A replica's whole state lives in LDS for the duration of the launch: positions
// (FP64), body frames, force / torque accumulators and the integrator state, so a step
// touches HBM not at all; the launch reads the replicas at its start and writes them
// back at its end. Replicas of up to NMAX particles run on an NT-thread workgroup
// (NT >= NMAX): thread t integrates particle t, all NT threads share the pairs. Pairs are enumerated once each (i < j, all pairs of the
// replica, bonded pairs excluded): the i half is evaluated by the same term code as
// the one-system kernel, and the j half follows from Newton's third law and the
// vanishing total torque of a rotation-invariant pair energy,
//     F_j = -F_i,    tau_j = -tau_i + (r_j - r_i) x F_i,
// accumulated with LDS float atomics. Bonds and the per-replica harmonic bias are
// extra work items of the same loop. Energies are reduced per replica in FP64.
#pragma once

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
enum { ENS_LIST = 4096 };  // pairs tested per chunk (LDS list entries)

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
};

// LDS of one replica of up to NMAX particles, on an NT-thread workgroup.
template<int NMAX, int NT>
struct EnsLds {
    double x[NMAX], y[NMAX], z[NMAX];
    float4 f0[NMAX], f1[NMAX], f2[NMAX];  // a1 + charge, a2 + type bits, a3
    float F[6][NMAX];                      // force and torque accumulators
    float E[NMAX];                         // half energies
    int bp[2][NMAX], bi[2][NMAX];          // local partner, role / type
    double4 v[NMAX], q[NMAX], p[NMAX], bd[NMAX], bdt[NMAX];
    double red[NT / 64][ENS_SLOTS];
    uint32_t list[ENS_LIST];               // compacted pairs: near from the front, far from the back
    uint32_t n_near, n_far;
};

// MULTI: some table differs between type pairs (each pair finds its sets through their
// offsets in the parameter block; otherwise the base sets, at fixed offsets).
template<int NMAX, int NT, bool COUNT, bool MULTI = false>
__global__ __launch_bounds__(NT) void k_ensemble(EnsArgs a) {
    static_assert(NT >= NMAX && NT % 64 == 0, "one thread per particle at least");
    // A grid of exactly one workgroup per replica, checked against the grid size: reading
    // the grid size also gives the kernel the hidden-argument block that every other
    // kernel of the library has (without it the argument segment ends at EnsArgs, and a
    // tool that reads a dispatch's hidden arguments reads past its end).
    if (gridDim.x != a.n_rep) return;
    __shared__ EnsLds<NMAX, NT> S;
    const uint32_t t = threadIdx.x;
    const uint32_t r = a.order[blockIdx.x];
    const uint32_t off = a.offsets[r];
    const uint32_t n = a.offsets[r + 1] - off;  // <= NMAX, checked by the host
    const double L[3] = {a.boxes[3 * r], a.boxes[3 * r + 1], a.boxes[3 * r + 2]};
    const double Linv[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
    const DParams& P = *a.P;
    const uint32_t mask = a.mask;
    const int nt = P.ntypes;
    const unsigned long long key = a.seed + r;
    EnsBias bias{0.0, 0.0, 0, 0};
    if (a.bias) bias = a.bias[r];
    const uint32_t has_bias = (bias.k != 0.0 && n > 1) ? 1u : 0u;

    // ---- load the replica
    if (t < n) {
        const uint32_t g = off + t;
        const double4 x = a.pos[g];
        S.x[t] = x.x; S.y[t] = x.y; S.z[t] = x.z;
        const double4 q = a.quat[g];
        S.q[t] = q;
        S.v[t] = a.vel[g];
        S.p[t] = a.angmom[g];
        S.bd[t] = a.bd[g];
        S.bdt[t] = a.bd_tq[g];
        float4 f0, f1, f2;
        quat_axes(q.x, q.y, q.z, q.w, f0, f1, f2);
        f0.w = a.charge ? (float)a.charge[g] : 1.f;
        f1.w = __int_as_float(type_of(x.w));
        f2.w = 0.f;
        S.f0[t] = f0; S.f1[t] = f1; S.f2[t] = f2;
        const float4 fo = a.force[g], to = a.torque[g];
        S.F[0][t] = fo.x; S.F[1][t] = fo.y; S.F[2][t] = fo.z;
        S.F[3][t] = to.x; S.F[4][t] = to.y; S.F[5][t] = to.z;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int o = a.bpart[2 * g + s];
            S.bp[s][t] = o >= 0 ? o - (int)off : -1;
            S.bi[s][t] = a.binfo[2 * g + s];
        }
    }

    const uint32_t npairs = n * (n - 1) / 2;
    const uint32_t total = a.nsteps > 0 ? a.nsteps : max(a.nevals, 1u);
    float et[ENS_SLOTS];
    uint32_t cnt[14];
#pragma unroll
    for (int s = 0; s < 14; ++s) cnt[s] = 0u;

    for (uint32_t step = 0; step < total; ++step) {
        const bool last = step + 1 == total;
        const bool want_e = last && a.want_e;
        // ---- first half of the step (integration) or plain re-evaluation; zero the
        //      accumulators of this thread's particle
        if (t < n) {
            if (a.nsteps > 0) {
                double4 x = make_double4(S.x[t], S.y[t], S.z[t], 0.0);
                double4 v = S.v[t], q = S.q[t], p = S.p[t];
                lv_first_half(a.lv, L, Linv, S.F[0][t], S.F[1][t], S.F[2][t], S.F[3][t], S.F[4][t], S.F[5][t],
                              S.bd[t], S.bdt[t], x, v, q, p);
                S.x[t] = x.x; S.y[t] = x.y; S.z[t] = x.z;
                S.v[t] = v; S.q[t] = q; S.p[t] = p;
                float4 f0, f1, f2;
                quat_axes(q.x, q.y, q.z, q.w, f0, f1, f2);
                f0.w = S.f0[t].w; f1.w = S.f1[t].w; f2.w = 0.f;
                S.f0[t] = f0; S.f1[t] = f1; S.f2[t] = f2;
            }
#pragma unroll
            for (int s = 0; s < 6; ++s) S.F[s][t] = 0.f;
            S.E[t] = 0.f;
        }
#pragma unroll
        for (int s = 0; s < ENS_SLOTS; ++s) et[s] = 0.f;
        __syncthreads();

        // ---- forces: unordered pairs (i, i + d mod n), d = 1 + k / n, i = k mod n, tested
        //      ENS_LIST at a time: a distance test sorts them into a near list (every
        //      term) and a far list (Debye-Hueckel only) in LDS, then whole wavefronts
        //      evaluate list entries, so lanes rarely idle on pairs beyond the cut-offs.
        {
            constexpr uint32_t CH = ENS_LIST / NT;  // pairs tested per thread per chunk
            const uint32_t lane = __lane_id();
            const unsigned long long below = (1ull << lane) - 1ull;
            uint32_t pi = t % n, pd = 1 + t / n;
            const uint32_t stq = NT / n, str = NT % n;
            for (uint32_t base = 0; base < npairs; base += ENS_LIST) {
                if (t == 0) { S.n_near = 0; S.n_far = 0; }
                __syncthreads();
#pragma unroll
                for (uint32_t q = 0; q < CH; ++q) {
                    const uint32_t k = base + q * NT + t;
                    const uint32_t i = pi;
                    uint32_t j = pi + pd;
                    if (j >= n) j -= n;
                    pi += str;
                    pd += stq;
                    if (pi >= n) { pi -= n; pd += 1; }
                    bool near = false, far = false;
                    if (k < npairs && (int)j != S.bp[0][i] && (int)j != S.bp[1][i]) {
                        // the same arithmetic as the evaluation below, so both see one rsq
                        const V dr = mk((float)min_image(S.x[i] - S.x[j], L[0], Linv[0]),
                                        (float)min_image(S.y[i] - S.y[j], L[1], Linv[1]),
                                        (float)min_image(S.z[i] - S.z[j], L[2], Linv[2]));
                        const float rsq = dot(dr, dr);
                        near = rsq < a.rnear2;
                        far = !near && rsq < a.rfar2;
                    }
                    const unsigned long long mn = __ballot(near), mf = __ballot(far);
                    uint32_t bn = 0, bf = 0;
                    if (lane == 0) {
                        if (mn) bn = atomicAdd(&S.n_near, (uint32_t)__popcll(mn));
                        if (mf) bf = atomicAdd(&S.n_far, (uint32_t)__popcll(mf));
                    }
                    bn = __shfl(bn, 0);
                    bf = __shfl(bf, 0);
                    const uint32_t packed = i | (j << 16);
                    if (near) S.list[bn + __popcll(mn & below)] = packed;
                    if (far) S.list[ENS_LIST - 1 - (bf + __popcll(mf & below))] = packed;
                }
                __syncthreads();
                const uint32_t nn = S.n_near, nf = S.n_far;
                for (uint32_t e = t; e < nn + nf; e += NT) {
                    const uint32_t packed = e < nn ? S.list[e] : S.list[ENS_LIST - 1 - (e - nn)];
                    const uint32_t i = packed & 0xffffu, j = packed >> 16;
                    const double dxd = min_image(S.x[i] - S.x[j], L[0], Linv[0]);
                    const double dyd = min_image(S.y[i] - S.y[j], L[1], Linv[1]);
                    const double dzd = min_image(S.z[i] - S.z[j], L[2], Linv[2]);
                    const V dr = mk((float)dxd, (float)dyd, (float)dzd);
                    const float rsq = dot(dr, dr);
                    const float4 g0 = S.f0[i], g1 = S.f1[i], g2 = S.f2[i];
                    const float4 h0 = S.f0[j], h1 = S.f1[j], h2 = S.f2[j];
                    const int tp = __float_as_int(g1.w) * nt + __float_as_int(h1.w);
                    const Frame fi{mk(g0.x, g0.y, g0.z), mk(g1.x, g1.y, g1.z), mk(g2.x, g2.y, g2.z)};
                    const Frame fj{mk(h0.x, h0.y, h0.z), mk(h1.x, h1.y, h1.z), mk(h2.x, h2.y, h2.z)};
                    const V Rc = -dr;
                    V Fp = zero3(), Tp = zero3();
                    const uint32_t m = e < nn ? mask : (mask & (1u << T_DEBYE));
                    const float ep = pair_terms<COUNT, MULTI>(P, m, tp, rsq, Rc, fi, fj, g0.w * h0.w, Fp, Tp, et, cnt);
                    if (Fp.x != 0.f || Fp.y != 0.f || Fp.z != 0.f || Tp.x != 0.f || Tp.y != 0.f || Tp.z != 0.f) {
                        const V Tj = cross(Rc, Fp) - Tp;
                        atomicAdd(&S.F[0][i], Fp.x); atomicAdd(&S.F[1][i], Fp.y); atomicAdd(&S.F[2][i], Fp.z);
                        atomicAdd(&S.F[3][i], Tp.x); atomicAdd(&S.F[4][i], Tp.y); atomicAdd(&S.F[5][i], Tp.z);
                        atomicAdd(&S.F[0][j], -Fp.x); atomicAdd(&S.F[1][j], -Fp.y); atomicAdd(&S.F[2][j], -Fp.z);
                        atomicAdd(&S.F[3][j], Tj.x); atomicAdd(&S.F[4][j], Tj.y); atomicAdd(&S.F[5][j], Tj.z);
                    }
                    if (want_e && ep != 0.f) {
                        atomicAdd(&S.E[i], 0.5f * ep);
                        atomicAdd(&S.E[j], 0.5f * ep);
                    }
                }
                __syncthreads();  // the list is rewritten by the next chunk
            }
        }
        // ---- bonds, each from its A end
        if (mask & (1u << T_BONDED)) {
            for (uint32_t A = t; A < n; A += NT) {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const int Bi = S.bp[s][A];
                    const int info = S.bi[s][A];
                    if (Bi < 0 || (info & 1) != 0) continue;  // A's bonds where it is the A end
                    const uint32_t B = (uint32_t)Bi;
                    const float4 x0 = S.f0[A], x1 = S.f1[A], x2 = S.f2[A];
                    const float4 y0 = S.f0[B], y1 = S.f1[B], y2 = S.f2[B];
                    const Frame fA{mk(x0.x, x0.y, x0.z), mk(x1.x, x1.y, x1.z), mk(x2.x, x2.y, x2.z)};
                    const Frame fB{mk(y0.x, y0.y, y0.z), mk(y1.x, y1.y, y1.z), mk(y2.x, y2.y, y2.z)};
                    BondRes br;
                    bond_eval(P.bond[info >> 1], min_image(S.x[B] - S.x[A], L[0], Linv[0]),
                              min_image(S.y[B] - S.y[A], L[1], Linv[1]), min_image(S.z[B] - S.z[A], L[2], Linv[2]),
                              fA, fB, __float_as_int(x1.w), __float_as_int(y1.w), br);
                    if (COUNT) { cnt[2 * T_BONDED]++; cnt[2 * T_BONDED + 1] += br.e_stack != 0.f; }
                    atomicAdd(&S.F[0][A], -br.FB.x); atomicAdd(&S.F[1][A], -br.FB.y); atomicAdd(&S.F[2][A], -br.FB.z);
                    atomicAdd(&S.F[3][A], br.TA.x); atomicAdd(&S.F[4][A], br.TA.y); atomicAdd(&S.F[5][A], br.TA.z);
                    atomicAdd(&S.F[0][B], br.FB.x); atomicAdd(&S.F[1][B], br.FB.y); atomicAdd(&S.F[2][B], br.FB.z);
                    atomicAdd(&S.F[3][B], br.TB.x); atomicAdd(&S.F[4][B], br.TB.y); atomicAdd(&S.F[5][B], br.TB.z);
                    const float ef = (float)br.e_fene;
                    et[S_FENE] += ef;
                    et[S_BEXC] += br.e_bexc;
                    et[S_STACK] += br.e_stack;
                    if (want_e) {
                        const float eh = 0.5f * (ef + br.e_bexc + br.e_stack);
                        atomicAdd(&S.E[A], eh);
                        atomicAdd(&S.E[B], eh);
                    }
                }
            }
        }
        // ---- harmonic bias between the centres of particles s and e (last thread)
        if (has_bias && t == NT - 1) {
            const uint32_t bs = (uint32_t)bias.s, be = (uint32_t)bias.e;
            const double dx = min_image(S.x[be] - S.x[bs], L[0], Linv[0]);
            const double dy = min_image(S.y[be] - S.y[bs], L[1], Linv[1]);
            const double dz = min_image(S.z[be] - S.z[bs], L[2], Linv[2]);
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            const double u = d - bias.d0;
            const double e = bias.k * u * u;
            const double c = d > 0.0 ? -2.0 * bias.k * u / d : 0.0;  // force on e = c * (r_e - r_s)
            const V f = mk((float)(c * dx), (float)(c * dy), (float)(c * dz));
            atomicAdd(&S.F[0][be], f.x); atomicAdd(&S.F[1][be], f.y); atomicAdd(&S.F[2][be], f.z);
            atomicAdd(&S.F[0][bs], -f.x); atomicAdd(&S.F[1][bs], -f.y); atomicAdd(&S.F[2][bs], -f.z);
            et[S_BIAS] += (float)e;
            if (want_e) {
                atomicAdd(&S.E[be], (float)(0.5 * e));
                atomicAdd(&S.E[bs], (float)(0.5 * e));
            }
        }
        __syncthreads();

        // ---- second half of the step
        if (a.nsteps > 0 && t < n) {
            double4 v = S.v[t], p = S.p[t], bd, bdt;
            lv_second_half(a.lv, t, a.step0 + step + 1, key, S.F[0][t], S.F[1][t], S.F[2][t], S.F[3][t],
                           S.F[4][t], S.F[5][t], S.q[t], v, p, bd, bdt);
            S.v[t] = v; S.p[t] = p; S.bd[t] = bd; S.bdt[t] = bdt;
        }
        // no barrier here: until the next one every thread touches only its own
        // particle's entries
    }

    // ---- write the replica back
    if (t < n) {
        const uint32_t g = off + t;
        if (a.nsteps > 0) {
            const double w = a.pos[g].w;
            a.pos[g] = make_double4(S.x[t], S.y[t], S.z[t], w);
            a.quat[g] = S.q[t];
            a.vel[g] = S.v[t];
            a.angmom[g] = S.p[t];
            a.bd[g] = S.bd[t];
            a.bd_tq[g] = S.bdt[t];
        }
        a.force[g] = make_float4(S.F[0][t], S.F[1][t], S.F[2][t], S.E[t]);
        a.torque[g] = make_float4(S.F[3][t], S.F[4][t], S.F[5][t], 0.f);
    }
    if (COUNT) {
#pragma unroll
        for (int s = 0; s < 14; ++s)
            if (cnt[s]) atomicAdd(&a.counts[s], (unsigned long long)cnt[s]);
    }
    if (a.want_e) {
        // per-replica totals in FP64: wave shuffle, then across the waves in LDS
        const int wid = t / 64, wl = t % 64;
#pragma unroll
        for (int s = 0; s < ENS_SLOTS; ++s) {
            double v = (double)et[s];
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
            if (wl == 0) S.red[wid][s] = v;
        }
        __syncthreads();
        if (t < ENS_SLOTS) {
            double v = 0.0;
#pragma unroll
            for (int w = 0; w < NT / 64; ++w) v += S.red[w][t];
            a.rep_energy[(size_t)r * ENS_SLOTS + t] = v;
        }
    }
}

}  // namespace oxd
