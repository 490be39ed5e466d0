// oxdna_device.h — device-side parameters and per-pair physics of the oxDNA terms
// for gfx950 (FP32 on the VALU; FP64 only for minimum-image displacements, the FENE
// log / 1/(delta^2 - dr^2) terms and reductions).
//
// Every particle's body axes (a1, a2, a3) are computed once per step from its FP64
// quaternion (kernel `frames`), so a pair evaluation never rotates: an interaction
// site s (body frame) is s.x a1 + s.y a2 + s.z a3. A pair evaluation returns only
// what particle i needs from a full neighbour list: force on i, lab-frame torque on
// i and the pair energy (j's share is computed by j's own evaluation, as in the
// host engine's GPU pair path the reference plugs into).
//
// The formulas follow the reference evaluators; each function cites the lines.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oxd {

// ------------------------------------------------------------------ float3 helpers
struct V { float x, y, z; };
__device__ __forceinline__ V mk(float x, float y, float z) { V r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V operator-(V a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V operator*(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V operator*(float s, V a) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V& operator+=(V& a, V b) { a = a + b; return a; }
__device__ __forceinline__ V& operator-=(V& a, V b) { a = a - b; return a; }
__device__ __forceinline__ float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V cross(V a, V b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V zero3() { return mk(0.f, 0.f, 0.f); }

struct Frame { V a1, a2, a3; };
__device__ __forceinline__ V site(const float s[3], const Frame& f) {
    return f.a1 * s[0] + f.a2 * s[1] + f.a3 * s[2];
}

// ------------------------------------------------------------------ math
// arccos on [-1, 1] by Abramowitz & Stegun 4.4.46: acos(x) = sqrt(1 - x) P7(x) for
// 0 <= x <= 1 with |error| <= 2e-8 rad, and acos(-x) = pi - acos(x). Also returns
// sin(acos(x)) = sqrt((1 - x)(1 + x)), which the f4 torque coefficient divides by.
__host__ __device__ __forceinline__ float acos_as(float x, float& sin_out) {
    float ax = fabsf(x);
    float p = -0.0012624911f;
    p = fmaf(p, ax, 0.0066700901f);
    p = fmaf(p, ax, -0.0170881256f);
    p = fmaf(p, ax, 0.0308918810f);
    p = fmaf(p, ax, -0.0501743046f);
    p = fmaf(p, ax, 0.0889789874f);
    p = fmaf(p, ax, -0.2145988016f);
    p = fmaf(p, ax, 1.5707963050f);
    float r = sqrtf(1.0f - ax) * p;
    sin_out = sqrtf((1.0f - ax) * (1.0f + ax));
    return x < 0.0f ? 3.14159265358979f - r : r;
}
__device__ __forceinline__ float clampc(float c) { return fminf(1.0f, fmaxf(-1.0f, c)); }

// ------------------------------------------------------------------ parameter blocks (FP32)
struct F4P { float a, b, t0, ts, tc, clo, chi; };  // angular modulation set; f4 != 0 iff clo < cos < chi
struct F5P { float a, b, xc, xs; };        // dihedral set
struct F1P { float a, r0, rlow, rhigh, rclow, rchigh, blow, bhigh; };
struct F2P { float k, r0, rc, rlow, rhigh, rclow, rchigh, blow, bhigh; };

struct DExcVol { float eps; float base[3], back[3]; float sigma[4], rstar[4], b[4], rc[4]; };
// Base-base block: hydrogen bonding (f1, per-type-pair eps) and cross-stacking (f2).
struct DBaseBase {
    float base[3];
    F1P f1; float shift_factor; float hb_eps[16];
    F4P hb_t1, hb_t23, hb_t4, hb_t78;
    F2P f2; F4P cs_t1, cs_t23, cs_t4, cs_t78; int cs_t4_enabled;
};
struct DCoax {
    float stack[3], backref[3], gamma;
    F2P f2; F4P t1, t4, t56; F5P phi3, phi4;
    int t1_mode, phi3_enabled, rna_coax; float t1_sa, t1_sb;
    float t1_cx;  // cos theta1 below which theta1's second piece (mode 0: f4(2 pi - theta1),
                  // mode 1: SA (theta1 - SB)^2) is nonzero; -2 if never
};
struct DDebye { float back[3]; float minus_kappa, prefactor, b_smooth, rc, rhigh; };
struct BondExc { float eps, sigma[3], rstar[3], b[3], rc[3]; };
struct DBonded {
    double eps_backbone, r0_backbone, delta_backbone, delta2;  // FENE stays FP64
    float back[3], base[3];
    BondExc bx;  // bonded excluded volume: base-base, base_A-back_B, back_A-base_B
    int stacking_enabled, rna_stacking;
    float stack[3], backref[3], gamma;
    F1P f1; float shift_factor; float st_eps[16];
    F4P t4, t5; F5P p1;
    float stack3[3], stack5[3], bbv3[3], bbv5[3];
    F4P tB1, tB2;
};

enum { MAX_BOND_TYPES = 4 };

// All term parameters of one force field: the parameter block is this struct followed
// by an extras block (in global memory and, in the force kernel, at the start of its
// dynamic LDS). Each nonbonded term has a base parameter set (the fields below) and,
// when its table differs between type pairs, further distinct sets in the extras:
// poff[tp][t] is the byte offset, from the start of the block, of type pair tp's set of
// term t (the base set's own offset when the tables are uniform). The hydrogen-bonding
// epsilon of every type pair stays in the base set's hb_eps table.
struct alignas(16) DParams {  // a multiple of 16 bytes: kernels copy it to LDS in 16-byte words
    uint32_t mask;            // bits of oxdna_term present
    int ntypes;
    float rcutsq[6][16];      // per nonbonded term, per type pair
    DExcVol ev;
    DBaseBase hb;             // used by OXDNA_TERM_HBOND (hb_*) ...
    DBaseBase cs;             // ... OXDNA_TERM_CROSS (cs_*) ...
    DBaseBase hbx;            // ... and the fused OXDNA_TERM_HBONDCROSS (both)
    DCoax cx;
    DDebye dh;
    int n_bond_types;
    DBonded bond[MAX_BOND_TYPES];
    // last, so that the sets above keep their offsets (and the alignment their reads
    // are merged on)
    uint16_t poff[16][8];     // per type pair: byte offset of each nonbonded term's set (0-5)
    uint32_t xwords;          // size of the extras block in 16-byte words (0: every table uniform)
};

// ------------------------------------------------------------------ building blocks
// The force kernel keeps the parameter sets in LDS. A field first read inside a branch
// is read there and waited for on the spot, one round trip per branch level; eager()
// reads a whole set (float fields only) up front in one batch with one wait: the empty
// asm statements take every value as an input, so no read can sink into the branches.
template<class T>
__device__ __forceinline__ T eager(const T& p) {
    static_assert(sizeof(T) % sizeof(float) == 0, "parameter sets hold float fields only");
    constexpr int n = (int)(sizeof(T) / sizeof(float));
    T r;
    const float* src = reinterpret_cast<const float*>(&p);
    float* dst = reinterpret_cast<float*>(&r);
#pragma unroll
    for (int k = 0; k < n; ++k) dst[k] = src[k];
#pragma unroll
    for (int k = 0; k < n; ++k) asm volatile("" ::"v"(dst[k]));
    return r;
}
// f1 Morse radial value and derivative (reference src/OxDNANumerics.h:114-155).
__device__ __forceinline__ void f1_vd(float r, const F1P& s_, float eps, float shift, float& v, float& d) {
    const F1P s = eager(s_);
    v = 0.f; d = 0.f;
    if (r < s.rchigh) {
        if (r > s.rhigh) {
            float t = r - s.rchigh;
            v = eps * s.bhigh * t * t;
            d = eps * 2.f * s.bhigh * t;
        } else if (r > s.rlow) {
            float e = __expf(-(r - s.r0) * s.a);
            float t = 1.f - e;
            v = eps * t * t - shift;
            d = eps * 2.f * t * e * s.a;
        } else if (r > s.rclow) {
            float t = r - s.rclow;
            v = eps * s.blow * t * t;
            d = eps * 2.f * s.blow * t;
        }
    }
}
// f2 harmonic radial value and derivative (reference src/OxDNANumerics.h:158-192).
__device__ __forceinline__ void f2_vd(float r, const F2P& s_, float& v, float& d) {
    const F2P s = eager(s_);
    v = 0.f; d = 0.f;
    if (r < s.rchigh) {
        if (r > s.rhigh) {
            float t = r - s.rchigh;
            v = s.k * s.bhigh * t * t;
            d = 2.f * s.k * s.bhigh * t;
        } else if (r > s.rlow) {
            float t = r - s.r0, c = s.rc - s.r0;
            v = 0.5f * s.k * (t * t - c * c);
            d = s.k * t;
        } else if (r > s.rclow) {
            float t = r - s.rclow;
            v = s.k * s.blow * t * t;
            d = 2.f * s.k * s.blow * t;
        }
    }
}
// f4 of the angle th and its torque coefficient (d f4/d th)/sin(th), given sin(th)
// (reference src/OxDNANumerics.h:201-244, inner-branch 1e-8 guard included).
// Branch-free: both pieces are computed and selected (a wavefront's lanes would run
// both branches anyway).
// inv = 1 / sth, passed in where several factors share one angle's sine.
__device__ __forceinline__ void f4_vdi(float th, float sth, float inv, const F4P& p, float& v, float& dsin) {
    const float t0 = th - p.t0;
    const float t = fabsf(t0);
    const float m = t0 < 0.f ? -1.f : 1.f;
    const float e = p.tc - t;
    const bool tail = t > p.ts;
    const float v_tail = p.b * e * e, d_tail = -m * 2.f * p.b * e * inv;
    const float v_in = 1.f - p.a * t * t;
    const float d_in = (sth * sth > 1e-8f) ? -m * 2.f * p.a * t * inv : -m * 2.f * p.a;
    const bool in = t < p.tc;
    v = in ? (tail ? v_tail : v_in) : 0.f;
    dsin = in ? (tail ? d_tail : d_in) : 0.f;
}
__device__ __forceinline__ void f4_vd(float th, float sth, const F4P& p, float& v, float& dsin) {
    f4_vdi(th, sth, 1.f / sth, p, v, dsin);
}
__device__ __forceinline__ float f4_v(float th, const F4P& p_) {
    const F4P p = eager(p_);
    float t = fabsf(th - p.t0);
    if (!(t < p.tc)) return 0.f;
    if (t > p.ts) { float e = p.tc - t; return p.b * e * e; }
    return 1.f - p.a * t * t;
}
// f5 of cos(phi) and its derivative (reference src/OxDNANumerics.h:261-290).
__device__ __forceinline__ void f5_vd(float c, const F5P& p_, float& v, float& d) {
    const F5P p = eager(p_);
    v = 0.f; d = 0.f;
    if (c > p.xc) {
        if (c < p.xs) { float e = p.xc - c; v = p.b * e * e; d = 2.f * p.b * (c - p.xc); }
        else if (c < 0.f) { v = 1.f - p.a * c * c; d = -2.f * p.a * c; }
        else { v = 1.f; d = 0.f; }
    }
}

// Whether the angular factor of cosine c can be nonzero (the reflected angle too when
// sym): a comparison in cosine space that spares the arccos of screened-out pairs.
// Slack of a few ulps keeps every pair whose f4 the full evaluation finds nonzero.
__device__ __forceinline__ bool f4_nz(float c, const F4P& p, bool sym) {
    const float slack = 4e-6f;
    c = clampc(c);
    bool nz = c > p.clo - slack && c < p.chi + slack;
    if (sym) nz = nz || (-c > p.clo - slack && -c < p.chi + slack);
    return nz;
}


// Smoothed LJ core (reference src/OxDNANumerics.h:76-107); f = force on the q site.
__device__ __forceinline__ float lj(V r, float eps, float sigma, float rstar, float b, float rc, V& f) {
    float rr = dot(r, r);
    f = zero3();
    if (!(rr < rc * rc)) return 0.f;
    if (rr > rstar * rstar) {
        float rm = sqrtf(rr);
        float d = rm - rc;
        f = r * (-2.f * eps * b * d / rm);
        return eps * b * d * d;
    }
    float s2 = sigma * sigma / rr;
    float s6 = s2 * s2 * s2;
    f = r * (-24.f * eps * (s6 - 2.f * s6 * s6) / rr);
    return 4.f * eps * (s6 * s6 - s6);
}

// Accumulator of one particle's share of a pair.
struct Acc { V F, T; float E; };

// ------------------------------------------------------------------ nonbonded terms
// Excluded volume (reference src/EvaluatorPairOxDNAExcVol.h:107-149).
__device__ __forceinline__ float ev_pair(V Rc, const Frame& fi, const Frame& fj, const DExcVol& p_, V& F, V& T) {
    const DExcVol p = eager(p_);
    V bk_i = site(p.back, fi), bs_i = site(p.base, fi);
    V bk_j = site(p.back, fj), bs_j = site(p.base, fj);
    float e = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        V si = (k == 0 || k == 3) ? bk_i : bs_i;
        V sj = (k == 0 || k == 2) ? bk_j : bs_j;
        V f;
        e += lj(Rc + sj - si, p.eps, p.sigma[k], p.rstar[k], p.b[k], p.rc[k], f);
        F -= f;
        T -= cross(si, f);
    }
    return e;
}

// Shared BASE-BASE geometry (reference src/OxDNABaseBase.h:84-121).
struct BB { V rdir, base_i; float rmod, irmod, c1, c2, c3, c4, c7, c8; };  // irmod = 1 / rmod
__device__ __forceinline__ bool bb_geom(V Rc, const Frame& fi, const Frame& fj, const float base[3], BB& g) {
    g.base_i = site(base, fi);
    V r = Rc + site(base, fj) - g.base_i;
    g.rmod = sqrtf(dot(r, r));
    if (!(g.rmod > 0.f)) return false;
    g.irmod = 1.f / g.rmod;
    g.rdir = r * g.irmod;
    g.c1 = -dot(fi.a1, fj.a1);
    g.c2 = -dot(fj.a1, g.rdir);
    g.c3 = dot(fi.a1, g.rdir);
    g.c4 = dot(fi.a3, fj.a3);
    g.c7 = -dot(fj.a3, g.rdir);
    g.c8 = dot(fi.a3, g.rdir);
    return true;
}

// The angles of the six BASE-BASE cosines (c1, c2, c3, c4, c7, c8) and their sines: one
// arccos each, shared by the two terms of the fused hydrogen-bonding + cross-stacking
// evaluator, which read the same geometry.
struct BBAng { float th[6], s[6], inv[6]; };  // angles, sines, 1 / sines
__device__ __forceinline__ void bb_angles(const BB& g, BBAng& a) {
    const float c[6] = {g.c1, g.c2, g.c3, g.c4, g.c7, g.c8};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        a.th[k] = acos_as(clampc(c[k]), a.s[k]);
        a.inv[k] = 1.f / a.s[k];
    }
}
// Angular factor with optional reflection (reference src/OxDNABaseBase.h:52-67):
// value f4(th) [+ f4(pi - th)], coefficient sign * (D(th) [- D(pi - th)]).
struct Ang { float v, d; };
// Angular factor from a precomputed angle, its sine and the sine's inverse.
__device__ __forceinline__ Ang angular_th(float th, float s, float inv, const F4P& p, bool sym, float sign) {
    Ang a;
    f4_vdi(th, s, inv, p, a.v, a.d);
    if (sym) {
        float v2, d2;
        f4_vdi(3.14159265358979f - th, s, inv, p, v2, d2);
        a.v += v2;
        a.d -= d2;
    }
    a.d *= sign;
    return a;
}
__device__ __forceinline__ Ang angular(float c, const F4P& p, bool sym, float sign) {
    float s;
    float th = acos_as(clampc(c), s);
    return angular_th(th, s, 1.f / s, p, sym, sign);
}

// Whether a BASE-BASE term can be nonzero: nonzero radial (the reference's early exit)
// and all six angular factors possibly nonzero, screened in cosine space.
__device__ __forceinline__ bool bb_active(const BB& g, float rv, const F4P& t1, const F4P& t23, const F4P& t4,
                                          const F4P& t78, bool sym, bool use_t4) {
    return rv != 0.f && f4_nz(g.c1, t1, false) && f4_nz(g.c2, t23, false) && f4_nz(g.c3, t23, false) &&
           (!use_t4 || f4_nz(g.c4, t4, sym)) && f4_nz(g.c7, t78, sym) && f4_nz(g.c8, t78, sym);
}

// A BASE-BASE term's energy and the scalar weights of its force and torque: the
// assembly (bb_assemble) is linear in them over vectors that depend only on the shared
// geometry, so the two terms of the fused evaluator add their weights and assemble once.
struct BBW { float e, wr, w1, w2, w3, w4, w7, w8; };
__device__ __forceinline__ void bbw_zero(BBW& w) { w = BBW{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ void bbw_add(BBW& w, const BBW& x) {
    w.e += x.e; w.wr += x.wr; w.w1 += x.w1; w.w2 += x.w2; w.w3 += x.w3; w.w4 += x.w4; w.w7 += x.w7; w.w8 += x.w8;
}
// Weights of one term given its radial value / derivative and the angles (reference
// src/OxDNABaseBase.h:128-219, the p = i half); zero weights when a factor vanishes.
__device__ __forceinline__ BBW bb_weights(const BB& g, const BBAng& an, float rv, float rd, const F4P& t1,
                                          const F4P& t23, const F4P& t4, const F4P& t78, bool sym, bool use_t4) {
    Ang A1 = angular_th(an.th[0], an.s[0], an.inv[0], t1, false, -1.f);
    Ang A2 = angular_th(an.th[1], an.s[1], an.inv[1], t23, false, -1.f);
    Ang A3 = angular_th(an.th[2], an.s[2], an.inv[2], t23, false, 1.f);
    Ang A4 = use_t4 ? angular_th(an.th[3], an.s[3], an.inv[3], t4, sym, 1.f) : Ang{1.f, 0.f};
    Ang A7 = angular_th(an.th[4], an.s[4], an.inv[4], t78, sym, -1.f);
    Ang A8 = angular_th(an.th[5], an.s[5], an.inv[5], t78, sym, 1.f);
    float p12 = A1.v * A2.v, p34 = A3.v * A4.v, p78 = A7.v * A8.v;
    float prod = p12 * p34 * p78;
    BBW w;
    w.e = rv * prod;
    if (w.e == 0.f) { bbw_zero(w); return w; }
    // products with one factor replaced by its coefficient
    const float inv = g.irmod;
    w.wr = -rd * prod;
    w.w1 = rv * A1.d * A2.v * p34 * p78;
    w.w2 = rv * A1.v * A2.d * p34 * p78 * inv;
    w.w3 = rv * p12 * A3.d * A4.v * p78 * inv;
    w.w4 = rv * p12 * A3.v * A4.d * p78;
    w.w7 = rv * p12 * p34 * A7.d * A8.v * inv;
    w.w8 = rv * p12 * p34 * A7.v * A8.d * inv;
    return w;
}
// Force and torque on i from summed weights (w3, w8 already divided by |r| for the
// force; the torque terms take them times |r|).
__device__ __forceinline__ void bb_assemble(const BB& g, const Frame& fi, const Frame& fj, const BBW& w, V& F, V& T) {
    V fq = g.rdir * w.wr;
    fq += (fj.a1 + g.rdir * g.c2) * w.w2;
    fq += (fi.a1 - g.rdir * g.c3) * w.w3;
    fq += (fj.a3 + g.rdir * g.c7) * w.w7;
    fq += (fi.a3 - g.rdir * g.c8) * w.w8;
    V tp = cross(fi.a3, fj.a3) * w.w4 + cross(fi.a1, fj.a1) * w.w1;
    tp -= cross(g.rdir, fi.a1) * (w.w3 * g.rmod);
    tp -= cross(g.rdir, fi.a3) * (w.w8 * g.rmod);
    tp -= cross(g.base_i, fq);
    F -= fq;
    T += tp;
}
// One BASE-BASE term given its radial value / derivative and the angles; adds force and
// torque on i.
__device__ __forceinline__ float bb_term_ang(const BB& g, const BBAng& an, const Frame& fi, const Frame& fj, float rv,
                                             float rd, const F4P& t1, const F4P& t23, const F4P& t4,
                                             const F4P& t78, bool sym, bool use_t4, V& F, V& T) {
    const BBW w = bb_weights(g, an, rv, rd, t1, t23, t4, t78, sym, use_t4);
    if (w.e == 0.f) return 0.f;
    bb_assemble(g, fi, fj, w, F, T);
    return w.e;
}

// One BASE-BASE term (screen, angles, evaluation).
__device__ __forceinline__ float bb_term(const BB& g, const Frame& fi, const Frame& fj, float rv, float rd,
                                         const F4P& t1, const F4P& t23, const F4P& t4, const F4P& t78,
                                         bool sym, bool use_t4, V& F, V& T) {
    if (!bb_active(g, rv, t1, t23, t4, t78, sym, use_t4)) return 0.f;
    BBAng an;
    bb_angles(g, an);
    return bb_term_ang(g, an, fi, fj, rv, rd, t1, t23, t4, t78, sym, use_t4, F, T);
}

// Coaxial stacking (reference src/EvaluatorPairOxDNACoaxStacking.h:134-372), p = i half.
__device__ __forceinline__ float coax_pair(V Rc, const Frame& fi, const Frame& fj, const DCoax& p, V& F, V& T) {
    const float PI = 3.14159265358979f;
    V st_i = site(p.stack, fi);
    V rs = Rc + site(p.stack, fj) - st_i;
    float rsm = sqrtf(dot(rs, rs));
    if (!(rsm > 0.f)) return 0.f;
    float fr, frD;
    f2_vd(rsm, p.f2, fr, frD);
    if (fr == 0.f) return 0.f;
    V rsd = rs * (1.f / rsm);
    float c1 = -dot(fi.a1, fj.a1), c4 = dot(fi.a3, fj.a3), c5 = dot(fi.a3, rsd), c6 = -dot(fj.a3, rsd);
    // theta4, theta5, theta6 and theta1 (its f4 or its mode's second piece) factors
    // screened in cosine space: a zero factor with a zero derivative zeroes the term
    if (!(f4_nz(c4, p.t4, false) && f4_nz(c5, p.t56, true) && f4_nz(c6, p.t56, true) &&
          (f4_nz(c1, p.t1, false) || clampc(c1) < p.t1_cx + 4e-6f)))
        return 0.f;
    float s1, s4, s5, s6;
    float th1 = acos_as(clampc(c1), s1), th4 = acos_as(clampc(c4), s4);
    float th5 = acos_as(clampc(c5), s5), th6 = acos_as(clampc(c6), s6);
    float g1, g1D, v, d;
    const float inv1 = 1.f / s1, inv5 = 1.f / s5, inv6 = 1.f / s6;  // each shared by two factors
    f4_vdi(th1, s1, inv1, p.t1, g1, g1D);
    g1D = -g1D;
    if (p.t1_mode == 0) {
        f4_vdi(2.f * PI - th1, -s1, -inv1, p.t1, v, d);  // sin(2 pi - t) = -sin(t)
        g1 += v; g1D -= d;
    } else if (th1 > p.t1_sb) {
        float t = th1 - p.t1_sb;
        g1 += p.t1_sa * t * t;
        g1D -= (s1 * s1 > 1e-8f) ? 2.f * p.t1_sa * t * inv1 : 2.f * p.t1_sa;
    }
    float g4, g4D;
    f4_vd(th4, s4, p.t4, g4, g4D);
    float g5, g5D, g6, g6D;
    f4_vdi(th5, s5, inv5, p.t56, g5, g5D);
    f4_vdi(PI - th5, s5, inv5, p.t56, v, d);
    g5 += v; g5D -= d;
    f4_vdi(th6, s6, inv6, p.t56, g6, g6D);
    g6D = -g6D;
    f4_vdi(PI - th6, s6, inv6, p.t56, v, d);
    g6 += v; g6D += d;
    V br_i = site(p.backref, fi), br_j = site(p.backref, fj);
    V rb = Rc + br_j - br_i;
    float rbm = sqrtf(dot(rb, rb));
    V rbd = rb * (1.f / rbm);
    float cp3 = dot(rsd, cross(rbd, fi.a1));
    float h3 = 0.f, h3D = 0.f, h4 = 0.f, h4D = 0.f, cp4 = 0.f, phi = 1.f;
    if (p.rna_coax) {
        cp4 = dot(rsd, cross(rbd, fj.a1));
        f5_vd(cp3, p.phi3, h3, h3D);
        f5_vd(cp4, p.phi4, h4, h4D);
        phi = h3 * h4;
    } else if (p.phi3_enabled) {
        f5_vd(cp3, p.phi3, h3, h3D);
        phi = h3 * h3;
    }
    float base = fr * g1 * g4 * g5 * g6;
    float e = base * phi;
    if (e == 0.f) return 0.f;
    V fq = rsd * (-frD * g1 * g4 * g5 * g6 * phi);
    float w1 = fr * g1D * g4 * g5 * g6 * phi;
    float w4 = fr * g1 * g4D * g5 * g6 * phi;
    float w5 = fr * g1 * g4 * g5D * g6 * phi;
    float w6 = fr * g1 * g4 * g5 * g6D * phi;
    V tp = cross(fi.a1, fj.a1) * w1 + cross(fi.a3, fj.a3) * w4;
    float inv = 1.f / rsm;
    fq += (fi.a3 - rsd * c5) * (w5 * inv);
    tp -= cross(rsd, fi.a3) * w5;
    fq += (fj.a3 + rsd * c6) * (w6 * inv);
    if (p.rna_coax) {
        tp -= cross(st_i, fq);
        float invb = 1.f / rbm;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            V vv = (w == 0) ? fi.a1 : fj.a1;
            float fc = (w == 0) ? base * h4 * h3D : base * h3 * h4D;
            V rbv = cross(rbd, vv);
            V vrs = cross(vv, rsd);
            V fst = (rbv - rsd * dot(rsd, rbv)) * (-fc * inv);
            V fbk = (vrs - rbd * dot(rbd, vrs)) * (-fc * invb);
            fq += fst + fbk;
            tp -= cross(st_i, fst) + cross(br_i, fbk);
            if (w == 0) tp -= cross(vv, cross(rsd, rbd)) * fc;
        }
        F -= fq;
        T += tp;
        return e;
    }
    if (p.phi3_enabled) {
        // oxDNA1 cos(phi3) (reference src/EvaluatorPairOxDNACoaxStacking.h:320-361).
        float gam = p.gamma;
        float rbm3 = rbm * rbm * rbm;
        float a2b1 = dot(fi.a2, fj.a1), a3b1 = dot(fi.a3, fj.a1);
        float ra1 = dot(rsd, fi.a1), ra2 = dot(rsd, fi.a2), ra3 = dot(rsd, fi.a3), rb1 = dot(rsd, fj.a1);
        float par = ra3 * a2b1 - ra2 * a3b1;
        float dcdr = -gam * par * (gam * (ra1 - rb1) + rsm) / rbm3;
        float dcda1b1 = gam * gam * gam * par / rbm3;
        float dcda2b1 = gam * ra3 / rbm;
        float dcda3b1 = -gam * ra2 / rbm;
        float dcdra1 = -gam * gam * par * rsm / rbm3;
        float dcdra2 = -gam * a3b1 / rbm;
        float dcdra3 = gam * a2b1 / rbm;
        float dcdrb1 = gam * gam * par * rsm / rbm3;
        float fc = base * 2.f * h3 * h3D;
        fq -= (rsd * dcdr + ((fi.a1 - rsd * ra1) * dcdra1 + (fi.a2 - rsd * ra2) * dcdra2 +
                             (fi.a3 - rsd * ra3) * dcdra3 + (fj.a1 - rsd * rb1) * dcdrb1) * inv) * fc;
        tp += (cross(rsd, fi.a1) * dcdra1 + cross(rsd, fi.a2) * dcdra2 + cross(rsd, fi.a3) * dcdra3) * fc;
        tp -= (cross(fi.a1, fj.a1) * dcda1b1 + cross(fi.a2, fj.a1) * dcda2b1 + cross(fi.a3, fj.a1) * dcda3b1) * fc;
    }
    tp -= cross(st_i, fq);
    F -= fq;
    T += tp;
    return e;
}

// Debye-Hueckel (reference src/EvaluatorPairOxDNADebye.h:85-141), p = i half.
__device__ __forceinline__ float dh_pair(V Rc, const Frame& fi, const Frame& fj, float qiqj, const DDebye& p_,
                                         V& F, V& T) {
    const DDebye p = eager(p_);
    V bi = site(p.back, fi);
    V rb = Rc + site(p.back, fj) - bi;
    float rm = sqrtf(dot(rb, rb));
    if (!(rm > 0.f) || !(rm < p.rc)) return 0.f;
    V rd = rb * (1.f / rm);
    float e, fmag;
    if (rm < p.rhigh) {
        float ex = __expf(rm * p.minus_kappa);
        e = ex * (p.prefactor / rm);
        fmag = -(p.prefactor * ex) * (p.minus_kappa / rm - 1.f / (rm * rm));
    } else {
        float d = rm - p.rc;
        e = p.b_smooth * d * d;
        fmag = -2.f * p.b_smooth * d;
    }
    V fq = rd * (fmag * qiqj);
    F -= fq;
    T -= cross(bi, fq);
    return e * qiqj;
}


// ------------------------------------------------------------------ bonded term
// Result of one bond (A, B) (reference src/OxDNABondedKernel.h:38-45): FB = force on
// B (force on A = -FB), lab-frame torques on A and B, energies, half virial.
struct BondRes { V FB, TA, TB; double e_fene; float e_bexc, e_stack; float vir[6]; };

__device__ __forceinline__ void vir_add(float vir[6], V r, V f) {
    vir[0] += 0.5f * r.x * f.x; vir[1] += 0.5f * r.x * f.y; vir[2] += 0.5f * r.x * f.z;
    vir[3] += 0.5f * r.y * f.y; vir[4] += 0.5f * r.y * f.z; vir[5] += 0.5f * r.z * f.z;
}

// oxDNA stacking, p = B, q = A (reference src/OxDNABondedKernel.h:69-231).
__device__ __forceinline__ float stack_dna(const DBonded& P, V Rc, const Frame& fA, const Frame& fB, int tA, int tB,
                                           BondRes& o) {
    const V a1 = fB.a1, a2 = fB.a2, a3 = fB.a3, b1 = fA.a1, b2 = fA.a2, b3 = fA.a3;
    V stP = site(P.stack, fB), stQ = site(P.stack, fA);
    V brP = site(P.backref, fB), brQ = site(P.backref, fA);
    V Rpq = -Rc;
    V rs = Rpq + stQ - stP;
    float rsm = sqrtf(dot(rs, rs));
    float eps = P.st_eps[tA * 4 + tB];
    float fr, frD;
    f1_vd(rsm, P.f1, eps, eps * P.shift_factor, fr, frD);
    if (fr == 0.f) return 0.f;
    V rbr = Rpq + brQ - brP;
    float rbm = sqrtf(dot(rbr, rbr));
    float inv = 1.f / rsm, invb = 1.f / rbm;
    V rsd = rs * inv;
    float c4 = dot(a3, b3), c5 = dot(a3, rsd), c6 = -dot(b3, rsd);
    float cp1 = dot(a2, rbr) * invb, cp2 = dot(b2, rbr) * invb;
    float s4, s5, s6;
    float th4 = acos_as(clampc(c4), s4), th5 = acos_as(clampc(-c5), s5), th6 = acos_as(clampc(c6), s6);
    float g4, g4D, g5, g5D, g6, g6D, h1, h1D, h2, h2D;
    f4_vd(th4, s4, P.t4, g4, g4D);
    f4_vd(th5, s5, P.t5, g5, g5D);
    f4_vd(th6, s6, P.t5, g6, g6D);
    f5_vd(cp1, P.p1, h1, h1D);
    f5_vd(cp2, P.p1, h2, h2D);
    float e = fr * g4 * g5 * g6 * h1 * h2;
    if (e == 0.f) return 0.f;
    V fq = rsd * (-frD * g4 * g5 * g6 * h1 * h2);
    fq -= (a3 - rsd * c5) * (fr * g4 * g5D * g6 * h1 * h2 * inv);
    fq -= (b3 + rsd * c6) * (fr * g4 * g5 * g6D * h1 * h2 * inv);
    float gam = P.gamma;
    float rbm3 = rbm * rbm * rbm;
    float ra2 = dot(rsd, a2), ra1 = dot(rsd, a1), rb1 = dot(rsd, b1), a2b1 = dot(a2, b1);
    float par = rsm * ra2 - a2b1 * gam;
    float d1dr = (rsm * rsm * ra2 - ra2 * rbm * rbm - rsm * (a2b1 + ra2 * (-ra1 + rb1)) * gam +
                  a2b1 * (-ra1 + rb1) * gam * gam) / rbm3;
    float d1dra1 = rsm * gam * par / rbm3;
    float d1dra2 = -rsm * invb;
    float d1drb1 = -rsm * gam * par / rbm3;
    float d1da1b1 = -gam * gam * par / rbm3;
    float d1da2b1 = gam * invb;
    float fp1 = -fr * g4 * g5 * g6 * h1D * h2;
    fq -= (rsd * d1dr + ((a2 - rsd * ra2) * d1dra2 + (a1 - rsd * ra1) * d1dra1 + (b1 - rsd * rb1) * d1drb1) * inv) * fp1;
    float sa2 = dot(rsd, b2), sa1 = dot(rsd, b1), sb1 = dot(rsd, a1), s21 = dot(b2, a1);
    float par2 = rsm * sa2 + s21 * gam;
    float d2dr = (par2 * (rsm + (sb1 - sa1) * gam) - sa2 * rbm * rbm) / rbm3;
    float d2dra1 = -rsm * gam * par2 / rbm3;
    float d2dra2 = -rsm * invb;
    float d2drb1 = rsm * gam * par2 / rbm3;
    float d2da1b1 = -gam * gam * par2 / rbm3;
    float d2da2b1 = -gam * invb;
    float fp2 = -fr * g4 * g5 * g6 * h1 * h2D;
    fq -= (rsd * d2dr + ((b2 - rsd * sa2) * d2dra2 + (b1 - rsd * sa1) * d2dra1 + (a1 - rsd * sb1) * d2drb1) * inv) * fp2;
    o.FB -= fq;
    vir_add(o.vir, rs, fq);
    V tp = -cross(stP, fq), tq = cross(stQ, fq);
    V t4d = cross(b3, a3);
    float tm = fr * g4D * g5 * g6 * h1 * h2;
    tp -= t4d * tm; tq += t4d * tm;
    tp += cross(rsd, a3) * (fr * g4 * g5D * g6 * h1 * h2);
    tq += cross(rsd, b3) * (fr * g4 * g5 * g6D * h1 * h2);
    tp += cross(rsd, a2) * (fp1 * d1dra2) + cross(rsd, a1) * (fp1 * d1dra1);
    tq += cross(rsd, b1) * (fp1 * d1drb1);
    V pure = cross(a2, b1) * (fp1 * d1da2b1) + cross(a1, b1) * (fp1 * d1da1b1);
    tp -= pure; tq += pure;
    tp += cross(rsd, a1) * (fp2 * d2drb1);
    tq += cross(rsd, b2) * (fp2 * d2dra2) + cross(rsd, b1) * (fp2 * d2dra1);
    pure = cross(a1, b2) * (fp2 * d2da2b1) + cross(a1, b1) * (fp2 * d2da1b1);
    tp -= pure; tq += pure;
    o.TB += tp;
    o.TA += tq;
    return e;
}

// oxRNA stacking, p = B, q = A (reference src/OxDNABondedKernel.h:238-353).
__device__ __forceinline__ float stack_rna(const DBonded& P, V Rc, const Frame& fA, const Frame& fB, int tA, int tB,
                                           BondRes& o) {
    const V a2 = fB.a2, a3 = fB.a3, b2 = fA.a2, b3 = fA.a3;
    V stP = site(P.stack3, fB), stQ = site(P.stack5, fA);
    V Rpq = -Rc;
    V rs = Rpq + stQ - stP;
    float rsm = sqrtf(dot(rs, rs));
    float eps = P.st_eps[tA * 4 + tB];
    float fr, frD;
    f1_vd(rsm, P.f1, eps, eps * P.shift_factor, fr, frD);
    if (fr == 0.f) return 0.f;
    V bkP = site(P.back, fB), bkQ = site(P.back, fA);
    V bvP = site(P.bbv3, fB), bvQ = site(P.bbv5, fA);
    float inv = 1.f / rsm;
    V rsd = rs * inv;
    V rb = Rpq + bkQ - bkP;
    float rbm = sqrtf(dot(rb, rb));
    float invb = 1.f / rbm;
    V rbd = rb * invb;
    float c5 = dot(a3, rsd), c6 = -dot(b3, rsd);
    float cB1 = -dot(rbd, bvP), cB2 = -dot(rbd, bvQ);
    float cp1 = dot(a2, rb) * invb, cp2 = dot(b2, rb) * invb;
    float s5, s6, sB1, sB2;
    float th5 = acos_as(clampc(-c5), s5), th6 = acos_as(clampc(c6), s6);
    float thB1 = acos_as(clampc(cB1), sB1), thB2 = acos_as(clampc(cB2), sB2);
    float g5, g5D, g6, g6D, gB1, gB1D, gB2, gB2D, h1, h1D, h2, h2D;
    f4_vd(th5, s5, P.t5, g5, g5D);
    f4_vd(th6, s6, P.t5, g6, g6D);
    f4_vd(thB1, sB1, P.tB1, gB1, gB1D);
    f4_vd(thB2, sB2, P.tB2, gB2, gB2D);
    f5_vd(cp1, P.p1, h1, h1D);
    f5_vd(cp2, P.p1, h2, h2D);
    float e = fr * g5 * g6 * gB1 * gB2 * h1 * h2;
    if (e == 0.f) return 0.f;
    V fq = rsd * (-frD * g5 * g6 * gB1 * gB2 * h1 * h2);
    fq -= (a3 - rsd * c5) * (fr * g5D * g6 * gB1 * gB2 * h1 * h2 * inv);
    fq -= (b3 + rsd * c6) * (fr * g5 * g6D * gB1 * gB2 * h1 * h2 * inv);
    V pb = (bvP + rbd * cB1) * (-fr * g5 * g6 * gB1D * gB2 * h1 * h2 * invb);
    pb -= (bvQ + rbd * cB2) * (fr * g5 * g6 * gB1 * gB2D * h1 * h2 * invb);
    float fp1 = -fr * g5 * g6 * gB1 * gB2 * h1D * h2;
    pb += (a2 - rb * (cp1 * invb)) * (fp1 * invb);
    float fp2 = -fr * g5 * g6 * gB1 * gB2 * h1 * h2D;
    pb += (b2 - rb * (cp2 * invb)) * (fp2 * invb);
    o.FB -= fq + pb;
    vir_add(o.vir, rs, fq);
    vir_add(o.vir, rb, pb);
    V tp = -cross(stP, fq) - cross(bkP, pb);
    V tq = cross(stQ, fq) + cross(bkQ, pb);
    tp += cross(rsd, a3) * (fr * g5D * g6 * gB1 * gB2 * h1 * h2);
    tq += cross(rsd, b3) * (fr * g5 * g6D * gB1 * gB2 * h1 * h2);
    tp += cross(rbd, bvP) * (fr * g5 * g6 * gB1D * gB2 * h1 * h2);
    tq += cross(rbd, bvQ) * (fr * g5 * g6 * gB1 * gB2D * h1 * h2);
    tp += cross(a2, rbd) * fp1;
    tq += cross(b2, rbd) * fp2;
    o.TB += tp;
    o.TA += tq;
    return e;
}

// Full bond (reference src/OxDNABondedKernel.h:357-431): FENE with FP64 arithmetic on
// the FP64 minimum-image vector Rc = r_B - r_A, then bonded excluded volume and
// stacking in FP32.
// energy = false: forces and torques only (the FENE energy, an FP64 logarithm, is left
// at 0 unless the bond is broken).
__device__ __forceinline__ void bond_eval(const DBonded& P, double Rx, double Ry, double Rz, const Frame& fA,
                                          const Frame& fB, int tA, int tB, BondRes& o, bool energy = true) {
    o.FB = zero3(); o.TA = zero3(); o.TB = zero3();
    o.e_fene = 0.0; o.e_bexc = 0.f; o.e_stack = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) o.vir[k] = 0.f;
    V bA = site(P.back, fA), bB = site(P.back, fB);
    double rx = Rx + (double)bB.x - (double)bA.x;
    double ry = Ry + (double)bB.y - (double)bA.y;
    double rz = Rz + (double)bB.z - (double)bA.z;
    double rm = sqrt(rx * rx + ry * ry + rz * rz);
    double d = rm - P.r0_backbone;
    if (!(fabs(d) < P.delta_backbone)) {
        o.e_fene = 1.0e12;  // broken bond: energy sentinel, no force
    } else {
        if (energy) o.e_fene = -0.5 * P.eps_backbone * log(1.0 - d * d / P.delta2);
        double pre = -(P.eps_backbone * d) / ((P.delta2 - d * d) * rm);
        V f = mk((float)(pre * rx), (float)(pre * ry), (float)(pre * rz));
        o.FB += f;
        o.TA -= cross(bA, f);
        o.TB += cross(bB, f);
        vir_add(o.vir, mk((float)rx, (float)ry, (float)rz), f);
    }
    V Rc = mk((float)Rx, (float)Ry, (float)Rz);
    V eA = site(P.base, fA), eB = site(P.base, fB);
    const BondExc bx = eager(P.bx);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        V si = (k == 2) ? bA : eA;
        V sj = (k == 1) ? bB : eB;
        V rc = Rc + sj - si;
        V f;
        o.e_bexc += lj(rc, bx.eps, bx.sigma[k], bx.rstar[k], bx.b[k], bx.rc[k], f);
        o.FB += f;
        o.TA -= cross(si, f);
        o.TB += cross(sj, f);
        vir_add(o.vir, rc, f);
    }
    if (P.stacking_enabled) {
        o.e_stack = P.rna_stacking ? stack_rna(P, Rc, fA, fB, tA, tB, o) : stack_dna(P, Rc, fA, fB, tA, tB, o);
    }
}

}  // namespace oxd
