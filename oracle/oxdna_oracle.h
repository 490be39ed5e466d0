// oxdna_oracle.h — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the oxDNA1 / oxDNA2 / oxRNA2 force evaluation of the reference
// plugin (glab-vbc/hoomd-oxDNA), written for checking the HIP engine. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
// the checker / the CPU baseline, never as the product path.
//
// Every routine is templated on the working real type R. R = double is the FP64
// oracle; R = census::Real counts floating-point operations (the op census that
// defines the algorithmic flops of each term). Each function cites the reference
// file:line it restates. Parameters arrive as the C-ABI POD structs of
// include/oxdna_hip.h (FP64 mirrors of the reference's parameter structs).
//
// Pinning (what checks each part; tests/test_oracle_*.py):
//   * primitives fene/repulsive_lj/f1/f2/f4/f5: closed forms at sample points and
//     C0/C1 continuity at every smoothing knot (the reference's derived constants are
//     built to make this hold);
//   * every term's energy: closed-form special cases (ideal geometries with every
//     angle at theta0, beyond-cut-off zeros, non-complementary HB = 0, zero charge);
//   * forces and torques: central finite differences of the oracle's own energy
//     (the reference's fd_check method, pytest/_helpers.py:137-181), Newton's third
//     law, total angular momentum, global rotation / translation invariance;
//   * neighbour pair sets: brute-force O(N^2) enumeration vs the cell-list builder.
//   * the f4Dsin small-angle fallback: equal to the limit of the inner branch as
//     sin(theta) -> 0 (theta0 at 0 or pi).
// Parity unpinned (no golden data exists offline): absolute agreement with oxDNA's
// split_energy.dat, and reference-specific conventions that no invariant fixes (the
// coaxial theta1 fallback 2 sa, where the branch it replaces diverges).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "oxdna_hip.h"

namespace oracle {

template<class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    template<class S> static V3 from(const S* p) { return V3(R(p[0]), R(p[1]), R(p[2])); }
};
template<class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template<class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template<class R> inline V3<R> operator-(const V3<R>& a) { return V3<R>(-a.x, -a.y, -a.z); }
template<class R> inline V3<R> operator*(const V3<R>& a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template<class R> inline V3<R> operator*(R s, const V3<R>& a) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template<class R> inline V3<R> operator/(const V3<R>& a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template<class R> inline V3<R>& operator+=(V3<R>& a, const V3<R>& b) { a = a + b; return a; }
template<class R> inline V3<R>& operator-=(V3<R>& a, const V3<R>& b) { a = a - b; return a; }
template<class R> inline R dot(const V3<R>& a, const V3<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template<class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Unit quaternion (w, x, y, z) rotating body-frame vectors into the lab frame:
// rotate(q, e_x) = a1, rotate(q, e_y) = a2 = a3 x a1, rotate(q, e_z) = a3
// (reference oxdna/io.py:1-8, 117-150). Same algebra as HOOMD's rotate(q, v).
template<class R> struct Quat {
    R s; V3<R> v;
    Quat() : s(1) {}
    explicit Quat(const double* q) : s(R(q[0])), v(R(q[1]), R(q[2]), R(q[3])) {}
};
template<class R> inline V3<R> rotate(const Quat<R>& q, const V3<R>& b) {
    return (q.s * q.s - dot(q.v, q.v)) * b + (R(2) * q.s) * cross(q.v, b) + (R(2) * dot(q.v, b)) * q.v;
}

// ---------------------------------------------------------------------------
// Building-block functions (reference src/OxDNANumerics.h).
// ---------------------------------------------------------------------------
using std::acos; using std::exp; using std::log; using std::sin; using std::sqrt; using std::fabs;

// reference src/OxDNANumerics.h:50-54
template<class R> inline R fene_energy(R dr, R eps, R delta2) {
    return R(-0.5) * eps * log(R(1) - dr * dr / delta2);
}
// reference src/OxDNANumerics.h:58-62 (force on the second site = prefactor * rvec)
template<class R> inline R fene_force_prefactor(R dr, R rmod, R eps, R delta2) {
    return -(eps * dr / (delta2 - dr * dr)) / rmod;
}
// reference src/OxDNANumerics.h:76-107; r = site_q - site_p, out = force on site q.
template<class R> inline R repulsive_lj(const V3<R>& r, R eps, R sigma, R rstar, R b, R rc, V3<R>& out) {
    R rr = dot(r, r);
    out = V3<R>();
    if (!(rr < rc * rc)) return R(0);
    if (rr > rstar * rstar) {
        R rmod = sqrt(rr);
        R d = rmod - rc;
        out = (R(-2) * eps * b * d / rmod) * r;
        return eps * b * d * d;
    }
    R s2 = sigma * sigma / rr;
    R s6 = s2 * s2 * s2;
    out = (R(-24) * eps * (s6 - R(2) * s6 * s6) / rr) * r;
    return R(4) * eps * (s6 * s6 - s6);
}
// Morse radial f1 and derivative (reference src/OxDNANumerics.h:114-155).
struct F1Shape { double a, r0, rlow, rhigh, rclow, rchigh, blow, bhigh; };
template<class R> inline R f1(R r, const F1Shape& s, R eps, R shift) {
    if (!(r < R(s.rchigh))) return R(0);
    if (r > R(s.rhigh)) { R d = r - R(s.rchigh); return eps * R(s.bhigh) * d * d; }
    if (r > R(s.rlow)) { R t = R(1) - exp(-(r - R(s.r0)) * R(s.a)); return eps * t * t - shift; }
    if (r > R(s.rclow)) { R d = r - R(s.rclow); return eps * R(s.blow) * d * d; }
    return R(0);
}
template<class R> inline R f1D(R r, const F1Shape& s, R eps) {
    if (!(r < R(s.rchigh))) return R(0);
    if (r > R(s.rhigh)) return eps * R(2) * R(s.bhigh) * (r - R(s.rchigh));
    if (r > R(s.rlow)) { R e = exp(-(r - R(s.r0)) * R(s.a)); return eps * R(2) * (R(1) - e) * e * R(s.a); }
    if (r > R(s.rclow)) return eps * R(2) * R(s.blow) * (r - R(s.rclow));
    return R(0);
}
// Harmonic radial f2 and derivative (reference src/OxDNANumerics.h:158-192).
struct F2Shape { double k, r0, rc, rlow, rhigh, rclow, rchigh, blow, bhigh; };
template<class R> inline R f2(R r, const F2Shape& s) {
    if (!(r < R(s.rchigh))) return R(0);
    if (r > R(s.rhigh)) { R d = r - R(s.rchigh); return R(s.k) * R(s.bhigh) * d * d; }
    if (r > R(s.rlow)) {
        R d = r - R(s.r0), dc = R(s.rc) - R(s.r0);
        return (R(s.k) / R(2)) * (d * d - dc * dc);
    }
    if (r > R(s.rclow)) { R d = r - R(s.rclow); return R(s.k) * R(s.blow) * d * d; }
    return R(0);
}
template<class R> inline R f2D(R r, const F2Shape& s) {
    if (!(r < R(s.rchigh))) return R(0);
    if (r > R(s.rhigh)) return R(2) * R(s.k) * R(s.bhigh) * (r - R(s.rchigh));
    if (r > R(s.rlow)) return R(s.k) * (r - R(s.r0));
    if (r > R(s.rclow)) return R(2) * R(s.k) * R(s.blow) * (r - R(s.rclow));
    return R(0);
}
// reference src/OxDNANumerics.h:194-198
template<class R> inline R clamp_cos(R c) { return c > R(1) ? R(1) : (c < R(-1) ? R(-1) : c); }
// Angular modulation f4 of the angle t, set p = (a, b, t0, ts, tc)
// (reference src/OxDNANumerics.h:201-215).
template<class R> inline R f4(R t, const double* p) {
    R d = t - R(p[2]);
    if (d < R(0)) d = -d;
    if (!(d < R(p[4]))) return R(0);
    if (d > R(p[3])) { R e = R(p[4]) - d; return R(p[1]) * e * e; }
    return R(1) - R(p[0]) * d * d;
}
// Torque coefficient (d f4 / d t) / sin(t) (reference src/OxDNANumerics.h:218-244);
// the 1e-8 guard applies only in the inner branch, as in the reference.
template<class R> inline R f4Dsin(R t, const double* p) {
    R d = t - R(p[2]);
    R m = R(1);
    if (d < R(0)) { d = -d; m = R(-1); }
    if (!(d < R(p[4]))) return R(0);
    R st = sin(t);
    if (d > R(p[3])) return m * R(2) * R(p[1]) * (d - R(p[4])) / st;
    if (st * st > R(1e-8)) return -m * R(2) * R(p[0]) * d / st;
    return -m * R(2) * R(p[0]);
}
// Dihedral factor f5 of cos(phi), set p = (a, b, xc, xs) (reference src/OxDNANumerics.h:261-290).
template<class R> inline R f5(R c, const double* p) {
    if (!(c > R(p[2]))) return R(0);
    if (c < R(p[3])) { R d = R(p[2]) - c; return R(p[1]) * d * d; }
    if (c < R(0)) return R(1) - R(p[0]) * c * c;
    return R(1);
}
template<class R> inline R f5D(R c, const double* p) {
    if (!(c > R(p[2]))) return R(0);
    if (c < R(p[3])) return R(2) * R(p[1]) * (c - R(p[2]));
    if (c < R(0)) return R(-2) * R(p[0]) * c;
    return R(0);
}

// ---------------------------------------------------------------------------
// One pair: result of a nonbonded evaluator (force on i, torques on i and j).
// ---------------------------------------------------------------------------
template<class R> struct PairOut {
    V3<R> F;          // force on i
    V3<R> Ti, Tj;     // lab-frame torques
    R E;              // full pair energy
    R E2;             // second part (fused HBondCross: cross-stacking part; E holds HB)
};

// Excluded volume (reference src/EvaluatorPairOxDNAExcVol.h:107-149).
template<class R>
bool eval_excvol(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, R rcutsq,
                 const oxdna_excvol_params& p, PairOut<R>& o) {
    o = PairOut<R>();
    if (!(dot(dr, dr) < rcutsq)) return false;
    V3<R> base_b = V3<R>::from(p.base_site), back_b = V3<R>::from(p.back_site);
    V3<R> base_i = rotate(qi, base_b), back_i = rotate(qi, back_b);
    V3<R> base_j = rotate(qj, base_b), back_j = rotate(qj, back_b);
    V3<R> Rc = -dr;  // r_j - r_i
    const V3<R>* si[4] = {&back_i, &base_i, &base_i, &back_i};
    const V3<R>* sj[4] = {&back_j, &base_j, &back_j, &base_j};
    for (int k = 0; k < 4; ++k) {
        V3<R> rc = Rc + *sj[k] - *si[k];
        V3<R> f;
        o.E = o.E + repulsive_lj(rc, R(p.eps), R(p.sigma[k]), R(p.rstar[k]), R(p.b[k]), R(p.rc[k]), f);
        o.F -= f;
        o.Ti += cross(*si[k], -f);
        o.Tj += cross(*sj[k], f);
    }
    return true;
}

// One angular factor with its torque coefficient; acos once, reflection as
// pi - theta with the derivative subtracted (reference src/OxDNABaseBase.h:52-67).
template<class R> struct Ang { R val, dsin; };
template<class R> inline Ang<R> angular(R cost, const double* t, bool sym, R sign) {
    const R PI = R(3.14159265358979323846);
    R th = acos(clamp_cos(cost));
    Ang<R> a;
    a.val = f4(th, t);
    R d = f4Dsin(th, t);
    if (sym) {
        R thn = PI - th;
        a.val = a.val + f4(thn, t);
        d = d - f4Dsin(thn, t);
    }
    a.dsin = sign * d;
    return a;
}

// Shared BASE-BASE geometry (reference src/OxDNABaseBase.h:74-121).
template<class R> struct BBGeom {
    bool valid;
    V3<R> a1, a3, b1, b3, base_i, base_j, rdir;
    R rmod, c1, c2, c3, c4, c7, c8;
};
template<class R>
BBGeom<R> bb_geometry(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, const double* base_site) {
    BBGeom<R> g;
    g.valid = false;
    g.a1 = rotate(qi, V3<R>(R(1), R(0), R(0)));
    g.a3 = rotate(qi, V3<R>(R(0), R(0), R(1)));
    g.b1 = rotate(qj, V3<R>(R(1), R(0), R(0)));
    g.b3 = rotate(qj, V3<R>(R(0), R(0), R(1)));
    V3<R> s = V3<R>::from(base_site);
    g.base_i = rotate(qi, s);
    g.base_j = rotate(qj, s);
    V3<R> r = -dr + g.base_j - g.base_i;
    g.rmod = sqrt(dot(r, r));
    if (!(g.rmod > R(0))) return g;
    g.rdir = r / g.rmod;
    g.c1 = -dot(g.a1, g.b1);
    g.c2 = -dot(g.b1, g.rdir);
    g.c3 = dot(g.a1, g.rdir);
    g.c4 = dot(g.a3, g.b3);
    g.c7 = -dot(g.b3, g.rdir);
    g.c8 = dot(g.a3, g.rdir);
    g.valid = true;
    return g;
}

// One BASE-BASE term accumulated into (F on i, E, Ti, Tj)
// (reference src/OxDNABaseBase.h:128-219). radial(r, &val, &deriv).
template<class R, class Radial>
void bb_term(const BBGeom<R>& g, const double* t1, const double* t23, const double* t4,
             const double* t78, bool sym, const Radial& radial, bool include_t4,
             V3<R>& F, R& E, V3<R>& Ti, V3<R>& Tj) {
    Ang<R> A1 = angular(g.c1, t1, false, R(-1));
    Ang<R> A2 = angular(g.c2, t23, false, R(-1));
    Ang<R> A3 = angular(g.c3, t23, false, R(1));
    Ang<R> A4 = include_t4 ? angular(g.c4, t4, sym, R(1)) : Ang<R>{R(1), R(0)};
    Ang<R> A7 = angular(g.c7, t78, sym, R(-1));
    Ang<R> A8 = angular(g.c8, t78, sym, R(1));
    R rv, rd;
    radial(g.rmod, rv, rd);
    R prod = A1.val * A2.val * A3.val * A4.val * A7.val * A8.val;
    R e = rv * prod;
    E = E + e;
    if (e == R(0)) return;
    V3<R> fq = -g.rdir * (rd * prod);  // force on q (= j)
    V3<R> tp, tq;
    // theta4
    V3<R> dir = cross(g.a3, g.b3);
    R tm = -rv * A1.val * A2.val * A3.val * A4.dsin * A7.val * A8.val;
    tp -= dir * tm; tq += dir * tm;
    // theta1
    dir = cross(g.a1, g.b1);
    tm = -rv * A1.dsin * A2.val * A3.val * A4.val * A7.val * A8.val;
    tp -= dir * tm; tq += dir * tm;
    // theta2
    R fact = rv * A1.val * A2.dsin * A3.val * A4.val * A7.val * A8.val;
    fq += (g.b1 + g.rdir * g.c2) * (fact / g.rmod);
    tq -= cross(g.rdir, g.b1) * fact;
    // theta3
    fact = rv * A1.val * A2.val * A3.dsin * A4.val * A7.val * A8.val;
    fq += (g.a1 - g.rdir * g.c3) * (fact / g.rmod);
    tp -= cross(g.rdir, g.a1) * fact;
    // theta7
    fact = rv * A1.val * A2.val * A3.val * A4.val * A7.dsin * A8.val;
    fq += (g.b3 + g.rdir * g.c7) * (fact / g.rmod);
    tq -= cross(g.rdir, g.b3) * fact;
    // theta8
    fact = rv * A1.val * A2.val * A3.val * A4.val * A7.val * A8.dsin;
    fq += (g.a3 - g.rdir * g.c8) * (fact / g.rmod);
    tp -= cross(g.rdir, g.a3) * fact;
    // site offsets
    tp -= cross(g.base_i, fq);
    tq += cross(g.base_j, fq);
    F -= fq;
    Ti += tp;
    Tj += tq;
}

template<class R> struct RadialF1 {
    F1Shape s; R eps, shift;
    void operator()(R r, R& v, R& d) const { v = f1(r, s, eps, shift); d = f1D(r, s, eps); }
};
template<class R> struct RadialF2 {
    F2Shape s;
    void operator()(R r, R& v, R& d) const { v = f2(r, s); d = f2D(r, s); }
};

// Hydrogen bonding (reference src/EvaluatorPairOxDNAHBond.h:104-116).
template<class R>
bool eval_hbond(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, R rcutsq,
                const oxdna_hbond_params& p, PairOut<R>& o) {
    o = PairOut<R>();
    if (!(dot(dr, dr) < rcutsq)) return false;
    BBGeom<R> g = bb_geometry(dr, qi, qj, p.base_site);
    if (!g.valid) return true;
    RadialF1<R> rad{{p.f1_a, p.f1_r0, p.f1_rlow, p.f1_rhigh, p.f1_rclow, p.f1_rchigh, p.f1_blow, p.f1_bhigh},
                    R(p.eps), R(p.eps * p.shift_factor)};
    bb_term(g, p.t1, p.t23, p.t4, p.t78, false, rad, true, o.F, o.E, o.Ti, o.Tj);
    return true;
}

// Cross-stacking (reference src/EvaluatorPairOxDNACrossStacking.h:101-113).
template<class R>
bool eval_cross(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, R rcutsq,
                const oxdna_crossstacking_params& p, PairOut<R>& o) {
    o = PairOut<R>();
    if (!(dot(dr, dr) < rcutsq)) return false;
    BBGeom<R> g = bb_geometry(dr, qi, qj, p.base_site);
    if (!g.valid) return true;
    RadialF2<R> rad{{p.f2_k, p.f2_r0, p.f2_rc, p.f2_rlow, p.f2_rhigh, p.f2_rclow, p.f2_rchigh, p.f2_blow, p.f2_bhigh}};
    bb_term(g, p.t1, p.t23, p.t4, p.t78, true, rad, p.t4_enabled != 0, o.F, o.E, o.Ti, o.Tj);
    return true;
}

// Fused HB + cross-stacking sharing one geometry (reference
// src/EvaluatorPairOxDNAHBondCross.h:261-301); o.E = HB part, o.E2 = CS part.
template<class R>
bool eval_hbondcross(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, R rcutsq,
                     const oxdna_hbondcross_params& p, PairOut<R>& o) {
    o = PairOut<R>();
    if (!(dot(dr, dr) < rcutsq)) return false;
    BBGeom<R> g = bb_geometry(dr, qi, qj, p.base_site);
    if (!g.valid) return true;
    RadialF1<R> hb{{p.f1_a, p.f1_r0, p.f1_rlow, p.f1_rhigh, p.f1_rclow, p.f1_rchigh, p.f1_blow, p.f1_bhigh},
                   R(p.eps), R(p.eps * p.shift_factor)};
    bb_term(g, p.hb_t1, p.hb_t23, p.hb_t4, p.hb_t78, false, hb, true, o.F, o.E, o.Ti, o.Tj);
    RadialF2<R> cs{{p.f2_k, p.f2_r0, p.f2_rc, p.f2_rlow, p.f2_rhigh, p.f2_rclow, p.f2_rchigh, p.f2_blow, p.f2_bhigh}};
    bb_term(g, p.cs_t1, p.cs_t23, p.cs_t4, p.cs_t78, true, cs, p.cs_t4_enabled != 0, o.F, o.E2, o.Ti, o.Tj);
    return true;
}

// Coaxial stacking, oxDNA1 / oxDNA2 / oxRNA forms
// (reference src/EvaluatorPairOxDNACoaxStacking.h:134-372).
template<class R>
bool eval_coax(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, R rcutsq,
               const oxdna_coaxstacking_params& p, PairOut<R>& o) {
    const R PI = R(3.14159265358979323846);
    o = PairOut<R>();
    if (!(dot(dr, dr) < rcutsq)) return false;
    V3<R> a1 = rotate(qi, V3<R>(R(1), R(0), R(0)));
    V3<R> a2 = rotate(qi, V3<R>(R(0), R(1), R(0)));
    V3<R> a3 = rotate(qi, V3<R>(R(0), R(0), R(1)));
    V3<R> b1 = rotate(qj, V3<R>(R(1), R(0), R(0)));
    V3<R> b3 = rotate(qj, V3<R>(R(0), R(0), R(1)));
    V3<R> ss = V3<R>::from(p.stack_site), bs = V3<R>::from(p.stack_back_ref);
    V3<R> st_i = rotate(qi, ss), st_j = rotate(qj, ss);
    V3<R> br_i = rotate(qi, bs), br_j = rotate(qj, bs);
    V3<R> Rc = -dr;
    V3<R> rs = Rc + st_j - st_i;
    R rsm = sqrt(dot(rs, rs));
    if (!(rsm > R(0))) return true;
    V3<R> rsd = rs / rsm;
    R c1 = -dot(a1, b1), c4 = dot(a3, b3), c5 = dot(a3, rsd), c6 = -dot(b3, rsd);
    V3<R> rb = Rc + br_j - br_i;
    R rbm = sqrt(dot(rb, rb));
    V3<R> rbd = rb / rbm;
    R cphi3 = dot(rsd, cross(rbd, a1));
    F2Shape s2{p.f2_k, p.f2_r0, p.f2_rc, p.f2_rlow, p.f2_rhigh, p.f2_rclow, p.f2_rchigh, p.f2_blow, p.f2_bhigh};
    R sa = R(p.t1_sa), sb = R(p.t1_sb), gam = R(p.gamma);
    R fr = f2(rsm, s2);
    R th1 = acos(clamp_cos(c1)), th4 = acos(clamp_cos(c4));
    R th5 = acos(clamp_cos(c5)), th6 = acos(clamp_cos(c6));
    R th5n = PI - th5, th6n = PI - th6;
    R g1 = f4(th1, p.t1);
    if (p.t1_mode == 0) g1 = g1 + f4(R(2) * PI - th1, p.t1);
    else if (th1 > sb) g1 = g1 + sa * (th1 - sb) * (th1 - sb);
    R g4 = f4(th4, p.t4);
    R g5 = f4(th5, p.t56) + f4(th5n, p.t56);
    R g6 = f4(th6, p.t56) + f4(th6n, p.t56);
    R f5p3 = R(0), f5p4 = R(0), cphi4 = R(0), phi = R(1);
    if (p.rna_coax) {
        cphi4 = dot(rsd, cross(rbd, b1));
        f5p3 = f5(cphi3, p.phi3);
        f5p4 = f5(cphi4, p.phi4);
        phi = f5p3 * f5p4;
    } else if (p.phi3_enabled) {
        f5p3 = f5(cphi3, p.phi3);
        phi = f5p3 * f5p3;
    }
    R e = fr * g1 * g4 * g5 * g6 * phi;
    o.E = e;
    if (e == R(0)) return true;
    R frD = f2D(rsm, s2);
    R g1D = -f4Dsin(th1, p.t1);
    if (p.t1_mode == 0) g1D = g1D - f4Dsin(R(2) * PI - th1, p.t1);
    else if (th1 > sb) {
        R st = sin(th1);
        R hd = (st * st > R(1e-8)) ? R(2) * sa * (th1 - sb) / st : R(2) * sa;
        g1D = g1D - hd;
    }
    R g4D = f4Dsin(th4, p.t4);
    R g5D = f4Dsin(th5, p.t56) - f4Dsin(th5n, p.t56);
    R g6D = -f4Dsin(th6, p.t56) + f4Dsin(th6n, p.t56);
    V3<R> fq = -rsd * (frD * g1 * g4 * g5 * g6 * phi);  // force on q (= j)
    V3<R> tp, tq;
    V3<R> dir = cross(a1, b1);
    R tm = -fr * g1D * g4 * g5 * g6 * phi;
    tp -= dir * tm; tq += dir * tm;
    dir = cross(a3, b3);
    tm = -fr * g1 * g4D * g5 * g6 * phi;
    tp -= dir * tm; tq += dir * tm;
    R fact = fr * g1 * g4 * g5D * g6 * phi;
    fq += fact * (a3 - rsd * c5) / rsm;
    tp -= cross(rsd, a3) * fact;
    fact = fr * g1 * g4 * g5 * g6D * phi;
    fq += (b3 + rsd * c6) * (fact / rsm);
    tq -= cross(rsd, b3) * fact;
    if (p.rna_coax) {
        // oxRNA: two dihedrals from the real BACK-BACK vector, compact cross-product
        // form (reference src/EvaluatorPairOxDNACoaxStacking.h:282-317).
        tp -= cross(st_i, fq);
        tq += cross(st_j, fq);
        R base = fr * g1 * g4 * g5 * g6;
        R d3 = f5D(cphi3, p.phi3), d4 = f5D(cphi4, p.phi4);
        for (int w = 0; w < 2; ++w) {
            V3<R> v = (w == 0) ? a1 : b1;
            R fc = (w == 0) ? base * f5p4 * d3 : base * f5p3 * d4;
            V3<R> rbv = cross(rbd, v);
            V3<R> vrs = cross(v, rsd);
            V3<R> fst = -(rbv - rsd * dot(rsd, rbv)) * (fc / rsm);
            V3<R> fbk = -(vrs - rbd * dot(rbd, vrs)) * (fc / rbm);
            fq += fst + fbk;
            tp -= cross(st_i, fst) + cross(br_i, fbk);
            tq += cross(st_j, fst) + cross(br_j, fbk);
            V3<R> pure = fc * cross(v, cross(rsd, rbd));
            if (w == 0) tp -= pure; else tq -= pure;
        }
        o.F = -fq; o.Ti = tp; o.Tj = tq;
        return true;
    }
    if (p.phi3_enabled) {
        // oxDNA1 cos(phi3) dihedral (reference src/EvaluatorPairOxDNACoaxStacking.h:320-361).
        R d3 = f5D(cphi3, p.phi3);
        R g3 = gam * gam * gam;
        R rbm3 = rbm * rbm * rbm;
        R a2b1 = dot(a2, b1), a3b1 = dot(a3, b1);
        R ra1 = dot(rsd, a1), ra2 = dot(rsd, a2), ra3 = dot(rsd, a3), rb1 = dot(rsd, b1);
        R par = ra3 * a2b1 - ra2 * a3b1;
        R dcdr = -gam * par * (gam * (ra1 - rb1) + rsm) / rbm3;
        R dcda1b1 = g3 * par / rbm3;
        R dcda2b1 = gam * ra3 / rbm;
        R dcda3b1 = -gam * ra2 / rbm;
        R dcdra1 = -gam * gam * par * rsm / rbm3;
        R dcdra2 = -gam * a3b1 / rbm;
        R dcdra3 = gam * a2b1 / rbm;
        R dcdrb1 = gam * gam * par * rsm / rbm3;
        R fc = fr * g1 * g4 * g5 * g6 * R(2) * f5p3 * d3;
        fq -= fc * (rsd * dcdr + ((a1 - rsd * ra1) * dcdra1 + (a2 - rsd * ra2) * dcdra2 +
                                  (a3 - rsd * ra3) * dcdra3 + (b1 - rsd * rb1) * dcdrb1) / rsm);
        tp += fc * (cross(rsd, a1) * dcdra1 + cross(rsd, a2) * dcdra2 + cross(rsd, a3) * dcdra3);
        tq += fc * (cross(rsd, b1) * dcdrb1);
        V3<R> pure = fc * (cross(a1, b1) * dcda1b1 + cross(a2, b1) * dcda2b1 + cross(a3, b1) * dcda3b1);
        tp -= pure; tq += pure;
    }
    tp -= cross(st_i, fq);
    tq += cross(st_j, fq);
    o.F = -fq; o.Ti = tp; o.Tj = tq;
    return true;
}

// Debye-Hueckel between BACK sites, times q_i q_j
// (reference src/EvaluatorPairOxDNADebye.h:85-141).
template<class R>
bool eval_debye(const V3<R>& dr, const Quat<R>& qi, const Quat<R>& qj, R rcutsq,
                const oxdna_debye_params& p, R chg_i, R chg_j, PairOut<R>& o) {
    o = PairOut<R>();
    if (!(dot(dr, dr) < rcutsq)) return false;
    R rc = R(p.rc), rh = R(p.rhigh), mk = R(p.minus_kappa), pf = R(p.prefactor), bs = R(p.b_smooth);
    V3<R> bsite = V3<R>::from(p.back_site);
    V3<R> bi = rotate(qi, bsite), bj = rotate(qj, bsite);
    V3<R> rb = -dr + bj - bi;
    R rm = sqrt(dot(rb, rb));
    if (!(rm > R(0)) || !(rm < rc)) return true;
    R cf = chg_i * chg_j;
    V3<R> rd = rb / rm;
    R e;
    V3<R> fq;
    if (rm < rh) {
        R ex = exp(rm * mk);
        e = ex * (pf / rm);
        fq = rd * (R(-1) * (pf * ex) * (mk / rm - R(1) / (rm * rm)));
    } else {
        R d = rm - rc;
        e = bs * d * d;
        fq = -rd * (R(2) * bs * d);
    }
    e = e * cf;
    fq = fq * cf;
    o.E = e;
    o.Ti = -cross(bi, fq);
    o.Tj = cross(bj, fq);
    o.F = -fq;
    return true;
}

// ---------------------------------------------------------------------------
// Bonded pair (reference src/OxDNABondedKernel.h:38-431).
// ---------------------------------------------------------------------------
template<class R> struct BondOut {
    V3<R> FB;  // force on B (force on A = -FB)
    V3<R> TA, TB;
    R e_fene, e_bexc, e_stack;
    R vir[6];
};
// reference src/OxDNABondedKernel.h:48-58
template<class R> inline void acc_virial(R vir[6], const V3<R>& r, const V3<R>& f) {
    vir[0] = vir[0] + R(0.5) * r.x * f.x; vir[1] = vir[1] + R(0.5) * r.x * f.y;
    vir[2] = vir[2] + R(0.5) * r.x * f.z; vir[3] = vir[3] + R(0.5) * r.y * f.y;
    vir[4] = vir[4] + R(0.5) * r.y * f.z; vir[5] = vir[5] + R(0.5) * r.z * f.z;
}

// oxDNA stacking; p = B, q = A (reference src/OxDNABondedKernel.h:69-231).
template<class R>
R bonded_stacking(const oxdna_bonded_params& P, const V3<R>& Rc, const Quat<R>& qA, const Quat<R>& qB,
                  int tA, int tB, BondOut<R>& o) {
    V3<R> a1 = rotate(qB, V3<R>(R(1), R(0), R(0))), a2 = rotate(qB, V3<R>(R(0), R(1), R(0)));
    V3<R> a3 = rotate(qB, V3<R>(R(0), R(0), R(1)));
    V3<R> b1 = rotate(qA, V3<R>(R(1), R(0), R(0))), b2 = rotate(qA, V3<R>(R(0), R(1), R(0)));
    V3<R> b3 = rotate(qA, V3<R>(R(0), R(0), R(1)));
    V3<R> ss = V3<R>::from(P.stack_site), bs = V3<R>::from(P.stack_back_ref);
    V3<R> stP = rotate(qB, ss), stQ = rotate(qA, ss), brP = rotate(qB, bs), brQ = rotate(qA, bs);
    V3<R> Rpq = -Rc;  // posA - posB
    V3<R> rbr = Rpq + brQ - brP;
    R rbm = sqrt(dot(rbr, rbr));
    V3<R> rs = Rpq + stQ - stP;
    R rsm = sqrt(dot(rs, rs));
    V3<R> rsd = rs / rsm;
    R c4 = dot(a3, b3), c5 = dot(a3, rsd), c6 = -dot(b3, rsd);
    R cp1 = dot(a2, rbr) / rbm, cp2 = dot(b2, rbr) / rbm;
    R eps = R(P.st_eps[tA * 4 + tB]);
    R shift = eps * R(P.st_shift_factor);
    R th4 = acos(clamp_cos(c4)), th5 = acos(clamp_cos(-c5)), th6 = acos(clamp_cos(c6));
    F1Shape s1{P.st_a, P.st_r0, P.st_rlow, P.st_rhigh, P.st_rclow, P.st_rchigh, P.st_blow, P.st_bhigh};
    R fr = f1(rsm, s1, eps, shift);
    R g4 = f4(th4, P.t4), g5 = f4(th5, P.t5), g6 = f4(th6, P.t5);
    R h1 = f5(cp1, P.p1), h2 = f5(cp2, P.p1);
    R e = fr * g4 * g5 * g6 * h1 * h2;
    if (e == R(0)) return e;
    R frD = f1D(rsm, s1, eps);
    R g4D = f4Dsin(th4, P.t4), g5D = f4Dsin(th5, P.t5), g6D = f4Dsin(th6, P.t5);
    R h1D = f5D(cp1, P.p1), h2D = f5D(cp2, P.p1);
    V3<R> fq = -rsd * (frD * g4 * g5 * g6 * h1 * h2);  // force on q (= A)
    fq -= (a3 - rsd * c5) * (fr * g4 * g5D * g6 * h1 * h2 / rsm);
    fq -= (b3 + rsd * c6) * (fr * g4 * g5 * g6D * h1 * h2 / rsm);
    R gam = R(P.st_gamma);
    R rbm3 = rbm * rbm * rbm;
    // cos(phi1)
    R ra2 = dot(rsd, a2), ra1 = dot(rsd, a1), rb1 = dot(rsd, b1), a2b1 = dot(a2, b1);
    R par = rsm * ra2 - a2b1 * gam;
    R d1dr = (rsm * rsm * ra2 - ra2 * rbm * rbm - rsm * (a2b1 + ra2 * (-ra1 + rb1)) * gam +
              a2b1 * (-ra1 + rb1) * gam * gam) / rbm3;
    R d1dra1 = rsm * gam * par / rbm3;
    R d1dra2 = -rsm / rbm;
    R d1drb1 = -rsm * gam * par / rbm3;
    R d1da1b1 = -gam * gam * par / rbm3;
    R d1da2b1 = gam / rbm;
    R fp1 = -fr * g4 * g5 * g6 * h1D * h2;
    fq -= (rsd * d1dr + ((a2 - rsd * ra2) * d1dra2 + (a1 - rsd * ra1) * d1dra1 + (b1 - rsd * rb1) * d1drb1) / rsm) * fp1;
    // cos(phi2) (p -> b, q -> a)
    R sa2 = dot(rsd, b2), sa1 = dot(rsd, b1), sb1 = dot(rsd, a1), s21 = dot(b2, a1);
    R par2 = rsm * sa2 + s21 * gam;
    R d2dr = (par2 * (rsm + (sb1 - sa1) * gam) - sa2 * rbm * rbm) / rbm3;
    R d2dra1 = -rsm * gam * (rsm * sa2 + s21 * gam) / rbm3;
    R d2dra2 = -rsm / rbm;
    R d2drb1 = rsm * gam * par2 / rbm3;
    R d2da1b1 = -gam * gam * par2 / rbm3;
    R d2da2b1 = -gam / rbm;
    R fp2 = -fr * g4 * g5 * g6 * h1 * h2D;
    fq -= fp2 * (rsd * d2dr + ((b2 - rsd * sa2) * d2dra2 + (b1 - rsd * sa1) * d2dra1 + (a1 - rsd * sb1) * d2drb1) / rsm);
    o.FB -= fq;
    acc_virial(o.vir, rs, fq);
    V3<R> tp, tq;
    tp -= cross(stP, fq);
    tq += cross(stQ, fq);
    V3<R> t4d = cross(b3, a3);
    R tmod = fr * g4D * g5 * g6 * h1 * h2;
    tp -= t4d * tmod; tq += t4d * tmod;
    V3<R> t5d = cross(rsd, a3);
    tmod = -fr * g4 * g5D * g6 * h1 * h2;
    tp -= t5d * tmod;
    V3<R> t6d = cross(rsd, b3);
    tmod = fr * g4 * g5 * g6D * h1 * h2;
    tq += t6d * tmod;
    tp += cross(rsd, a2) * (fp1 * d1dra2) + cross(rsd, a1) * (fp1 * d1dra1);
    tq += cross(rsd, b1) * (fp1 * d1drb1);
    V3<R> pure = cross(a2, b1) * (fp1 * d1da2b1) + cross(a1, b1) * (fp1 * d1da1b1);
    tp -= pure; tq += pure;
    tp += cross(rsd, a1) * (fp2 * d2drb1);
    tq += cross(rsd, b2) * (fp2 * d2dra2) + cross(rsd, b1) * (fp2 * d2dra1);
    pure = cross(a1, b2) * (fp2 * d2da2b1) + cross(a1, b1) * (fp2 * d2da1b1);
    tp -= pure; tq += pure;
    o.TB += tp;
    o.TA += tq;
    return e;
}

// oxRNA stacking (reference src/OxDNABondedKernel.h:238-353). Keeps the reference's
// deliberate f4(thetaB1) f4(thetaB2) factor in force_part_phi2 (:320-321).
template<class R>
R bonded_stacking_rna(const oxdna_bonded_params& P, const V3<R>& Rc, const Quat<R>& qA, const Quat<R>& qB,
                      int tA, int tB, BondOut<R>& o) {
    V3<R> a2 = rotate(qB, V3<R>(R(0), R(1), R(0))), a3 = rotate(qB, V3<R>(R(0), R(0), R(1)));
    V3<R> b2 = rotate(qA, V3<R>(R(0), R(1), R(0))), b3 = rotate(qA, V3<R>(R(0), R(0), R(1)));
    V3<R> stP = rotate(qB, V3<R>::from(P.stack_3_site)), stQ = rotate(qA, V3<R>::from(P.stack_5_site));
    V3<R> bsite = V3<R>::from(P.back_site);
    V3<R> bkP = rotate(qB, bsite), bkQ = rotate(qA, bsite);
    V3<R> bvP = rotate(qB, V3<R>::from(P.bbvector_3)), bvQ = rotate(qA, V3<R>::from(P.bbvector_5));
    V3<R> Rpq = -Rc;
    V3<R> rs = Rpq + stQ - stP;
    R rsm = sqrt(dot(rs, rs));
    V3<R> rsd = rs / rsm;
    V3<R> rb = Rpq + bkQ - bkP;
    R rbm = sqrt(dot(rb, rb));
    V3<R> rbd = rb / rbm;
    R c5 = dot(a3, rsd), c6 = -dot(b3, rsd);
    R cB1 = -dot(rbd, bvP), cB2 = -dot(rbd, bvQ);
    R cp1 = dot(a2, rb) / rbm, cp2 = dot(b2, rb) / rbm;
    R eps = R(P.st_eps[tA * 4 + tB]);
    R shift = eps * R(P.st_shift_factor);
    R th5 = acos(clamp_cos(-c5)), th6 = acos(clamp_cos(c6));
    R thB1 = acos(clamp_cos(cB1)), thB2 = acos(clamp_cos(cB2));
    F1Shape s1{P.st_a, P.st_r0, P.st_rlow, P.st_rhigh, P.st_rclow, P.st_rchigh, P.st_blow, P.st_bhigh};
    R fr = f1(rsm, s1, eps, shift);
    R g5 = f4(th5, P.t5), g6 = f4(th6, P.t5), gB1 = f4(thB1, P.tB1), gB2 = f4(thB2, P.tB2);
    R h1 = f5(cp1, P.p1), h2 = f5(cp2, P.p1);
    R e = fr * g5 * g6 * gB1 * gB2 * h1 * h2;
    if (e == R(0)) return e;
    R frD = f1D(rsm, s1, eps);
    R g5D = f4Dsin(th5, P.t5), g6D = f4Dsin(th6, P.t5), gB1D = f4Dsin(thB1, P.tB1), gB2D = f4Dsin(thB2, P.tB2);
    R h1D = f5D(cp1, P.p1), h2D = f5D(cp2, P.p1);
    V3<R> fq = -rsd * (frD * g5 * g6 * gB1 * gB2 * h1 * h2);
    fq -= (a3 - rsd * c5) * (fr * g5D * g6 * gB1 * gB2 * h1 * h2 / rsm);
    fq -= (b3 + rsd * c6) * (fr * g5 * g6D * gB1 * gB2 * h1 * h2 / rsm);
    V3<R> pb;
    pb -= (bvP + rbd * cB1) * (fr * g5 * g6 * gB1D * gB2 * h1 * h2 / rbm);
    pb -= (bvQ + rbd * cB2) * (fr * g5 * g6 * gB1 * gB2D * h1 * h2 / rbm);
    R fp1 = -fr * g5 * g6 * gB1 * gB2 * h1D * h2;
    pb += (a2 - rb * (cp1 / rbm)) * (fp1 / rbm);
    R fp2 = -fr * g5 * g6 * gB1 * gB2 * h1 * h2D;
    pb += (b2 - rb * (cp2 / rbm)) * (fp2 / rbm);
    o.FB -= fq + pb;
    acc_virial(o.vir, rs, fq);
    acc_virial(o.vir, rb, pb);
    V3<R> tp, tq;
    tp -= cross(stP, fq); tq += cross(stQ, fq);
    tp -= cross(bkP, pb); tq += cross(bkQ, pb);
    tp -= cross(rsd, a3) * (-fr * g5D * g6 * gB1 * gB2 * h1 * h2);
    tq += cross(rsd, b3) * (fr * g5 * g6D * gB1 * gB2 * h1 * h2);
    tp += cross(rbd, bvP) * (fr * g5 * g6 * gB1D * gB2 * h1 * h2);
    tq += cross(rbd, bvQ) * (fr * g5 * g6 * gB1 * gB2D * h1 * h2);
    tp += cross(a2, rbd) * fp1;
    tq += cross(b2, rbd) * fp2;
    o.TB += tp;
    o.TA += tq;
    return e;
}

// Full bonded pair: FENE, bonded excluded volume, stacking
// (reference src/OxDNABondedKernel.h:357-431). Rc = r_B - r_A (minimum image).
// The oracle is FP64 throughout, so the reference's FP64-FENE / FP32-rest split
// collapses to one precision here.
template<class R>
void bonded_pair(const oxdna_bonded_params& P, const V3<R>& Rc, const Quat<R>& qA, const Quat<R>& qB,
                 int tA, int tB, BondOut<R>& o) {
    o = BondOut<R>();
    for (int k = 0; k < 6; ++k) o.vir[k] = R(0);
    o.e_fene = o.e_bexc = o.e_stack = R(0);
    V3<R> bsite = V3<R>::from(P.back_site), esite = V3<R>::from(P.base_site);
    V3<R> bA = rotate(qA, bsite), bB = rotate(qB, bsite);
    V3<R> rb = Rc + bB - bA;
    R rm = sqrt(dot(rb, rb));
    R d = rm - R(P.r0_backbone);
    if (!(fabs(d) < R(P.delta_backbone))) {
        o.e_fene = R(1.0e12);  // broken bond: energy sentinel, no force (reference :381-384)
    } else {
        o.e_fene = fene_energy(d, R(P.eps_backbone), R(P.delta2));
        V3<R> f = fene_force_prefactor(d, rm, R(P.eps_backbone), R(P.delta2)) * rb;
        o.FB += f;
        o.TA += cross(bA, -f);
        o.TB += cross(bB, f);
        acc_virial(o.vir, rb, f);
    }
    V3<R> eA = rotate(qA, esite), eB = rotate(qB, esite);
    const V3<R>* si[3] = {&eA, &eA, &bA};
    const V3<R>* sj[3] = {&eB, &bB, &eB};
    for (int k = 0; k < 3; ++k) {
        V3<R> rc = Rc + *sj[k] - *si[k];
        V3<R> f;
        o.e_bexc = o.e_bexc + repulsive_lj(rc, R(P.excl_eps), R(P.bx_sigma[k]), R(P.bx_rstar[k]),
                                           R(P.bx_b[k]), R(P.bx_rc[k]), f);
        o.FB += f;
        o.TA += cross(*si[k], -f);
        o.TB += cross(*sj[k], f);
        acc_virial(o.vir, rc, f);
    }
    if (P.stacking_enabled) {
        if (P.rna_stacking) o.e_stack = bonded_stacking_rna(P, Rc, qA, qB, tA, tB, o);
        else o.e_stack = bonded_stacking(P, Rc, qA, qB, tA, tB, o);
    }
}

}  // namespace oracle
