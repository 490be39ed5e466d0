// oxdna_oracle.cpp — TEST INFRASTRUCTURE ONLY (see oxdna_oracle.h).
//
// C entry points of the CPU oracle: single pair / single bond evaluation, the
// whole-system driver with the reference's accumulation conventions, neighbour pair
// enumeration (brute force and cell list), and the op census.
//
// System conventions restated here (host MD engine of the reference, HOOMD-blue 6.x;
// parity of these rests on the reference's call sites, not on its source):
//   * dr = r_i - r_j, minimum image dr -= L * rint(dr / L) per axis (FP64, no FMA);
//   * a term is evaluated for every pair i < j that is not bonded
//     (exclusions=("bond",), reference oxdna/model/dna1.py:307-309) with
//     |dr|^2 < rcut(ti,tj)^2; force on i = F, on j = -F, torque Ti / Tj, and each
//     particle receives half the pair energy;
//   * per-particle virial: 1/2 dr_ij (x) F_ij, added to i and (dr_ij, F_ij itself)
//     to j as well, since dr and F both flip sign;
//   * bonded: per bond (A, B) the force on A is -FB, on B FB; each endpoint gets half
//     of each energy and the whole sub-interaction virial
//     (reference src/OxDNABondedForceCompute.cc:76-131).
#include "oxdna_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <algorithm>
#include <array>
#include <cstring>
#include <set>
#include <utility>

using namespace oracle;

extern "C" {

typedef struct oracle_system {
    uint32_t n;
    const double* pos;     // 3n
    const double* quat;    // 4n (w, x, y, z)
    const int32_t* type;   // n
    const double* charge;  // n (NULL = 1)
    double L[3];
    uint32_t nb;
    const int32_t* bonds;  // 2nb (A, B)
    const int32_t* btype;  // nb (NULL = 0)
    int32_t ntypes;
    const void* pair_params[6];  // per nonbonded term id: ntypes^2 table or NULL
    const double* rcut[6];       // per nonbonded term id: ntypes^2 COM cut-offs
    uint32_t nbt;
    const oxdna_bonded_params* bond_params;
} oracle_system;

}  // extern "C"

namespace {

inline double min_image(double d, double L, double Linv) {
    double img = std::rint(d * Linv);
    double t = img * L;
    return d - t;
}

struct PairList {
    std::vector<std::pair<int, int>> pairs;  // i < j
};

double rsq_of(const oracle_system* s, int i, int j, double dr[3]) {
    for (int a = 0; a < 3; ++a) {
        double Linv = 1.0 / s->L[a];
        dr[a] = min_image(s->pos[3 * i + a] - s->pos[3 * j + a], s->L[a], Linv);
    }
    double r2 = dr[0] * dr[0];
    r2 = r2 + dr[1] * dr[1];
    r2 = r2 + dr[2] * dr[2];
    return r2;
}

std::set<std::pair<int, int>> exclusion_set(const oracle_system* s) {
    std::set<std::pair<int, int>> ex;
    for (uint32_t b = 0; b < s->nb; ++b) {
        int a = s->bonds[2 * b], c = s->bonds[2 * b + 1];
        ex.insert(std::make_pair(std::min(a, c), std::max(a, c)));
    }
    return ex;
}

// All non-excluded pairs i < j with |dr|^2 < r_list^2, brute force.
void pairs_brute(const oracle_system* s, double r_list, PairList& out) {
    auto ex = exclusion_set(s);
    double rl2 = r_list * r_list;
    for (uint32_t i = 0; i < s->n; ++i)
        for (uint32_t j = i + 1; j < s->n; ++j) {
            double dr[3];
            if (rsq_of(s, i, j, dr) < rl2 && !ex.count(std::make_pair((int)i, (int)j)))
                out.pairs.emplace_back(i, j);
        }
}

// Same set through a cell list (cells of side >= r_list; needs >= 3 cells per axis).
void pairs_cells(const oracle_system* s, double r_list, PairList& out) {
    int nc[3];
    for (int a = 0; a < 3; ++a) {
        nc[a] = (int)std::floor(s->L[a] / r_list);
        if (nc[a] < 3) { pairs_brute(s, r_list, out); return; }
    }
    auto ex = exclusion_set(s);
    size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
    std::vector<std::vector<int>> cells(ncell);
    auto cell_of = [&](int i, int c[3]) {
        for (int a = 0; a < 3; ++a) {
            double x = s->pos[3 * i + a] / s->L[a] + 0.5;  // box centred at the origin
            x -= std::floor(x);
            int k = (int)(x * nc[a]);
            if (k >= nc[a]) k = nc[a] - 1;
            if (k < 0) k = 0;
            c[a] = k;
        }
    };
    for (uint32_t i = 0; i < s->n; ++i) {
        int c[3];
        cell_of(i, c);
        cells[((size_t)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back(i);
    }
    double rl2 = r_list * r_list;
    for (uint32_t i = 0; i < s->n; ++i) {
        int c[3];
        cell_of(i, c);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int x = (c[0] + dx + nc[0]) % nc[0], y = (c[1] + dy + nc[1]) % nc[1],
                        z = (c[2] + dz + nc[2]) % nc[2];
                    for (int j : cells[((size_t)z * nc[1] + y) * nc[0] + x]) {
                        if (j <= (int)i) continue;
                        double dr[3];
                        if (rsq_of(s, i, j, dr) < rl2 && !ex.count(std::make_pair((int)i, j)))
                            out.pairs.emplace_back(i, j);
                    }
                }
    }
    std::sort(out.pairs.begin(), out.pairs.end());
}

double max_rcut(const oracle_system* s, uint32_t mask) {
    double m = 0;
    int nt2 = s->ntypes * s->ntypes;
    for (int t = 0; t < 6; ++t)
        if ((mask >> t) & 1u)
            if (s->pair_params[t] && s->rcut[t])
                for (int k = 0; k < nt2; ++k) m = std::max(m, s->rcut[t][k]);
    return m;
}

template<class R>
bool eval_term(int term, const oracle_system* s, int i, int j, const V3<R>& dr, PairOut<R>& o) {
    int k = s->type[i] * s->ntypes + s->type[j];
    R rc = R(s->rcut[term][k]);
    R rcsq = rc * rc;
    Quat<R> qi(s->quat + 4 * i), qj(s->quat + 4 * j);
    switch (term) {
    case OXDNA_TERM_EXCVOL:
        return eval_excvol(dr, qi, qj, rcsq, ((const oxdna_excvol_params*)s->pair_params[term])[k], o);
    case OXDNA_TERM_HBOND:
        return eval_hbond(dr, qi, qj, rcsq, ((const oxdna_hbond_params*)s->pair_params[term])[k], o);
    case OXDNA_TERM_CROSS:
        return eval_cross(dr, qi, qj, rcsq, ((const oxdna_crossstacking_params*)s->pair_params[term])[k], o);
    case OXDNA_TERM_COAX:
        return eval_coax(dr, qi, qj, rcsq, ((const oxdna_coaxstacking_params*)s->pair_params[term])[k], o);
    case OXDNA_TERM_DEBYE: {
        R ci = R(s->charge ? s->charge[i] : 1.0), cj = R(s->charge ? s->charge[j] : 1.0);
        return eval_debye(dr, qi, qj, rcsq, ((const oxdna_debye_params*)s->pair_params[term])[k], ci, cj, o);
    }
    case OXDNA_TERM_HBONDCROSS:
        return eval_hbondcross(dr, qi, qj, rcsq, ((const oxdna_hbondcross_params*)s->pair_params[term])[k], o);
    }
    return false;
}

const int kSlot[6] = {OXDNA_E_NEXC, OXDNA_E_HB, OXDNA_E_CRSTACK, OXDNA_E_CXSTACK, OXDNA_E_DEBYE, OXDNA_E_HB};

// Bonded driver (reference src/OxDNABondedForceCompute.cc:76-131).
template<class R>
void bond_eval(const oracle_system* s, uint32_t b, BondOut<R>& r) {
    int A = s->bonds[2 * b], B = s->bonds[2 * b + 1];
    int bt = s->btype ? s->btype[b] : 0;
    double Rc[3];
    for (int a = 0; a < 3; ++a)
        Rc[a] = min_image(s->pos[3 * B + a] - s->pos[3 * A + a], s->L[a], 1.0 / s->L[a]);
    Quat<R> qA(s->quat + 4 * A), qB(s->quat + 4 * B);
    bonded_pair(s->bond_params[bt], V3<R>::from(Rc), qA, qB, s->type[A], s->type[B], r);
}

}  // namespace

extern "C" {

// Whole-system evaluation of the terms in mask (bits of oxdna_term). Outputs are
// overwritten: F[3n], T[3n], Eh[n] (sum of half energies), Eterm[8] (split_energy
// order), Vir[6n]; any output may be NULL. use_cells selects the cell-list pair
// enumeration (same pair set as brute force; checked by tests). Returns 0 or -1.
int oracle_compute(const oracle_system* s, uint32_t mask, int use_cells, double* F, double* T,
                   double* Eh, double* Eterm, double* Vir) {
    uint32_t n = s->n;
    std::vector<double> f(3 * n, 0.0), t(3 * n, 0.0), eh(n, 0.0), vir(6 * n, 0.0);
    double et[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double rl = max_rcut(s, mask);
    if (rl > 0) {
        PairList pl;
        if (use_cells) pairs_cells(s, rl, pl);
        else pairs_brute(s, rl, pl);
        for (auto& pr : pl.pairs) {
            int i = pr.first, j = pr.second;
            double dr[3];
            rsq_of(s, i, j, dr);
            V3<double> d = V3<double>::from(dr);
            for (int term = 0; term < 6; ++term) {
                if (!((mask >> term) & 1u) || !s->pair_params[term]) continue;
                PairOut<double> o;
                if (!eval_term(term, s, i, j, d, o)) continue;
                double e = o.E + o.E2;
                et[kSlot[term]] += o.E;
                if (term == OXDNA_TERM_HBONDCROSS) et[OXDNA_E_CRSTACK] += o.E2;
                eh[i] += 0.5 * e;
                eh[j] += 0.5 * e;
                double Fv[3] = {o.F.x, o.F.y, o.F.z};
                double vv[6] = {dr[0] * Fv[0], dr[0] * Fv[1], dr[0] * Fv[2], dr[1] * Fv[1], dr[1] * Fv[2], dr[2] * Fv[2]};
                for (int a = 0; a < 3; ++a) {
                    f[3 * i + a] += Fv[a];
                    f[3 * j + a] -= Fv[a];
                }
                t[3 * i + 0] += o.Ti.x; t[3 * i + 1] += o.Ti.y; t[3 * i + 2] += o.Ti.z;
                t[3 * j + 0] += o.Tj.x; t[3 * j + 1] += o.Tj.y; t[3 * j + 2] += o.Tj.z;
                for (int k = 0; k < 6; ++k) {
                    vir[6 * i + k] += 0.5 * vv[k];
                    vir[6 * j + k] += 0.5 * vv[k];
                }
            }
        }
    }
    if (((mask >> OXDNA_TERM_BONDED) & 1u) && s->bond_params) {
        for (uint32_t b = 0; b < s->nb; ++b) {
            int A = s->bonds[2 * b], B = s->bonds[2 * b + 1];
            BondOut<double> r;
            bond_eval(s, b, r);
            et[OXDNA_E_FENE] += r.e_fene;
            et[OXDNA_E_BEXC] += r.e_bexc;
            et[OXDNA_E_STACK] += r.e_stack;
            double half = 0.5 * (r.e_fene + r.e_bexc + r.e_stack);
            eh[A] += half; eh[B] += half;
            f[3 * A + 0] -= r.FB.x; f[3 * A + 1] -= r.FB.y; f[3 * A + 2] -= r.FB.z;
            f[3 * B + 0] += r.FB.x; f[3 * B + 1] += r.FB.y; f[3 * B + 2] += r.FB.z;
            t[3 * A + 0] += r.TA.x; t[3 * A + 1] += r.TA.y; t[3 * A + 2] += r.TA.z;
            t[3 * B + 0] += r.TB.x; t[3 * B + 1] += r.TB.y; t[3 * B + 2] += r.TB.z;
            for (int k = 0; k < 6; ++k) {
                vir[6 * A + k] += r.vir[k];
                vir[6 * B + k] += r.vir[k];
            }
        }
    }
    if (F) std::memcpy(F, f.data(), sizeof(double) * 3 * n);
    if (T) std::memcpy(T, t.data(), sizeof(double) * 3 * n);
    if (Eh) std::memcpy(Eh, eh.data(), sizeof(double) * n);
    if (Vir) std::memcpy(Vir, vir.data(), sizeof(double) * 6 * n);
    if (Eterm) std::memcpy(Eterm, et, sizeof(et));
    return 0;
}

// CPU baseline (bench.py's cpu_baseline leg): forces, torques and per-term energies of
// the terms in mask over a given pair list (i < j, e.g. from oracle_pairs with a skin,
// reused between list rebuilds as an MD code does) and the bonds, on nthreads OpenMP
// threads (each with its own accumulators, summed at the end; serial without OpenMP).
// F[3n], T[3n], Eterm[8] are overwritten; any may be NULL.
int oracle_compute_list(const oracle_system* s, uint32_t mask, const int32_t* pairs, size_t npairs, int nthreads,
                        double* F, double* T, double* Eterm) {
    const uint32_t n = s->n;
    int nt = nthreads > 0 ? nthreads : 1;
#ifndef _OPENMP
    nt = 1;
#endif
    std::vector<std::vector<double>> fa(nt, std::vector<double>(3 * n, 0.0)), ta(nt, std::vector<double>(3 * n, 0.0));
    std::vector<std::array<double, 8>> ea(nt);
    for (auto& e : ea) e.fill(0.0);
    const bool bonded = ((mask >> OXDNA_TERM_BONDED) & 1u) && s->bond_params;
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        const int me = omp_get_thread_num();
#else
        const int me = 0;
#endif
        double* f = fa[me].data();
        double* t = ta[me].data();
        double* et = ea[me].data();
#pragma omp for schedule(dynamic, 256) nowait
        for (long long k = 0; k < (long long)npairs; ++k) {
            const int i = pairs[2 * k], j = pairs[2 * k + 1];
            double dr[3];
            rsq_of(s, i, j, dr);
            const V3<double> d = V3<double>::from(dr);
            for (int term = 0; term < 6; ++term) {
                if (!((mask >> term) & 1u) || !s->pair_params[term]) continue;
                PairOut<double> o;
                if (!eval_term(term, s, i, j, d, o)) continue;
                et[kSlot[term]] += o.E;
                if (term == OXDNA_TERM_HBONDCROSS) et[OXDNA_E_CRSTACK] += o.E2;
                f[3 * i + 0] += o.F.x; f[3 * i + 1] += o.F.y; f[3 * i + 2] += o.F.z;
                f[3 * j + 0] -= o.F.x; f[3 * j + 1] -= o.F.y; f[3 * j + 2] -= o.F.z;
                t[3 * i + 0] += o.Ti.x; t[3 * i + 1] += o.Ti.y; t[3 * i + 2] += o.Ti.z;
                t[3 * j + 0] += o.Tj.x; t[3 * j + 1] += o.Tj.y; t[3 * j + 2] += o.Tj.z;
            }
        }
        if (bonded) {
#pragma omp for schedule(static) nowait
            for (long long b = 0; b < (long long)s->nb; ++b) {
                const int A = s->bonds[2 * b], B = s->bonds[2 * b + 1];
                BondOut<double> r;
                bond_eval(s, (uint32_t)b, r);
                et[OXDNA_E_FENE] += r.e_fene; et[OXDNA_E_BEXC] += r.e_bexc; et[OXDNA_E_STACK] += r.e_stack;
                f[3 * A + 0] -= r.FB.x; f[3 * A + 1] -= r.FB.y; f[3 * A + 2] -= r.FB.z;
                f[3 * B + 0] += r.FB.x; f[3 * B + 1] += r.FB.y; f[3 * B + 2] += r.FB.z;
                t[3 * A + 0] += r.TA.x; t[3 * A + 1] += r.TA.y; t[3 * A + 2] += r.TA.z;
                t[3 * B + 0] += r.TB.x; t[3 * B + 1] += r.TB.y; t[3 * B + 2] += r.TB.z;
            }
        }
    }
    for (int w = 1; w < nt; ++w) {
        for (size_t k = 0; k < 3 * (size_t)n; ++k) { fa[0][k] += fa[w][k]; ta[0][k] += ta[w][k]; }
        for (int k = 0; k < 8; ++k) ea[0][k] += ea[w][k];
    }
    if (F) std::memcpy(F, fa[0].data(), sizeof(double) * 3 * n);
    if (T) std::memcpy(T, ta[0].data(), sizeof(double) * 3 * n);
    if (Eterm) std::memcpy(Eterm, ea[0].data(), sizeof(double) * 8);
    return 0;
}

// Neighbour pairs i < j (bond-excluded) with |dr|^2 < r_list^2; brute force or cells.
int oracle_pairs(const oracle_system* s, double r_list, int use_cells, int32_t* out, size_t max_pairs,
                 size_t* count) {
    PairList pl;
    if (use_cells) pairs_cells(s, r_list, pl);
    else pairs_brute(s, r_list, pl);
    std::sort(pl.pairs.begin(), pl.pairs.end());
    *count = pl.pairs.size();
    if (pl.pairs.size() > max_pairs) return -1;
    for (size_t k = 0; k < pl.pairs.size(); ++k) {
        out[2 * k] = pl.pairs[k].first;
        out[2 * k + 1] = pl.pairs[k].second;
    }
    return 0;
}

// One nonbonded pair: dr = r_i - r_j (already minimum-imaged), returns 1 when the
// pair is inside rcut. out: F[3], Ti[3], Tj[3], E, E2.
int oracle_pair(int term, const double* dr, const double* qi, const double* qj, double rcut,
                const void* params, double chg_i, double chg_j, double* out) {
    V3<double> d = V3<double>::from(dr);
    Quat<double> a(qi), b(qj);
    PairOut<double> o;
    double rcsq = rcut * rcut;
    bool in = false;
    switch (term) {
    case OXDNA_TERM_EXCVOL: in = eval_excvol(d, a, b, rcsq, *(const oxdna_excvol_params*)params, o); break;
    case OXDNA_TERM_HBOND: in = eval_hbond(d, a, b, rcsq, *(const oxdna_hbond_params*)params, o); break;
    case OXDNA_TERM_CROSS: in = eval_cross(d, a, b, rcsq, *(const oxdna_crossstacking_params*)params, o); break;
    case OXDNA_TERM_COAX: in = eval_coax(d, a, b, rcsq, *(const oxdna_coaxstacking_params*)params, o); break;
    case OXDNA_TERM_DEBYE: in = eval_debye(d, a, b, rcsq, *(const oxdna_debye_params*)params, chg_i, chg_j, o); break;
    case OXDNA_TERM_HBONDCROSS: in = eval_hbondcross(d, a, b, rcsq, *(const oxdna_hbondcross_params*)params, o); break;
    default: return -1;
    }
    double v[11] = {o.F.x, o.F.y, o.F.z, o.Ti.x, o.Ti.y, o.Ti.z, o.Tj.x, o.Tj.y, o.Tj.z, o.E, o.E2};
    std::memcpy(out, v, sizeof(v));
    return in ? 1 : 0;
}

// One bond (A, B), Rc = r_B - r_A. out: FB[3], TA[3], TB[3], e_fene, e_bexc, e_stack, vir[6].
int oracle_bond(const oxdna_bonded_params* p, const double* Rc, const double* qA, const double* qB,
                int tA, int tB, double* out) {
    BondOut<double> r;
    bonded_pair(*p, V3<double>::from(Rc), Quat<double>(qA), Quat<double>(qB), tA, tB, r);
    double v[18] = {r.FB.x, r.FB.y, r.FB.z, r.TA.x, r.TA.y, r.TA.z, r.TB.x, r.TB.y, r.TB.z,
                    r.e_fene, r.e_bexc, r.e_stack, r.vir[0], r.vir[1], r.vir[2], r.vir[3], r.vir[4], r.vir[5]};
    std::memcpy(out, v, sizeof(v));
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Op census: the same routines instantiated on a counting scalar.
// Rule: add / sub / mul / div / sqrt = 1 flop each (a fused multiply-add is its mul
// plus its add = 2), acos / sin / exp / log = 1 special op each (reported apart);
// comparisons, negation and conversions are free.
// ---------------------------------------------------------------------------
namespace census {
struct Counts { double flops = 0, special = 0; };
static Counts g;
struct Real {
    double v;
    Real() : v(0) {}
    Real(double x) : v(x) {}
    Real(int x) : v(x) {}
    explicit operator double() const { return v; }
};
inline Real operator+(Real a, Real b) { g.flops += 1; return Real(a.v + b.v); }
inline Real operator-(Real a, Real b) { g.flops += 1; return Real(a.v - b.v); }
inline Real operator*(Real a, Real b) { g.flops += 1; return Real(a.v * b.v); }
inline Real operator/(Real a, Real b) { g.flops += 1; return Real(a.v / b.v); }
inline Real operator-(Real a) { return Real(-a.v); }
inline bool operator<(Real a, Real b) { return a.v < b.v; }
inline bool operator>(Real a, Real b) { return a.v > b.v; }
inline bool operator==(Real a, Real b) { return a.v == b.v; }
inline Real sqrt(Real a) { g.flops += 1; return Real(std::sqrt(a.v)); }
inline Real acos(Real a) { g.special += 1; return Real(std::acos(a.v)); }
inline Real sin(Real a) { g.special += 1; return Real(std::sin(a.v)); }
inline Real exp(Real a) { g.special += 1; return Real(std::exp(a.v)); }
inline Real log(Real a) { g.special += 1; return Real(std::log(a.v)); }
inline Real fabs(Real a) { return Real(std::fabs(a.v)); }
}  // namespace census

extern "C" {

// Census over the system's pairs for each term in mask, plus the bonded term.
// out[t*6 + k] for t = term id 0..6:
//   k=0 pairs inside the term's cut-off, k=1 of those with nonzero energy,
//   k=2 flops spent on zero-energy pairs, k=3 flops on nonzero pairs,
//   k=4 / k=5 special ops on zero / nonzero pairs.
// Counts are per unordered pair (bonds: per bond); a full-list kernel evaluates each twice.
int oracle_census(const oracle_system* s, uint32_t mask, double* out) {
    using census::Real;
    std::memset(out, 0, sizeof(double) * 7 * 6);
    double rl = max_rcut(s, mask);
    if (rl > 0) {
        PairList pl;
        pairs_cells(s, rl, pl);
        for (auto& pr : pl.pairs) {
            int i = pr.first, j = pr.second;
            double dr[3];
            rsq_of(s, i, j, dr);
            V3<Real> d = V3<Real>::from(dr);
            for (int term = 0; term < 6; ++term) {
                if (!((mask >> term) & 1u) || !s->pair_params[term]) continue;
                census::g = census::Counts();
                PairOut<Real> o;
                bool in = eval_term(term, s, i, j, d, o);
                if (!in) continue;
                bool nz = (o.E.v != 0.0) || (o.E2.v != 0.0);
                double* row = out + term * 6;
                row[0] += 1;
                row[1] += nz ? 1 : 0;
                row[nz ? 3 : 2] += census::g.flops;
                row[nz ? 5 : 4] += census::g.special;
            }
        }
    }
    if (((mask >> OXDNA_TERM_BONDED) & 1u) && s->bond_params) {
        for (uint32_t b = 0; b < s->nb; ++b) {
            census::g = census::Counts();
            BondOut<Real> r;
            bond_eval(s, b, r);
            bool nz = r.e_stack.v != 0.0;
            double* row = out + OXDNA_TERM_BONDED * 6;
            row[0] += 1;
            row[1] += nz ? 1 : 0;
            row[nz ? 3 : 2] += census::g.flops;
            row[nz ? 5 : 4] += census::g.special;
        }
    }
    return 0;
}

}  // extern "C"

extern "C" {

// Building blocks for the primitive checks. which: 0 f1 (p = a, r0, rlow, rhigh,
// rclow, rchigh, blow, bhigh, eps, shift), 1 f2 (p = k, r0, rc, rlow, rhigh, rclow,
// rchigh, blow, bhigh), 2 f4 of an angle (p = a, b, t0, ts, tc) -> (f4, f4Dsin),
// 3 f5 of cos phi (p = a, b, xc, xs), 4 repulsive LJ at distance x along one axis
// (p = eps, sigma, rstar, b, rc) -> (energy, force on the second site along the axis),
// 5 FENE (p = eps, r0, delta) at separation x -> (energy, force magnitude along r).
int oracle_prim(int which, double x, const double* p, double* out) {
    switch (which) {
    case 0: {
        F1Shape s{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
        out[0] = f1(x, s, p[8], p[9]);
        out[1] = f1D(x, s, p[8]);
        return 0;
    }
    case 1: {
        F2Shape s{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]};
        out[0] = f2(x, s);
        out[1] = f2D(x, s);
        return 0;
    }
    case 2:
        out[0] = f4(x, p);
        out[1] = f4Dsin(x, p);
        return 0;
    case 3:
        out[0] = f5(x, p);
        out[1] = f5D(x, p);
        return 0;
    case 4: {
        V3<double> f;
        out[0] = repulsive_lj(V3<double>(x, 0.0, 0.0), p[0], p[1], p[2], p[3], p[4], f);
        out[1] = f.x;
        return 0;
    }
    case 5: {
        double d = x - p[1];
        out[0] = fene_energy(d, p[0], p[2] * p[2]);
        out[1] = fene_force_prefactor(d, x, p[0], p[2] * p[2]) * x;
        return 0;
    }
    }
    return -1;
}

}  // extern "C"
