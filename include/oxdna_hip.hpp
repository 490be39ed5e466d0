// oxdna_hip.hpp — C++ host layer over the C ABI of oxdna_hip.h (header only, C++17).
//
// Mirrors the reference plugin's C++ classes so that host code written against them
// reads the same:
//   * OxDNABondedForceCompute (reference src/OxDNABondedForceCompute.h:38-77):
//     setParams / getParams per bond type, computeForces(timestep),
//     getFeneEnergy / getBondedExcVolEnergy / getStackingEnergy;
//   * AnisoPotentialPairOxDNA<Term, Params> — the six nonbonded force computes the
//     reference instantiates from HOOMD-blue's AnisoPotentialPair<Evaluator>
//     (src/AnisoPotentialPairOxDNA*.cc:14-20), with per-type-pair setParams / setRcut,
//     computeForces(timestep) and the evaluator names of getName()
//     (src/EvaluatorPairOxDNA*.h, e.g. EvaluatorPairOxDNAHBondCross.h:181).
// A SystemDefinition (host arrays: positions, orientations, types, charges, box,
// bonds) stands in for the host engine's; each force compute owns one engine context
// on the chosen device and uploads the state when it changed. Errors are thrown as
// oxdna::Error (std::runtime_error) carrying the engine's message, like the
// reference's host std::runtime_error.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "oxdna_hip.h"

namespace oxdna {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const oxdna_ctx* ctx, const char* what) {
    if (rc != 0) throw Error(std::string(what) + ": " + oxdna_last_error(ctx));
}

// Host-side system: n particles with (x, y, z), (w, x, y, z) orientations, type ids
// (0..3 = A, G, C, T), charges, a periodic box centred on the origin, and bonds
// (A, B, type) with B the 5' neighbour of A.
struct SystemDefinition {
    std::vector<double> pos, quat, charge;
    std::vector<int32_t> type;
    std::array<double, 3> box{{0.0, 0.0, 0.0}};
    std::vector<std::array<int32_t, 3>> bonds;
    std::vector<std::string> bond_types{"backbone"};
    uint64_t version = 0;  // bump after changing any array

    unsigned int getN() const { return (unsigned int)type.size(); }
    void modified() { ++version; }
};

// Per-particle results of the last computeForces.
struct ForceArrays {
    std::vector<double> force;   // 4 per particle: x, y, z, half energy
    std::vector<double> torque;  // 3 per particle
    std::vector<double> virial;  // 6 per particle: xx, xy, xz, yy, yz, zz
};

// RAII engine context bound to one system.
class Engine {
  public:
    explicit Engine(int device = 0) {
        oxdna_ctx* c = nullptr;
        check(oxdna_ctx_create(device, nullptr, &c), nullptr, "oxdna_ctx_create");
        m_ctx.reset(c);
    }
    oxdna_ctx* get() const { return m_ctx.get(); }

    void upload(const SystemDefinition& s) {
        if (m_version == s.version && m_have_state) return;
        const unsigned int n = s.getN();
        if (s.pos.size() != 3 * (size_t)n || s.quat.size() != 4 * (size_t)n)
            throw Error("SystemDefinition: pos / quat sizes do not match the type array");
        oxdna_box b{{s.box[0], s.box[1], s.box[2]}};
        check(oxdna_set_state(get(), n, s.pos.data(), s.quat.data(), s.type.data(),
                              s.charge.empty() ? nullptr : s.charge.data(), &b),
              get(), "oxdna_set_state");
        std::vector<int32_t> ab, bt;
        for (const auto& x : s.bonds) {
            ab.push_back(x[0]);
            ab.push_back(x[1]);
            bt.push_back(x[2]);
        }
        check(oxdna_set_bonds(get(), (unsigned int)s.bonds.size(), ab.data(), bt.data()), get(), "oxdna_set_bonds");
        m_version = s.version;
        m_have_state = true;
        m_n = n;
    }

    ForceArrays forces() const {
        const unsigned int n = m_n;
        ForceArrays f;
        f.force.resize(4 * (size_t)n);
        f.torque.resize(3 * (size_t)n);
        f.virial.resize(6 * (size_t)n);
        check(oxdna_get_forces(get(), f.force.data(), f.torque.data(), f.virial.data()), get(), "oxdna_get_forces");
        return f;
    }

    std::array<double, OXDNA_N_ENERGY> termEnergies() const {
        std::array<double, OXDNA_N_ENERGY> e{};
        check(oxdna_get_term_energies(get(), e.data()), get(), "oxdna_get_term_energies");
        return e;
    }

  private:
    struct Deleter {
        void operator()(oxdna_ctx* c) const { oxdna_ctx_destroy(c); }
    };
    std::unique_ptr<oxdna_ctx, Deleter> m_ctx;
    uint64_t m_version = 0;
    bool m_have_state = false;
    unsigned int m_n = 0;
};

// Common part of a force compute: the system, its engine context, the last results.
class ForceCompute {
  public:
    ForceCompute(std::shared_ptr<SystemDefinition> sysdef, int device)
        : m_sysdef(std::move(sysdef)), m_engine(device) {
        if (!m_sysdef) throw Error("ForceCompute: NULL system definition");
    }
    virtual ~ForceCompute() = default;

    // Evaluates the force at the system's current state (the timestep is carried for
    // interface parity with the host engine; the engine is stateless across calls).
    void compute(uint64_t timestep) { computeForces(timestep); }
    const ForceArrays& getForces() const { return m_out; }
    // Total energy of this force (sum of its per-term slots).
    double getEnergy() const { return m_energy; }

  protected:
    virtual void computeForces(uint64_t timestep) = 0;
    void prepare() { m_engine.upload(*m_sysdef); }
    void evaluate(unsigned int term_bit) {
        check(oxdna_compute(m_engine.get(), term_bit, 1), m_engine.get(), "oxdna_compute");
        m_e = m_engine.termEnergies();
        m_out = m_engine.forces();
    }

    std::shared_ptr<SystemDefinition> m_sysdef;
    Engine m_engine;
    ForceArrays m_out;
    std::array<double, OXDNA_N_ENERGY> m_e{};
    double m_energy = 0.0;
};

// FENE + bonded excluded volume + stacking (reference src/OxDNABondedForceCompute.h:38-77).
class OxDNABondedForceCompute : public ForceCompute {
  public:
    explicit OxDNABondedForceCompute(std::shared_ptr<SystemDefinition> sysdef, int device = 0)
        : ForceCompute(std::move(sysdef), device) {}

    // Parameters of one bond type (the reference takes a Python dict with the same keys).
    void setParams(const std::string& type, const oxdna_bonded_params& p) {
        const auto& names = m_sysdef->bond_types;
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == type) {
                m_params[k] = p;
                m_dirty = true;
                return;
            }
        throw Error("OxDNABondedForceCompute: unknown bond type " + type);
    }
    oxdna_bonded_params getParams(const std::string& type) const {
        const auto& names = m_sysdef->bond_types;
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == type) {
                auto it = m_params.find(k);
                if (it == m_params.end()) throw Error("OxDNABondedForceCompute: parameters of " + type + " not set");
                return it->second;
            }
        throw Error("OxDNABondedForceCompute: unknown bond type " + type);
    }

    double getFeneEnergy() const { return m_e[OXDNA_E_FENE]; }
    double getBondedExcVolEnergy() const { return m_e[OXDNA_E_BEXC]; }
    double getStackingEnergy() const { return m_e[OXDNA_E_STACK]; }

  protected:
    void computeForces(uint64_t) override {
        prepare();
        if (m_dirty) {
            const size_t nt = m_sysdef->bond_types.size();
            std::vector<oxdna_bonded_params> table(nt);
            for (size_t k = 0; k < nt; ++k) {
                auto it = m_params.find(k);
                if (it == m_params.end())
                    throw Error("OxDNABondedForceCompute: parameters of " + m_sysdef->bond_types[k] + " not set");
                table[k] = it->second;
            }
            check(oxdna_set_bond_params(m_engine.get(), (unsigned int)nt, table.data()), m_engine.get(),
                  "oxdna_set_bond_params");
            m_dirty = false;
        }
        evaluate(1u << OXDNA_TERM_BONDED);
        m_energy = m_e[OXDNA_E_FENE] + m_e[OXDNA_E_BEXC] + m_e[OXDNA_E_STACK];
    }

  private:
    std::map<size_t, oxdna_bonded_params> m_params;
    bool m_dirty = true;
};

// One nonbonded oxDNA term: HOOMD's AnisoPotentialPair<Evaluator> as the reference
// instantiates it, per type pair parameters and cut-offs, a full neighbour list with
// skin `r_buff` and bonded pairs excluded.
template<int TERM, class Params>
class AnisoPotentialPairOxDNA : public ForceCompute {
  public:
    using param_type = Params;

    AnisoPotentialPairOxDNA(std::shared_ptr<SystemDefinition> sysdef, double r_buff = 0.4, int device = 0)
        : ForceCompute(std::move(sysdef), device), m_rbuff(r_buff) {}

    static std::string getName() {
        switch (TERM) {
        case OXDNA_TERM_EXCVOL: return "oxdna_excvol";
        case OXDNA_TERM_HBOND: return "oxdna_hbond";
        case OXDNA_TERM_CROSS: return "oxdna_crossstacking";
        case OXDNA_TERM_HBONDCROSS: return "oxdna_hbondcross";
        case OXDNA_TERM_COAX: return "oxdna_coaxstacking";
        default: return "oxdna_debye";
        }
    }

    // Parameters of the unordered type pair (typ1, typ2); type ids 0..3.
    void setParams(unsigned int typ1, unsigned int typ2, const Params& p) {
        checkTypes(typ1, typ2);
        m_params[typ1 * 4 + typ2] = p;
        m_params[typ2 * 4 + typ1] = p;
        m_set[typ1 * 4 + typ2] = m_set[typ2 * 4 + typ1] = true;
        m_dirty = true;
    }
    const Params& getParams(unsigned int typ1, unsigned int typ2) const {
        checkTypes(typ1, typ2);
        return m_params[typ1 * 4 + typ2];
    }
    // Centre-of-mass cut-off of the pair (typ1, typ2).
    void setRcut(unsigned int typ1, unsigned int typ2, double rcut) {
        checkTypes(typ1, typ2);
        if (!(rcut >= 0.0)) throw Error(getName() + ": negative cut-off");
        m_rcut[typ1 * 4 + typ2] = m_rcut[typ2 * 4 + typ1] = rcut;
        m_dirty = true;
    }
    double getRcut(unsigned int typ1, unsigned int typ2) const {
        checkTypes(typ1, typ2);
        return m_rcut[typ1 * 4 + typ2];
    }

  protected:
    void computeForces(uint64_t) override {
        prepare();
        if (m_dirty) {
            for (int k = 0; k < 16; ++k)
                if (!m_set[k]) throw Error(getName() + ": parameters of every type pair must be set");
            check(oxdna_set_pair_params(m_engine.get(), TERM, 4, m_params.data(), m_rcut.data()), m_engine.get(),
                  "oxdna_set_pair_params");
            check(oxdna_set_nlist(m_engine.get(), m_rbuff), m_engine.get(), "oxdna_set_nlist");
            m_dirty = false;
        }
        evaluate(1u << TERM);
        m_energy = 0.0;
        for (double e : m_e) m_energy += e;
    }

  private:
    void checkTypes(unsigned int a, unsigned int b) const {
        if (a >= 4 || b >= 4) throw Error(getName() + ": type id out of range");
    }
    std::array<Params, 16> m_params{};
    std::array<bool, 16> m_set{};
    std::array<double, 16> m_rcut{};
    double m_rbuff;
    bool m_dirty = true;
};

using AnisoPotentialPairOxDNAExcVol = AnisoPotentialPairOxDNA<OXDNA_TERM_EXCVOL, oxdna_excvol_params>;
using AnisoPotentialPairOxDNAHBond = AnisoPotentialPairOxDNA<OXDNA_TERM_HBOND, oxdna_hbond_params>;
using AnisoPotentialPairOxDNACrossStacking = AnisoPotentialPairOxDNA<OXDNA_TERM_CROSS, oxdna_crossstacking_params>;
using AnisoPotentialPairOxDNAHBondCross = AnisoPotentialPairOxDNA<OXDNA_TERM_HBONDCROSS, oxdna_hbondcross_params>;
using AnisoPotentialPairOxDNACoaxStacking = AnisoPotentialPairOxDNA<OXDNA_TERM_COAX, oxdna_coaxstacking_params>;
using AnisoPotentialPairOxDNADebye = AnisoPotentialPairOxDNA<OXDNA_TERM_DEBYE, oxdna_debye_params>;

}  // namespace oxdna
