// host_api_check.cpp — drives the C++ host layer (include/oxdna_hip.hpp) from a system
// file and writes every force compute's energies, forces and torques. Used by the GPU
// tests (tests/test_gpu_host_api.py), which compare its output with the oracle.
//
// Input (native-endian binary): uint32 n, nb; double pos[3n], quat[4n]; int32 type[n];
// double charge[n], box[3]; int32 bonds[3 nb] (A, B, type); then, per nonbonded term
// 0..5, int32 present and, if present, 16 parameter structs + double rcut[16]; then
// int32 bonded_present and, if present, one oxdna_bonded_params.
// Output: for each present force in the order read, int32 term, double energies[8],
// double force[4n], double torque[3n].
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "oxdna_hip.hpp"

namespace {

struct Reader {
    FILE* f;
    template<class T> void get(T* p, size_t n) {
        if (n && std::fread(p, sizeof(T), n, f) != n) throw oxdna::Error("truncated input");
    }
    template<class T> T one() {
        T v;
        get(&v, 1);
        return v;
    }
};

template<class Pair>
std::unique_ptr<oxdna::ForceCompute> make_pair(Reader& in, std::shared_ptr<oxdna::SystemDefinition> sys) {
    typename Pair::param_type table[16];
    double rcut[16];
    in.get(table, 16);
    in.get(rcut, 16);
    auto f = std::make_unique<Pair>(sys, 0.4);
    for (unsigned a = 0; a < 4; ++a)
        for (unsigned b = a; b < 4; ++b) {
            f->setParams(a, b, table[a * 4 + b]);
            f->setRcut(a, b, rcut[a * 4 + b]);
        }
    return f;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s system.bin out.bin\n", argv[0]);
        return 2;
    }
    try {
        FILE* fi = std::fopen(argv[1], "rb");
        if (!fi) throw oxdna::Error(std::string("cannot open ") + argv[1]);
        Reader in{fi};
        auto sys = std::make_shared<oxdna::SystemDefinition>();
        const uint32_t n = in.one<uint32_t>(), nb = in.one<uint32_t>();
        sys->pos.resize(3 * (size_t)n);
        sys->quat.resize(4 * (size_t)n);
        sys->type.resize(n);
        sys->charge.resize(n);
        in.get(sys->pos.data(), sys->pos.size());
        in.get(sys->quat.data(), sys->quat.size());
        in.get(sys->type.data(), n);
        in.get(sys->charge.data(), n);
        in.get(sys->box.data(), 3);
        sys->bonds.resize(nb);
        for (auto& b : sys->bonds) in.get(b.data(), 3);
        sys->modified();

        std::vector<std::pair<int, std::unique_ptr<oxdna::ForceCompute>>> forces;
        for (int t = 0; t < OXDNA_TERM_BONDED; ++t) {
            if (!in.one<int32_t>()) continue;
            switch (t) {
            case OXDNA_TERM_EXCVOL: forces.emplace_back(t, make_pair<oxdna::AnisoPotentialPairOxDNAExcVol>(in, sys)); break;
            case OXDNA_TERM_HBOND: forces.emplace_back(t, make_pair<oxdna::AnisoPotentialPairOxDNAHBond>(in, sys)); break;
            case OXDNA_TERM_CROSS: forces.emplace_back(t, make_pair<oxdna::AnisoPotentialPairOxDNACrossStacking>(in, sys)); break;
            case OXDNA_TERM_COAX: forces.emplace_back(t, make_pair<oxdna::AnisoPotentialPairOxDNACoaxStacking>(in, sys)); break;
            case OXDNA_TERM_DEBYE: forces.emplace_back(t, make_pair<oxdna::AnisoPotentialPairOxDNADebye>(in, sys)); break;
            case OXDNA_TERM_HBONDCROSS: forces.emplace_back(t, make_pair<oxdna::AnisoPotentialPairOxDNAHBondCross>(in, sys)); break;
            }
        }
        if (in.one<int32_t>()) {
            oxdna_bonded_params p;
            in.get(&p, 1);
            auto b = std::make_unique<oxdna::OxDNABondedForceCompute>(sys);
            b->setParams("backbone", p);
            forces.emplace_back(OXDNA_TERM_BONDED, std::move(b));
        }
        std::fclose(fi);

        FILE* fo = std::fopen(argv[2], "wb");
        if (!fo) throw oxdna::Error(std::string("cannot open ") + argv[2]);
        for (auto& [t, f] : forces) {
            f->compute(0);
            std::array<double, OXDNA_N_ENERGY> e{};
            if (t == OXDNA_TERM_BONDED) {
                auto* b = static_cast<oxdna::OxDNABondedForceCompute*>(f.get());
                e[OXDNA_E_FENE] = b->getFeneEnergy();
                e[OXDNA_E_BEXC] = b->getBondedExcVolEnergy();
                e[OXDNA_E_STACK] = b->getStackingEnergy();
            } else {
                e[0] = f->getEnergy();  // total of the term (its slots are checked from Python)
            }
            const int32_t term = t;
            std::fwrite(&term, sizeof(term), 1, fo);
            std::fwrite(e.data(), sizeof(double), e.size(), fo);
            const auto& out = f->getForces();
            std::fwrite(out.force.data(), sizeof(double), out.force.size(), fo);
            std::fwrite(out.torque.data(), sizeof(double), out.torque.size(), fo);
            std::printf("%s: energy %.9g\n", t == OXDNA_TERM_BONDED ? "bonded" : std::to_string(t).c_str(),
                        f->getEnergy());
        }
        std::fclose(fo);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
