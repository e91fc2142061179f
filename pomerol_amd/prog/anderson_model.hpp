// Full ED of the single-impurity Anderson model (reference: prog/anderson_model.hpp:27-118):
// impurity "A" with Coulomb U and level ed, bath sites b0, b1, ... with the given levels and
// hoppings to A.
#pragma once

#include "quantum_model.hpp"

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

class anderson_model : public quantum_model {
    std::vector<double> levels, hoppings;

public:
    anderson_model(int argc, char* argv[]) : quantum_model(argc, argv, "Full-ED of the Anderson model") {
        args_parser.value("U", "Interaction constant U", "10");
        args_parser.value("ed", "Energy level of the impurity", "0");
        args_parser.value("levels", "Energy levels of the bath sites", "");
        args_parser.value("hoppings", "Hopping to the bath sites", "");
        parse_args(argc, argv);
        levels = args_parser.get_list<double>("levels");
        hoppings = args_parser.get_list<double>("hoppings");
        if(levels.size() != hoppings.size()) {
            std::cerr << "Number of levels != number of hoppings\n";
            std::exit(2);
        }
        std::cout << "Diagonalization of 1+" << levels.size() << " sites\n";
        using namespace Pomerol::LatticePresets;
        HExpr += CoulombS("A", args_parser.get_double("U"), args_parser.get_double("ed"));
        for(std::size_t i = 0; i < levels.size(); ++i) {
            std::string const name = "b" + std::to_string(i);
            HExpr += Hopping("A", name, hoppings[i]);
            HExpr += Level(name, levels[i]);
        }
        std::cout << "Hamiltonian:\n" << HExpr << '\n';
    }

    std::pair<Pomerol::ParticleIndex, Pomerol::ParticleIndex> get_node(IndexInfoType const& IndexInfo) override {
        using Pomerol::LatticePresets::down;
        using Pomerol::LatticePresets::up;
        return {IndexInfo.getIndex("A", 0, down), IndexInfo.getIndex("A", 0, up)};
    }

    // G_{down down} only
    void prepare_indices(Pomerol::ParticleIndex d0, Pomerol::ParticleIndex, std::set<Pomerol::IndexCombination2>& indices2,
                         std::set<Pomerol::ParticleIndex>&, IndexInfoType const&) override {
        indices2.insert(Pomerol::IndexCombination2(d0, d0));
    }
};
