// Full ED of a periodic size_x x size_y Hubbard cluster (reference: prog/hubbard2d_model.hpp:27-172):
// sites S0, S1, ... (S_{y size_x + x}), Coulomb U and chemical potential mu (default U/2) on every
// site, nearest-neighbour hopping -t to the right and upper neighbours, and t' on the two
// diagonals when |t'| > 1e-10.
#pragma once

#include "quantum_model.hpp"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

class hubbard2d_model : public quantum_model {
    int size_x = 2, size_y = 2;
    std::vector<std::string> names;

    std::size_t SiteIndexF(std::size_t x, std::size_t y) const { return y * size_x + x; }

public:
    hubbard2d_model(int argc, char* argv[]) : quantum_model(argc, argv, "Full-ED of the MxM Hubbard cluster") {
        args_parser.value("U", "Hubbard constant U", "10");
        args_parser.value("mu", "Chemical potential", "U/2");
        args_parser.value("t", "NN hopping constant t", "1");
        args_parser.value("tp", "NNN hopping constant t'", "0");
        args_parser.value("x", "Size over x", "2");
        args_parser.value("y", "Size over y", "2");
        parse_args(argc, argv);
        size_x = args_parser.get_int("x");
        size_y = args_parser.get_int("y");

        std::size_t const L = std::size_t(size_x * size_y);
        names.resize(L);
        for(std::size_t y = 0; y < std::size_t(size_y); ++y)
            for(std::size_t x = 0; x < std::size_t(size_x); ++x)
                names[SiteIndexF(x, y)] = "S" + std::to_string(SiteIndexF(x, y));

        using namespace Pomerol::LatticePresets;
        double const U = args_parser.get_double("U");
        double mu = std::numeric_limits<double>::quiet_NaN();
        try {
            mu = args_parser.get_double("mu");
        } catch(std::invalid_argument const&) { // the default "U/2"
        }
        if(std::isnan(mu))
            mu = U / 2;
        for(std::size_t i = 0; i < L; ++i)
            HExpr += CoulombS(names[i], U, -mu);

        double const t = args_parser.get_double("t"), tp = args_parser.get_double("tp");
        for(std::size_t y = 0; y < std::size_t(size_y); ++y) {
            for(std::size_t x = 0; x < std::size_t(size_x); ++x) {
                auto pos = SiteIndexF(x, y);
                auto pos_right = SiteIndexF((x + 1) % size_x, y);
                auto pos_up = SiteIndexF(x, (y + 1) % size_y);
                auto pos_dia_right = SiteIndexF((x + 1) % size_x, (y + 1) % size_y);
                // unsigned arithmetic as in the reference: (y - 1) wraps around before the modulo
                auto pos_dia_left = SiteIndexF((x + 1) % size_x, (y - 1) % size_y);
                std::cout << pos_dia_right << " " << pos_dia_left << '\n';
                if(size_x > 1)
                    HExpr += Hopping(std::min(names[pos], names[pos_right]), std::max(names[pos], names[pos_right]), -t);
                if(size_y > 1)
                    HExpr += Hopping(std::min(names[pos], names[pos_up]), std::max(names[pos], names[pos_up]), -t);
                if(std::abs(tp) > 1e-10 && size_x > 1 && size_y > 1) {
                    HExpr += Hopping(std::min(names[pos], names[pos_dia_right]), std::max(names[pos], names[pos_dia_right]),
                                     tp);
                    HExpr += Hopping(std::min(names[pos], names[pos_dia_left]), std::max(names[pos], names[pos_dia_left]),
                                     tp);
                }
            }
        }
        std::cout << "Hamiltonian:\n" << HExpr << '\n';
    }

    std::pair<Pomerol::ParticleIndex, Pomerol::ParticleIndex> get_node(IndexInfoType const& IndexInfo) override {
        using Pomerol::LatticePresets::down;
        using Pomerol::LatticePresets::up;
        return {IndexInfo.getIndex("S0", 0, down), IndexInfo.getIndex("S0", 0, up)};
    }

    // G between S0 and every site of the first row, spin down
    void prepare_indices(Pomerol::ParticleIndex d0, Pomerol::ParticleIndex, std::set<Pomerol::IndexCombination2>& indices2,
                         std::set<Pomerol::ParticleIndex>& f, IndexInfoType const& IndexInfo) override {
        for(std::size_t x = 0; x < std::size_t(size_x); ++x) {
            Pomerol::ParticleIndex ind = IndexInfo.getIndex(names[SiteIndexF(x, 0)], 0, Pomerol::LatticePresets::down);
            f.insert(ind);
            indices2.insert(Pomerol::IndexCombination2(d0, ind));
        }
    }
};
