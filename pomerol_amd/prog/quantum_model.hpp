// Command-line ED programs over the pomerol-compatible API (reference: prog/quantum_model.hpp,
// prog/quantum_model.cpp): build a model Hamiltonian, diagonalize it, and optionally write the
// Green's functions on Matsubara and real frequencies and one 2PGF component on a
// (bosonic, fermionic, fermionic) Matsubara grid as text files in the working directory.
#pragma once

#include <pomerol.hpp>

#include "args.hpp"

#include <set>
#include <string>
#include <utility>

class quantum_model {
public:
    using IndexInfoType =
        Pomerol::IndexClassification<std::string, unsigned short, Pomerol::LatticePresets::spin>;

    quantum_model(int argc, char* argv[], std::string const& prog_desc);
    virtual ~quantum_model() = default;

    /// Diagonalize the model; Green's functions and the 2PGF as requested on the command line.
    void compute();

    /// The (spin down, spin up) indices of the node the Green's functions are taken on.
    virtual std::pair<Pomerol::ParticleIndex, Pomerol::ParticleIndex> get_node(IndexInfoType const& IndexInfo) = 0;
    /// Index pairs of the Green's functions to write, and the indices whose c, c^+ they need.
    virtual void prepare_indices(Pomerol::ParticleIndex d0, Pomerol::ParticleIndex u0,
                                 std::set<Pomerol::IndexCombination2>& indices2, std::set<Pomerol::ParticleIndex>& f,
                                 IndexInfoType const& IndexInfo) = 0;

protected:
    progargs::Parser args_parser;
    Pomerol::LatticePresets::RealExpr HExpr;

    void parse_args(int argc, char const* const argv[]);
    void print_section(std::string const& str) const;

private:
    Pomerol::RealType beta = 1.0;
    bool calc_gf = false;
    bool calc_2pgf = false;
};
