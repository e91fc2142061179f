// The common driver of the ED programs (reference: prog/quantum_model.cpp:27-263).
//
// Output files (one text line per grid point, values written with 16 significant digits):
//   spectrum.dat                 "k  E_k" for the sorted eigenvalues
//   gw_imfreq_<ij>.dat           "w_n  Re G  Im G",  n = wf_min .. 4 wf_max
//   gw_refreq_<ij>.dat           "w  Re G  Im G",    w = -D .. D in steps of gf.step, G(w + i eta)
//   chi<ijkl>.dat                "W  w3  w2  Re  Im" for the 2PGF at (W + w3, w2, w3)
//   chi<ijkl>_W<Im W>.dat        "w3  w2  Re  Im" for one bosonic frequency W
// with Matsubara frequencies written as their imaginary parts; the fermionic grids run over
// n = wf_min .. wf_max, the bosonic one over m = wb_min .. wb_max, both ends included.
#include "quantum_model.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <tuple>
#include <vector>

using namespace Pomerol;

namespace {

std::ofstream open_out(std::string const& name) {
    std::ofstream f(name);
    if(!f)
        throw std::runtime_error("cannot write " + name);
    f << std::setprecision(16) << std::scientific;
    return f;
}

} // namespace

quantum_model::quantum_model(int, char*[], std::string const& prog_desc) : args_parser(prog_desc) {
    args_parser.value("beta", "Inverse temperature", "1");
    args_parser.flag("calc_gf", "Calculate Green's functions");
    args_parser.flag("calc_2gf", "Calculate 2-particle Green's functions");
    args_parser.value("gf.eta", "GF: Offset from the real axis for Green's function calculation", "0.05");
    args_parser.value("gf.step", "GF: step of the real frequency grid", "0.01");
    args_parser.value("gf.D", "GF: length of the real frequency grid", "6");
    args_parser.value("wf_min", "Minimum fermionic Matsubara frequency", "-20");
    args_parser.value("wf_max", "Maximum fermionic Matsubara frequency (4x for GF)", "20");
    args_parser.value("wb_min", "Minimum bosonic Matsubara frequency", "0");
    args_parser.value("wb_max", "Maximum bosonic Matsubara frequency", "0");
    args_parser.value("2pgf.indices", "2PGF index combination", "0,0,0,0");
    args_parser.value("2pgf.reduce_tol", "Energy resonance resolution in 2PGF", "1e-05");
    args_parser.value("2pgf.coeff_tol", "Tolerance on numerators in 2PGF", "1e-12");
}

void quantum_model::parse_args(int argc, char const* const argv[]) {
    args_parser.parse(argc, argv);
    beta = args_parser.get_double("beta");
    calc_gf = args_parser.get_flag("calc_gf");
    calc_2pgf = args_parser.get_flag("calc_2gf");
    calc_gf = calc_gf || calc_2pgf;
}

void quantum_model::print_section(std::string const& str) const {
    std::cout << std::string(str.size(), '=') << '\n' << str << '\n' << std::string(str.size(), '=') << '\n';
}

void quantum_model::compute() {
    IndexInfoType IndexInfo = MakeIndexClassification(HExpr);
    print_section("Indices");
    std::cout << IndexInfo << '\n';

    auto HS = MakeHilbertSpace(IndexInfo, HExpr);
    HS.compute();
    StatesClassification S;
    S.compute(HS);
    Hamiltonian H(S);
    H.prepare(HExpr, HS, MPI_COMM_WORLD);
    H.compute(MPI_COMM_WORLD);

    {
        RealVectorType evals = H.getEigenValues();
        std::sort(evals.begin(), evals.end());
        auto f = open_out("spectrum.dat");
        for(std::size_t k = 0; k < evals.size(); ++k)
            f << k << "  " << evals[k] << '\n';
    }

    DensityMatrix rho(S, H, beta);
    rho.prepare();
    rho.compute();

    auto node = get_node(IndexInfo);
    ParticleIndex d0 = node.first, u0 = node.second;
    if(!calc_gf)
        return;

    print_section("1-particle Green's functions calc");
    std::set<ParticleIndex> f{u0, d0};
    std::set<IndexCombination2> indices2;
    double const eta = args_parser.get_double("gf.eta");
    double const step = args_parser.get_double("gf.step");
    double const hbw = args_parser.get_double("gf.D");
    int const wf_min = args_parser.get_int("wf_min"), wf_max = args_parser.get_int("wf_max");
    int const wb_min = args_parser.get_int("wb_min"), wb_max = args_parser.get_int("wb_max");
    prepare_indices(d0, u0, indices2, f, IndexInfo);

    FieldOperatorContainer Operators(IndexInfo, HS, S, H, f);
    Operators.prepareAll(HS);
    Operators.computeAll();
    GFContainer G(IndexInfo, S, H, rho, Operators);
    G.prepareAll(indices2);
    G.computeAll();

    auto fermionic = [this](int n) { return ComplexType(0.0, M_PI * (2 * n + 1) / beta); };
    auto bosonic = [this](int m) { return ComplexType(0.0, M_PI * 2 * m / beta); };

    for(auto const& ind2 : indices2) {
        GreensFunction const& GF = G(ind2);
        std::cout << "Saving imfreq G" << ind2 << " on " << 4 * wf_max << " Matsubara freqs.\n";
        std::string const ind_str = std::to_string(ind2.Index1) + std::to_string(ind2.Index2);
        std::vector<ComplexType> z;
        for(int n = wf_min; n <= 4 * wf_max; ++n)
            z.push_back(fermionic(n));
        std::vector<ComplexType> g = GF.evaluate(z);
        auto fi = open_out("gw_imfreq_" + ind_str + ".dat");
        for(std::size_t k = 0; k < z.size(); ++k)
            fi << z[k].imag() << "  " << g[k].real() << "  " << g[k].imag() << '\n';

        std::size_t const npts = 2 * static_cast<std::size_t>(hbw / step) + 1;
        std::vector<ComplexType> zr;
        for(std::size_t k = 0; k < npts; ++k) {
            double const w = npts > 1 ? -hbw + 2.0 * hbw * double(k) / double(npts - 1) : -hbw;
            zr.push_back(ComplexType(w, eta));
        }
        std::vector<ComplexType> gr = GF.evaluate(zr);
        auto fr = open_out("gw_refreq_" + ind_str + ".dat");
        for(std::size_t k = 0; k < zr.size(); ++k)
            fr << zr[k].real() << "  " << gr[k].real() << "  " << gr[k].imag() << '\n';
    }

    if(!calc_2pgf)
        return;

    print_section("2-Particle Green's function calc");
    std::vector<std::size_t> idx = args_parser.get_list<std::size_t>("2pgf.indices");
    if(idx.size() != 4)
        throw std::runtime_error("Need 4 indices for 2PGF");
    IndexCombination4 const index_comb{ParticleIndex(idx[0]), ParticleIndex(idx[1]), ParticleIndex(idx[2]),
                                       ParticleIndex(idx[3])};
    std::string const ind_str = std::to_string(index_comb.Index1) + std::to_string(index_comb.Index2) +
                                std::to_string(index_comb.Index3) + std::to_string(index_comb.Index4);
    TwoParticleGF G4(S, H, Operators.getAnnihilationOperator(index_comb.Index1),
                     Operators.getAnnihilationOperator(index_comb.Index2), Operators.getCreationOperator(index_comb.Index3),
                     Operators.getCreationOperator(index_comb.Index4), rho);
    G4.PoleResolution = args_parser.get_double("2pgf.reduce_tol");
    G4.CoefficientTolerance = args_parser.get_double("2pgf.coeff_tol");
    G4.prepare();

    FreqVec3 freqs;
    for(int m = wb_min; m <= wb_max; ++m)
        for(int n3 = wf_min; n3 <= wf_max; ++n3)
            for(int n2 = wf_min; n2 <= wf_max; ++n2)
                freqs.emplace_back(bosonic(m) + fermionic(n3), fermionic(n2), fermionic(n3));
    std::cout << "2PGF : " << freqs.size() << " freqs to evaluate\n";
    std::vector<ComplexType> chi = G4.compute(true, freqs);
    if(chi.empty()) // a vanishing 2PGF (no parts) comes back empty: it is zero
        chi.assign(freqs.size(), 0.0);

    std::cout << "Saving 2PGF " << index_comb << '\n';
    auto full = open_out("chi" + ind_str + ".dat");
    std::size_t w = 0;
    for(int m = wb_min; m <= wb_max; ++m) {
        double const W = bosonic(m).imag();
        auto one = open_out("chi" + ind_str + "_W" + std::to_string(W) + ".dat");
        for(int n3 = wf_min; n3 <= wf_max; ++n3)
            for(int n2 = wf_min; n2 <= wf_max; ++n2, ++w) {
                double const w3 = fermionic(n3).imag(), w2 = fermionic(n2).imag();
                full << W << "  " << w3 << "  " << w2 << "  " << chi[w].real() << "  " << chi[w].imag() << '\n';
                one << w3 << "  " << w2 << "  " << chi[w].real() << "  " << chi[w].imag() << '\n';
            }
    }
}
