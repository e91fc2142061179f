// A pomerol caller written against the reference's header names: compiled and linked against
// this library by tests/test_library.py (no device is needed to build it).
#include <pomerol/DensityMatrix.hpp>
#include <pomerol/EnsembleAverage.hpp>
#include <pomerol/FieldOperatorContainer.hpp>
#include <pomerol/GFContainer.hpp>
#include <pomerol/GreensFunction.hpp>
#include <pomerol/Hamiltonian.hpp>
#include <pomerol/HilbertSpace.hpp>
#include <pomerol/IndexClassification.hpp>
#include <pomerol/LatticePresets.hpp>
#include <pomerol/Misc.hpp>
#include <pomerol/StatesClassification.hpp>
#include <pomerol/Susceptibility.hpp>
#include <pomerol/ThreePointSusceptibility.hpp>
#include <pomerol/ThreePointSusceptibilityContainer.hpp>
#include <pomerol/TwoParticleGFContainer.hpp>
#include <pomerol/Vertex4.hpp>

#include <complex>
#include <iostream>
#include <set>

using namespace Pomerol;

int main() {
    using namespace LatticePresets;
    RealType const U = 1.0, mu = 0.4, beta = 10.0;
    auto HExpr = CoulombS("A", U, -mu) + Level("B", 0.3) + Hopping("A", "B", 0.5);
    auto IndexInfo = MakeIndexClassification(HExpr);
    auto HS = MakeHilbertSpace(IndexInfo, HExpr);
    HS.compute();
    StatesClassification S;
    S.compute(HS);
    Hamiltonian H(S);
    H.prepare(HExpr, HS, MPI_COMM_WORLD);
    H.compute(MPI_COMM_WORLD);
    DensityMatrix rho(S, H, beta);
    rho.prepare();
    rho.compute();
    rho.truncateBlocks(1e-12, false);

    FieldOperatorContainer Operators(IndexInfo, HS, S, H);
    Operators.prepareAll(HS);
    Operators.computeAll();
    ParticleIndex iu = IndexInfo.getIndex("A", 0, up), id = IndexInfo.getIndex("A", 0, down);

    GreensFunction GF(S, H, Operators.getAnnihilationOperator(id), Operators.getCreationOperator(id), rho);
    GF.prepare();
    GF.compute();
    std::complex<double> g0 = GF(0), gt = GF.of_tau(0.5 * beta);

    GFContainer G(IndexInfo, S, H, rho, Operators);
    std::set<IndexCombination2> pairs = {IndexCombination2(iu, iu), IndexCombination2(id, id)};
    G.prepareAll(pairs);
    G.computeAll();

    QuadraticOperator SpOp(IndexInfo, HS, S, H, iu, id), SmOp(IndexInfo, HS, S, H, id, iu);
    SpOp.prepare(HS);
    SpOp.compute();
    SmOp.prepare(HS);
    SmOp.compute();
    Susceptibility Chi(S, H, SpOp, SmOp, rho);
    Chi.prepare();
    Chi.compute();
    Chi.subtractDisconnected();
    std::complex<double> chi0 = Chi(0), chit = Chi.of_tau(0.0);
    EnsembleAverage Ave(SpOp, rho);
    Ave.compute();

    TwoParticleGFContainer Chi4(IndexInfo, S, H, rho, Operators);
    Chi4.PoleResolution = 1e-8;
    Chi4.CoefficientTolerance = 1e-16;
    std::set<IndexCombination4> quads = {IndexCombination4(iu, id, iu, id)};
    Chi4.prepareAll(quads);
    Chi4.computeAll(false, FreqVec3{}, MPI_COMM_WORLD);
    std::complex<double> x = Chi4(IndexCombination4(iu, id, iu, id))(0, 0, 0);
    GreensFunction const& Gu = G(iu, iu);
    GreensFunction const& Gd = G(id, id);
    Vertex4 Gamma(Chi4(IndexCombination4(iu, id, iu, id)), Gu, Gd, G(iu, id), G(id, iu));
    std::set<IndexCombination2> mixed = {IndexCombination2(iu, id), IndexCombination2(id, iu)};
    G.prepareAll(mixed);
    G.computeAll();
    Gamma.compute(2);
    x += Gamma(0, 0, 0) - Gamma.value(0, 0, 0);

    ThreePointSusceptibility Chi3(Channel::PH, S, H, Operators.getCreationOperator(iu),
                                  Operators.getAnnihilationOperator(iu), Operators.getCreationOperator(id),
                                  Operators.getAnnihilationOperator(id), rho);
    Chi3.prepare();
    Chi3.compute();
    x += Chi3(0, 0);
    ThreePointSusceptibilityContainer Chi3s(Channel::xPH, IndexInfo, S, H, rho, Operators);
    Chi3s.prepareAll({IndexCombination4(iu, id, id, iu)});
    RealType const w0 = M_PI / beta;
    auto chi3vals = Chi3s.computeAll(false, FreqVec2{FreqTuple2(ComplexType(0, w0), ComplexType(0, -w0))});
    x += chi3vals.begin()->second.at(0) - Chi3s(IndexCombination4(iu, id, id, iu)).element()->operator()(0, -1);

    std::cout << g0 << gt << G(iu, iu)(3) << chi0 << chit << Ave() << x << std::endl;
    return 0;
}
