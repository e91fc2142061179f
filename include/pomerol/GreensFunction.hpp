// Single-particle Green's function, dynamical susceptibility and their containers.
//
// References: GreensFunction.hpp/.cpp (parts :41-86, evaluation :123-147), GreensFunctionPart
// (.cpp:24-110, term merge rules .hpp:75-127), GFContainer.hpp/.cpp (:20-48), IndexContainer2.hpp,
// Susceptibility.hpp/.cpp (:42-119, :143-171), SusceptibilityPart.hpp/.cpp (:24-127, :216-224).
//
// compute() generates and merges the terms of every part in one device pipeline
// (pomx_pair_terms); operator() and evaluate() run the Lehmann sums on the device
// (pomx_pair_eval).  Terms stay on the device; getTerms() copies them to the host.
#pragma once

#include "MonomialOperator.hpp"

#include <map>
#include <memory>
#include <set>

namespace Pomerol {

/// One (pole, residue) term of a two-operator Lehmann representation.
struct PoleTerm {
    ComplexType Residue;
    RealType Pole;
};

class GreensFunctionPart : public Thermal {
    friend class GreensFunction;
    HamiltonianPart const& HpartInner;
    HamiltonianPart const& HpartOuter;
    DensityMatrixPart const& DMpartInner;
    DensityMatrixPart const& DMpartOuter;
    MonomialOperatorPart const& F1;
    MonomialOperatorPart const& F2;
    double const* d_res = nullptr;
    double const* d_pole = nullptr;
    int64_t n_terms = 0;

public:
    RealType PoleResolution;
    RealType CoefficientTolerance;
    GreensFunctionPart(MonomialOperatorPart const& F1, MonomialOperatorPart const& F2, HamiltonianPart const& HpartInner,
                       HamiltonianPart const& HpartOuter, DensityMatrixPart const& DMpartInner,
                       DensityMatrixPart const& DMpartOuter, RealType PoleResolution, RealType CoefficientTolerance);
    ComplexType operator()(ComplexType z) const;
    ComplexType operator()(long MatsubaraNumber) const { return (*this)(MatsubaraSpacing * RealType(2 * MatsubaraNumber + 1)); }
    /// G_part(tau), 0 <= tau <= beta (GreensFunctionPart.hpp:219-221).
    ComplexType of_tau(RealType tau) const;
    std::size_t getNumTerms() const { return std::size_t(n_terms); }
    std::vector<PoleTerm> getTerms() const;
    pomx_pair_part devicePart() const;
};

class GreensFunction : public Thermal, public ComputableObject {
    StatesClassification const& S;
    Hamiltonian const& H;
    FieldOperator const& F1;
    FieldOperator const& F2;
    DensityMatrix const& DM;
    bool Vanishing = true;
    std::vector<GreensFunctionPart> parts;
    std::shared_ptr<detail::DeviceArray<double>> d_res, d_pole;
    int64_t n_terms = 0;

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-8;

    GreensFunction(StatesClassification const& S, Hamiltonian const& H, FieldOperator const& F1, FieldOperator const& F2,
                   DensityMatrix const& DM);
    GreensFunction(GreensFunction const& GF);

    void prepare();
    void compute();
    ParticleIndex getIndex(std::size_t Position) const;
    ComplexType operator()(long MatsubaraNumber) const { return (*this)(MatsubaraSpacing * RealType(2 * MatsubaraNumber + 1)); }
    ComplexType operator()(ComplexType z) const;
    /// G(z) at many points in one device launch.
    std::vector<ComplexType> evaluate(std::vector<ComplexType> const& z) const;
    /// G(tau), 0 <= tau <= beta (GreensFunction.hpp:138-146).
    ComplexType of_tau(RealType tau) const;
    /// G(tau) at many points in one device launch.
    std::vector<ComplexType> evaluateTau(std::vector<RealType> const& tau) const;
    bool isVanishing() const { return Vanishing; }
    std::size_t getNumParts() const { return parts.size(); }
    std::size_t getNumTerms() const { return std::size_t(n_terms); }
    std::vector<GreensFunctionPart> const& getParts() const { return parts; }
};

struct IndexCombination2 {
    ParticleIndex Index1, Index2;
    IndexCombination2(ParticleIndex i1, ParticleIndex i2) : Index1(i1), Index2(i2) {}
    bool operator<(IndexCombination2 const& o) const { return std::tie(Index1, Index2) < std::tie(o.Index1, o.Index2); }
    bool operator==(IndexCombination2 const& o) const { return Index1 == o.Index1 && Index2 == o.Index2; }
    bool operator!=(IndexCombination2 const& o) const { return !(*this == o); }
    friend std::ostream& operator<<(std::ostream& os, IndexCombination2 const& ic) {
        return os << "(" << ic.Index1 << ic.Index2 << ")";
    }
};

/// G_ij for a set of index pairs (GFContainer.cpp:20-48; every pair if the set is empty).
class GFContainer : public Thermal {
    ParticleIndex NumIndices;
    StatesClassification const& S;
    Hamiltonian const& H;
    DensityMatrix const& DM;
    FieldOperatorContainer const& Operators;
    std::map<IndexCombination2, std::shared_ptr<GreensFunction>> ElementsMap;

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-8;

    /// Anomalous: elements <c_i c_j> (two annihilators) instead of <c_i c^+_j> (GFContainer.cpp:34-48).
    bool const Anomalous;

    template <typename... IndexTypes>
    GFContainer(IndexClassification<IndexTypes...> const& IndexInfo, StatesClassification const& S, Hamiltonian const& H,
                DensityMatrix const& DM, FieldOperatorContainer const& Ops, bool Anomalous = false)
        : Thermal(DM.beta), NumIndices(IndexInfo.getIndexSize()), S(S), H(H), DM(DM), Operators(Ops),
          Anomalous(Anomalous) {}

    void prepareAll(std::set<IndexCombination2> const& Indices = {});
    void computeAll();
    bool isInContainer(IndexCombination2 const& Indices) const { return ElementsMap.count(Indices) > 0; }
    GreensFunction& operator()(IndexCombination2 const& Indices);
    GreensFunction& operator()(ParticleIndex Index1, ParticleIndex Index2) { return (*this)(IndexCombination2(Index1, Index2)); }
};

class SusceptibilityPart : public Thermal {
    friend class Susceptibility;
    HamiltonianPart const& HpartInner;
    HamiltonianPart const& HpartOuter;
    DensityMatrixPart const& DMpartInner;
    DensityMatrixPart const& DMpartOuter;
    MonomialOperatorPart const& A;
    MonomialOperatorPart const& B;
    double const* d_res = nullptr;
    double const* d_pole = nullptr;
    int64_t n_terms = 0;
    ComplexType ZeroPoleWeight = 0;

public:
    RealType PoleResolution;
    RealType CoefficientTolerance;
    SusceptibilityPart(MonomialOperatorPart const& A, MonomialOperatorPart const& B, HamiltonianPart const& HpartInner,
                       HamiltonianPart const& HpartOuter, DensityMatrixPart const& DMpartInner,
                       DensityMatrixPart const& DMpartOuter, RealType PoleResolution, RealType CoefficientTolerance);
    ComplexType operator()(ComplexType z) const;
    ComplexType operator()(long MatsubaraNumber) const { return (*this)(MatsubaraSpacing * RealType(2 * MatsubaraNumber)); }
    ComplexType getZeroPoleWeight() const { return ZeroPoleWeight; }
    /// chi_part(tau), 0 <= tau <= beta, zero-pole weight included (SusceptibilityPart.hpp:222-224).
    ComplexType of_tau(RealType tau) const;
    std::size_t getNumTerms() const { return std::size_t(n_terms); }
    std::vector<PoleTerm> getTerms() const;
    pomx_pair_part devicePart() const;
};

class Susceptibility : public Thermal, public ComputableObject {
    StatesClassification const& S;
    Hamiltonian const& H;
    MonomialOperator const& A;
    MonomialOperator const& B;
    DensityMatrix const& DM;
    bool Vanishing = true;
    std::vector<SusceptibilityPart> parts;
    std::shared_ptr<detail::DeviceArray<double>> d_res, d_pole;
    int64_t n_terms = 0;
    ComplexType zero_pole_total = 0;
    bool SubtractDisconnected = false;
    ComplexType ave_A = {}, ave_B = {};

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-8;

    Susceptibility(StatesClassification const& S, Hamiltonian const& H, MonomialOperator const& A, MonomialOperator const& B,
                   DensityMatrix const& DM);
    Susceptibility(Susceptibility const& Chi);

    void prepare();
    void compute();
    void subtractDisconnected();
    void subtractDisconnected(ComplexType ave_A, ComplexType ave_B);
    void subtractDisconnected(EnsembleAverage& EA_A, EnsembleAverage& EA_B);
    ComplexType operator()(long MatsubaraNumber) const { return (*this)(MatsubaraSpacing * RealType(2 * MatsubaraNumber)); }
    ComplexType operator()(ComplexType z) const;
    std::vector<ComplexType> evaluate(std::vector<ComplexType> const& z) const;
    /// chi(tau), 0 <= tau <= beta (Susceptibility.hpp:160-170).
    ComplexType of_tau(RealType tau) const;
    /// chi(tau) at many points in one device launch.
    std::vector<ComplexType> evaluateTau(std::vector<RealType> const& tau) const;
    bool isVanishing() const { return Vanishing; }
    std::size_t getNumParts() const { return parts.size(); }
    std::vector<SusceptibilityPart> const& getParts() const { return parts; }
    std::size_t getNumTerms() const { return std::size_t(n_terms); }
};

} // namespace Pomerol
