// Three-point susceptibilities in the PP, PH and xPH channels, their parts and the index
// container.
//
// References: ThreePointSusceptibility.hpp/.cpp (operator roles :37-66, parts :68-142, compute
// :148-241, evaluation .hpp:162-177), ThreePointSusceptibilityPart.hpp/.cpp (term kinds
// .hpp:66-370, generation .cpp:243-348, evaluation .hpp:527-540, .cpp:350-359) and
// ThreePointSusceptibilityContainer.hpp/.cpp (:19-48), IndexContainer4.hpp (:171-273).
//
// Each part's terms are generated, merged and (for compute(clear, freqs)) evaluated on the
// frequency list on the device (pomx_chi3_part_compute).  With clear == false the merged terms
// stay on the device so operator() and evaluate() can use them later.
#pragma once

#include "TwoParticleGF.hpp"

#include <array>
#include <map>
#include <memory>
#include <set>
#include <tuple>

namespace Pomerol {

using FreqTuple2 = std::tuple<ComplexType, ComplexType>;
using FreqVec2 = std::vector<FreqTuple2>;

class ThreePointSusceptibilityPart : public Thermal, public ComputableObject {
    friend class ThreePointSusceptibility;

public:
    /// Fermion-fermion term C / ((z1 - P0)(z2 - P1)).
    struct NonResonantFFTerm {
        ComplexType Coeff = 0;
        std::array<RealType, 2> Poles = {{0, 0}};
        long Weight = 0;
        ComplexType operator()(ComplexType z1, ComplexType z2) const {
            return Coeff / ((z1 - Poles[0]) * (z2 - Poles[1]));
        }
    };
    /// Fermion-boson term C / ((z1 - P1)(z1 - xi z2 - P12)).
    struct NonResonantFBTerm {
        ComplexType Coeff = 0;
        RealType P1 = 0;
        RealType P12 = 0;
        int xi = 1;
        long Weight = 0;
        ComplexType operator()(ComplexType z1, ComplexType z2) const {
            return Coeff / ((z1 - P1) * (z1 - double(xi) * z2 - P12));
        }
    };
    /// Resonant term C / (z1 - P), present where |z1 - xi z2| <= DeltaTolerance.
    struct ResonantTerm {
        ComplexType Coeff = 0;
        RealType P = 0;
        int xi = 1;
        long Weight = 0;
        ComplexType operator()(ComplexType z1, ComplexType z2, RealType DeltaTolerance = 1e-16) const {
            return (std::abs(z1 - double(xi) * z2) <= DeltaTolerance) ? (Coeff / (z1 - P)) : ComplexType(0);
        }
    };

private:
    MonomialOperatorPart const& F1;
    MonomialOperatorPart const& F2;
    MonomialOperatorPart const& B1;
    MonomialOperatorPart const& B2;
    HamiltonianPart const &Hpart1, &Hpart2, &Hpart3;
    DensityMatrixPart const &DMpart1, &DMpart2, &DMpart3;
    Channel channel;
    bool SwappedFermionOps;
    RealType PoleResolution;
    RealType CoefficientTolerance;
    std::shared_ptr<pomx_chi3_terms> terms; // device terms (kept when compute() is not cleared)
    int64_t counts[3] = {0, 0, 0};
    int64_t raw_terms = 0;

    pomx_chi3_part devicePart() const;
    void run(int nfreq, double const* d_freq, double* d_acc, bool keep);
    std::vector<double> hostTerms() const;

public:
    ThreePointSusceptibilityPart(MonomialOperatorPart const& F1, MonomialOperatorPart const& F2,
                                 MonomialOperatorPart const& B1, MonomialOperatorPart const& B2,
                                 HamiltonianPart const& Hpart1, HamiltonianPart const& Hpart2,
                                 HamiltonianPart const& Hpart3, DensityMatrixPart const& DMpart1,
                                 DensityMatrixPart const& DMpart2, DensityMatrixPart const& DMpart3, Channel channel,
                                 bool SwappedFermionOps, RealType PoleResolution, RealType CoefficientTolerance);

    void compute();
    void clear();
    ComplexType operator()(ComplexType z1, ComplexType z2) const;
    ComplexType operator()(long MatsubaraNumber1, long MatsubaraNumber2) const;

    std::size_t getNumResonantTerms() const { return std::size_t(counts[2]); }
    std::size_t getNumNonResonantFFTerms() const { return std::size_t(counts[0]); }
    std::size_t getNumNonResonantFBTerms() const { return std::size_t(counts[1]); }
    /// Terms emitted before merging.
    int64_t getNumRawTerms() const { return raw_terms; }
    bool getSwappedFermionOps() const { return SwappedFermionOps; }
    std::array<BlockNumber, 3> getBlocks() const {
        return {Hpart1.getBlockNumber(), Hpart2.getBlockNumber(), Hpart3.getBlockNumber()};
    }
    /// Copies of the merged terms (Weight is not kept and reads 0).
    std::vector<ResonantTerm> getResonantTerms() const;
    std::vector<NonResonantFFTerm> getNonResonantFFTerms() const;
    std::vector<NonResonantFBTerm> getNonResonantFBTerms() const;
};

class ThreePointSusceptibility : public Thermal, public ComputableObject {
    friend class ThreePointSusceptibilityContainer;

    Channel channel;
    StatesClassification const& S;
    Hamiltonian const& H;
    CreationOperator const& CX1;
    AnnihilationOperator const& C2;
    CreationOperator const& CX3;
    AnnihilationOperator const& C4;
    DensityMatrix const& DM;
    std::vector<ThreePointSusceptibilityPart> parts;

    CreationOperator const& getF1() const;
    MonomialOperator const& getF2() const;
    MonomialOperator const& getB1() const;
    MonomialOperator const& getB2() const;

protected:
    bool Vanishing = true;

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-16;

    ThreePointSusceptibility(Channel channel, StatesClassification const& S, Hamiltonian const& H,
                             CreationOperator const& CX1, AnnihilationOperator const& C2, CreationOperator const& CX3,
                             AnnihilationOperator const& C4, DensityMatrix const& DM);

    void prepare();
    std::vector<ComplexType> compute(bool clear = false, FreqVec2 const& freqs = {}, MPI_Comm const& comm = MPI_COMM_WORLD);
    ParticleIndex getIndex(std::size_t Position) const;
    ComplexType operator()(ComplexType z1, ComplexType z2) const;
    ComplexType operator()(long MatsubaraNumber1, long MatsubaraNumber2) const;
    /// Values at many pairs from the kept terms (compute() with clear == false), one device
    /// evaluation per part.
    std::vector<ComplexType> evaluate(FreqVec2 const& freqs) const;
    bool isVanishing() const { return Vanishing; }
    Channel getChannel() const { return channel; }
    std::vector<ThreePointSusceptibilityPart> const& getParts() const { return parts; }
    std::size_t getNumParts() const { return parts.size(); }
};

/// Three-point susceptibilities for a set of index combinations plus the permutation images the
/// reference's IndexContainer4 adds (IndexContainer4.hpp:171-273,
/// ThreePointSusceptibilityContainer.cpp:19-48).
class ThreePointSusceptibilityContainer : public Thermal {
    Channel channel;
    ParticleIndex NumIndices;
    StatesClassification const& S;
    Hamiltonian const& H;
    DensityMatrix const& DM;
    FieldOperatorContainer const& Operators;
    std::map<IndexCombination4, ElementWithPermFreq<ThreePointSusceptibility>> ElementsMap;
    std::map<IndexCombination4, std::shared_ptr<ThreePointSusceptibility>> NonTrivialElements;

    ElementWithPermFreq<ThreePointSusceptibility>& create(IndexCombination4 const& Indices);

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-16;

    template <typename... IndexTypes>
    ThreePointSusceptibilityContainer(Channel channel, IndexClassification<IndexTypes...> const& IndexInfo,
                                      StatesClassification const& S, Hamiltonian const& H, DensityMatrix const& DM,
                                      FieldOperatorContainer const& Ops)
        : Thermal(DM.beta), channel(channel), NumIndices(IndexInfo.getIndexSize()), S(S), H(H), DM(DM), Operators(Ops) {}

    void fill(std::set<IndexCombination4> Indices = {});
    void prepareAll(std::set<IndexCombination4> const& Indices = {});
    std::map<IndexCombination4, std::vector<ComplexType>> computeAll(bool clearTerms = false, FreqVec2 const& freqs = {},
                                                                     MPI_Comm const& comm = MPI_COMM_WORLD);
    bool isInContainer(IndexCombination4 const& Indices) const { return ElementsMap.count(Indices) > 0; }
    bool isInContainer(ParticleIndex i1, ParticleIndex i2, ParticleIndex i3, ParticleIndex i4) const {
        return isInContainer(IndexCombination4(i1, i2, i3, i4));
    }
    ElementWithPermFreq<ThreePointSusceptibility>& operator()(IndexCombination4 const& Indices);
    ElementWithPermFreq<ThreePointSusceptibility>& operator()(ParticleIndex i1, ParticleIndex i2, ParticleIndex i3,
                                                              ParticleIndex i4) {
        return (*this)(IndexCombination4(i1, i2, i3, i4));
    }
    std::map<IndexCombination4, std::shared_ptr<ThreePointSusceptibility>> const& getNonTrivialElements() const {
        return NonTrivialElements;
    }
};

} // namespace Pomerol
