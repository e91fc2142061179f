// Two-particle Green's function <c_i c_j c^+_k c^+_l>, its parts and the index container.
//
// References: TwoParticleGF.hpp/.cpp (:36-204), TwoParticleGFPart.hpp/.cpp (term kinds
// :68-308, generation :189-298, evaluation :300-322, .hpp:475-495), TwoParticleGFContainer
// .hpp/.cpp (:23-136), IndexContainer4.hpp (:137-273).
//
// Each part's terms are generated, merged and (for compute(clear, freqs)) evaluated on the
// frequency list on the device (pomx_g2_part_compute).  With clear == false the merged terms
// stay on the device so operator() can evaluate them later.
#pragma once

#include "GreensFunction.hpp"

#include <array>
#include <map>
#include <memory>
#include <set>
#include <tuple>

namespace Pomerol {

class SpectrumTangent;

using FreqTuple3 = std::tuple<ComplexType, ComplexType, ComplexType>;
using FreqVec3 = std::vector<FreqTuple3>;

class TwoParticleGFPart : public Thermal, public ComputableObject {
    friend class TwoParticleGF;

public:
    /// Non-resonant term C/((z1-P1)(z2-P2)(z3-P3)), z2 -> z1+z2+z3-P1-P2-P3 if isz4.
    struct NonResonantTerm {
        ComplexType Coeff = 0;
        std::array<RealType, 3> Poles = {{0, 0, 0}};
        bool isz4 = false;
        long Weight = 0;
        ComplexType operator()(ComplexType z1, ComplexType z2, ComplexType z3) const {
            return isz4 ? Coeff / ((z1 - Poles[0]) * (z1 + z2 + z3 - Poles[0] - Poles[1] - Poles[2]) * (z3 - Poles[2]))
                        : Coeff / ((z1 - Poles[0]) * (z2 - Poles[1]) * (z3 - Poles[2]));
        }
    };
    /// Resonant term (TwoParticleGFPart.hpp:178-307).
    struct ResonantTerm {
        ComplexType ResCoeff = 0;
        ComplexType NonResCoeff = 0;
        std::array<RealType, 3> Poles = {{0, 0, 0}};
        bool isz1z2 = false;
        long Weight = 0;
        ComplexType operator()(ComplexType z1, ComplexType z2, ComplexType z3, RealType DeltaTolerance = 1e-16) const {
            ComplexType Diff = isz1z2 ? z1 + z2 - Poles[0] - Poles[1] : z2 + z3 - Poles[1] - Poles[2];
            return (std::abs(Diff) <= DeltaTolerance ? ResCoeff : (NonResCoeff / Diff)) / ((z1 - Poles[0]) * (z3 - Poles[2]));
        }
    };

private:
    MonomialOperatorPart const& O1;
    MonomialOperatorPart const& O2;
    MonomialOperatorPart const& O3;
    MonomialOperatorPart const& CX4;
    HamiltonianPart const &Hpart1, &Hpart2, &Hpart3, &Hpart4;
    DensityMatrixPart const &DMpart1, &DMpart2, &DMpart3, &DMpart4;
    Permutation3 Permutation;
    RealType PoleResolution;
    RealType CoefficientTolerance;
    std::shared_ptr<pomx_g2_terms> terms; // device terms (kept when compute() is not cleared)
    int64_t counts[4] = {0, 0, 0, 0};
    int64_t raw_terms = 0;
    int perm_index = 0;

    pomx_g2_part devicePart() const;
    void run(int nfreq, double const* d_freq, double* d_acc, bool keep);
    void runTangent(SpectrumTangent& T, int nfreq, double const* d_freq, double* d_acc, double* d_dacc);

public:
    TwoParticleGFPart(MonomialOperatorPart const& O1, MonomialOperatorPart const& O2, MonomialOperatorPart const& O3,
                      MonomialOperatorPart const& CX4, HamiltonianPart const& Hpart1, HamiltonianPart const& Hpart2,
                      HamiltonianPart const& Hpart3, HamiltonianPart const& Hpart4, DensityMatrixPart const& DMpart1,
                      DensityMatrixPart const& DMpart2, DensityMatrixPart const& DMpart3, DensityMatrixPart const& DMpart4,
                      Permutation3 Permutation, RealType PoleResolution, RealType CoefficientTolerance);

    void compute();
    void clear();
    ComplexType operator()(ComplexType z1, ComplexType z2, ComplexType z3) const;
    ComplexType operator()(long MatsubaraNumber1, long MatsubaraNumber2, long MatsubaraNumber3) const;

    std::size_t getNumResonantTerms() const { return std::size_t(counts[2] + counts[3]); }
    std::size_t getNumNonResonantTerms() const { return std::size_t(counts[0] + counts[1]); }
    /// Terms emitted before merging.
    int64_t getNumRawTerms() const { return raw_terms; }
    Permutation3 const& getPermutation() const { return Permutation; }
    std::array<BlockNumber, 4> getBlocks() const {
        return {Hpart1.getBlockNumber(), Hpart2.getBlockNumber(), Hpart3.getBlockNumber(), Hpart4.getBlockNumber()};
    }
    std::vector<NonResonantTerm> getNonResonantTerms() const;
    std::vector<ResonantTerm> getResonantTerms() const;
};

class TwoParticleGF : public Thermal, public ComputableObject {
    friend class TwoParticleGFContainer;
    StatesClassification const& S;
    Hamiltonian const& H;
    AnnihilationOperator const& C1;
    AnnihilationOperator const& C2;
    CreationOperator const& CX3;
    CreationOperator const& CX4;
    DensityMatrix const& DM;
    std::vector<TwoParticleGFPart> parts;

protected:
    bool Vanishing = true;
    MonomialOperatorPart const& OperatorPartAtPosition(std::size_t PermutationNumber, std::size_t OperatorPosition,
                                                      BlockNumber LeftIndex) const;
    BlockNumber getLeftIndex(std::size_t PermutationNumber, std::size_t OperatorPosition, BlockNumber RightIndex) const;
    BlockNumber getRightIndex(std::size_t PermutationNumber, std::size_t OperatorPosition, BlockNumber LeftIndex) const;

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-16;

    TwoParticleGF(StatesClassification const& S, Hamiltonian const& H, AnnihilationOperator const& C1,
                  AnnihilationOperator const& C2, CreationOperator const& CX3, CreationOperator const& CX4,
                  DensityMatrix const& DM);

    void prepare();
    std::vector<ComplexType> compute(bool clear = false, FreqVec3 const& freqs = {}, MPI_Comm const& comm = MPI_COMM_WORLD);
    /// As compute(), restricted to the listed parts (bounded samples of very large workloads).
    std::vector<ComplexType> computeParts(std::vector<std::size_t> const& which, bool clear, FreqVec3 const& freqs);
    /// Derivative along the direction of T (computed) at the triplets, from the terms compute()
    /// keeps, with their thresholds and merge groups held fixed; *values (if given) receives the
    /// value from the same terms.  No counterpart in the reference.
    std::vector<ComplexType> computeTangent(SpectrumTangent& T, FreqVec3 const& freqs,
                                            std::vector<ComplexType>* values = nullptr);
    ParticleIndex getIndex(std::size_t Position) const;
    ComplexType operator()(ComplexType z1, ComplexType z2, ComplexType z3) const;
    ComplexType operator()(long MatsubaraNumber1, long MatsubaraNumber2, long MatsubaraNumber3) const;
    /// Values at many triplets from the kept terms (compute() with clear == false), one device
    /// evaluation per part.
    std::vector<ComplexType> evaluate(FreqVec3 const& freqs) const;
    bool isVanishing() const { return Vanishing; }
    std::vector<TwoParticleGFPart> const& getParts() const { return parts; }
    std::size_t getNumParts() const { return parts.size(); }
};

struct IndexCombination4 {
    ParticleIndex Index1, Index2, Index3, Index4;
    IndexCombination4(ParticleIndex i1, ParticleIndex i2, ParticleIndex i3, ParticleIndex i4)
        : Index1(i1), Index2(i2), Index3(i3), Index4(i4) {}
    bool operator<(IndexCombination4 const& o) const {
        return std::tie(Index1, Index2, Index3, Index4) < std::tie(o.Index1, o.Index2, o.Index3, o.Index4);
    }
    bool operator==(IndexCombination4 const& o) const {
        return Index1 == o.Index1 && Index2 == o.Index2 && Index3 == o.Index3 && Index4 == o.Index4;
    }
    bool operator!=(IndexCombination4 const& o) const { return !(*this == o); }
    friend std::ostream& operator<<(std::ostream& os, IndexCombination4 const& ic) {
        return os << "(" << ic.Index1 << ic.Index2 << ic.Index3 << ic.Index4 << ")";
    }
};

/// An element with a frequency permutation: a crossing-symmetry image of a computed element
/// (IndexContainer4.hpp:137-167).
template <typename ElementType> class ElementWithPermFreq {
    std::shared_ptr<ElementType> pElement;

public:
    Permutation4 FrequenciesPermutation;
    ElementWithPermFreq(std::shared_ptr<ElementType> p, Permutation4 const& perm) : pElement(p), FrequenciesPermutation(perm) {}
    ComplexType operator()(long n1, long n2, long n3) const {
        long n[4] = {n1, n2, n3, n1 + n2 - n3};
        return (*pElement)(n[FrequenciesPermutation.perm[0]], n[FrequenciesPermutation.perm[1]],
                           n[FrequenciesPermutation.perm[2]]) *
               RealType(FrequenciesPermutation.sign);
    }
    operator ElementType&() { return *pElement; }
    operator ElementType const&() const { return *pElement; }
    std::shared_ptr<ElementType> element() const { return pElement; }
};

/// 2PGF elements for a set of index combinations plus their permutation images
/// (TwoParticleGFContainer.cpp:23-136, IndexContainer4.hpp:171-273).
class TwoParticleGFContainer : public Thermal {
    ParticleIndex NumIndices;
    StatesClassification const& S;
    Hamiltonian const& H;
    DensityMatrix const& DM;
    FieldOperatorContainer const& Operators;
    std::map<IndexCombination4, ElementWithPermFreq<TwoParticleGF>> ElementsMap;
    std::map<IndexCombination4, std::shared_ptr<TwoParticleGF>> NonTrivialElements;

    ElementWithPermFreq<TwoParticleGF>& create(IndexCombination4 const& Indices);

public:
    RealType PoleResolution = 1e-8;
    RealType CoefficientTolerance = 1e-16;

    template <typename... IndexTypes>
    TwoParticleGFContainer(IndexClassification<IndexTypes...> const& IndexInfo, StatesClassification const& S,
                           Hamiltonian const& H, DensityMatrix const& DM, FieldOperatorContainer const& Ops)
        : Thermal(DM.beta), NumIndices(IndexInfo.getIndexSize()), S(S), H(H), DM(DM), Operators(Ops) {}

    void fill(std::set<IndexCombination4> Indices = {});
    void prepareAll(std::set<IndexCombination4> const& Indices = {});
    std::map<IndexCombination4, std::vector<ComplexType>> computeAll(bool clearTerms = false, FreqVec3 const& freqs = {},
                                                                     MPI_Comm const& comm = MPI_COMM_WORLD, bool split = true);
    bool isInContainer(IndexCombination4 const& Indices) const { return ElementsMap.count(Indices) > 0; }
    ElementWithPermFreq<TwoParticleGF>& operator()(IndexCombination4 const& Indices);
    ElementWithPermFreq<TwoParticleGF>& operator()(ParticleIndex i1, ParticleIndex i2, ParticleIndex i3, ParticleIndex i4) {
        return (*this)(IndexCombination4(i1, i2, i3, i4));
    }
    std::map<IndexCombination4, std::shared_ptr<TwoParticleGF>> const& getNonTrivialElements() const {
        return NonTrivialElements;
    }
};

} // namespace Pomerol
