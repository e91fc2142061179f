// Monomial operators in the Hamiltonian eigenbasis.
//
// References: MonomialOperator.hpp:40-439 (block connectivity via find_connections, the
// Creation/Annihilation/Field/Quadratic/Quartic operator classes), MonomialOperator.cpp:33-78,
// MonomialOperatorPart.hpp/.cpp:32-111 (rotation U_L^dagger (M U_R), sparseView(Tolerance),
// setFromAdjoint), FieldOperatorContainer.hpp/.cpp:22-37, EnsembleAverage.hpp/.cpp:27-67.
//
// All parts of an operator are rotated in one batched FP64 MFMA launch (pomx_rotate_batched)
// into a device arena.  An annihilation part computed by setFromAdjoint is a transposed view of
// the creation part's storage (no copy).  getRowMajorValue/getColMajorValue return the block as
// a dense host matrix (entries dropped by sparseView are exact zeros).
#pragma once

#include "Hamiltonian.hpp"

#include <map>
#include <memory>
#include <set>
#include <tuple>

namespace Pomerol {

using BlockMapping = std::pair<BlockNumber, BlockNumber>; // (left, right) = (to, from)

/// Two-sided block map of an operator: left[to] = from, right[from] = to (the views of the
/// reference's boost::bimap, MonomialOperator.hpp:60).
struct BlocksBimap {
    std::map<BlockNumber, BlockNumber> left, right;
    void insert(BlockMapping const& m) {
        left.emplace(m.first, m.second);
        right.emplace(m.second, m.first);
    }
    std::size_t size() const { return left.size(); }
};

class MonomialOperatorPart : public ComputableObject {
    std::shared_ptr<FermionOperator const> owned; // the operator, when the part was given a temporary
    FermionOperator const* MOp;
    StatesClassification const& S;
    HamiltonianPart const& HFrom;
    HamiltonianPart const& HTo;
    std::shared_ptr<detail::DeviceArray<double>> store;
    int64_t off = 0, rs = 0, cs = 0;
    mutable std::shared_ptr<MatrixType<false>> host;

    friend class MonomialOperator;

public:
    MonomialOperatorPart(FermionOperator const& MOp, StatesClassification const& S, HamiltonianPart const& HFrom,
                         HamiltonianPart const& HTo);
    /// From a temporary operator: the part keeps its own copy.
    MonomialOperatorPart(FermionOperator&& MOp, StatesClassification const& S, HamiltonianPart const& HFrom,
                         HamiltonianPart const& HTo);

    /// Rotate this part alone (MonomialOperatorPart.cpp:32-92).
    void compute(RealType Tolerance = 1e-8);
    /// This part becomes the adjoint of `part` (MonomialOperatorPart.cpp:94-111).
    void setFromAdjoint(MonomialOperatorPart const& part);

    bool isComplex() const { return false; }
    BlockNumber getLeftIndex() const { return HTo.getBlockNumber(); }
    BlockNumber getRightIndex() const { return HFrom.getBlockNumber(); }
    long rows() const { return long(HTo.getSize()); }
    long cols() const { return long(HFrom.getSize()); }
    template <bool C> MatrixType<C> const& getRowMajorValue() const;
    template <bool C> MatrixType<C> const& getColMajorValue() const { return getRowMajorValue<C>(); }
    /// Device view of the rotated block: element (r, c) at base[r * rs + c * cs].
    pomx_view deviceView() const { return pomx_view{store->data() + off, rs, cs}; }

    /// The rotated block, as MonomialOperatorPart.hpp:130-136 prints it.
    friend std::ostream& operator<<(std::ostream& os, MonomialOperatorPart const& part);
};

template <> MatrixType<false> const& MonomialOperatorPart::getRowMajorValue<false>() const;
template <> MatrixType<true> const& MonomialOperatorPart::getRowMajorValue<true>() const;
inline std::ostream& operator<<(std::ostream& os, MonomialOperatorPart const& part) {
    return os << part.getRowMajorValue<false>() << '\n';
}

class MonomialOperator : public ComputableObject {
protected:
    friend class FieldOperatorContainer;
    FermionOperator MOp;
    StatesClassification const& S;
    Hamiltonian const& H;
    std::map<BlockNumber, std::size_t> mapPartsFromRight, mapPartsFromLeft;
    BlocksBimap LeftRightBlocks;
    std::vector<MonomialOperatorPart> parts;
    std::shared_ptr<detail::DeviceArray<double>> store;

public:
    template <typename ScalarType, typename... IndexTypes>
    MonomialOperator(Operators::expression<ScalarType, IndexTypes...> const& MO, HilbertSpace<IndexTypes...> const& HS,
                     StatesClassification const& S, Hamiltonian const& H)
        : MOp(HS.getIndexInfo().toOperator(MO)), S(S), H(H) {
        if(MO.size() > 1)
            throw std::runtime_error("Only monomial expressions are supported");
    }
    MonomialOperator(MonomialOperator&&) = default;

    bool isComplex() const { return false; }

    template <typename... IndexTypes> void prepare(HilbertSpace<IndexTypes...> const& HS) {
        if(getStatus() >= Prepared)
            return;
        prepareImpl(HS.getStatus() == Computed ? &HS.getSpacePartition() : nullptr);
    }
    void compute(RealType Tolerance = 1e-8, MPI_Comm const& comm = MPI_COMM_WORLD);

    MonomialOperatorPart& getPartFromLeftIndex(BlockNumber LeftIndex);
    MonomialOperatorPart const& getPartFromLeftIndex(BlockNumber LeftIndex) const;
    MonomialOperatorPart& getPartFromRightIndex(BlockNumber RightIndex);
    MonomialOperatorPart const& getPartFromRightIndex(BlockNumber RightIndex) const;
    BlockNumber getLeftIndex(BlockNumber RightIndex) const;
    BlockNumber getRightIndex(BlockNumber LeftIndex) const;
    BlocksBimap const& getBlockMapping() const;
    std::vector<MonomialOperatorPart> const& getParts() const { return parts; }

    /// Rotate the parts of several operators in one batched launch.
    static void computeBatch(std::vector<MonomialOperator*> const& ops, RealType Tolerance);

private:
    void prepareImpl(SpacePartition const* partition);
    void checkPrepared() const;
};

class FieldOperator : public MonomialOperator {
    ParticleIndex Index;

public:
    template <typename ScalarType, typename... IndexTypes>
    FieldOperator(Operators::expression<ScalarType, IndexTypes...> const& FO, HilbertSpace<IndexTypes...> const& HS,
                  StatesClassification const& S, Hamiltonian const& H, ParticleIndex Index)
        : MonomialOperator(FO, HS, S, H), Index(Index) {
        if(FO.size() != 1 || FO.get_monomials().begin()->first.size() != 1)
            throw std::runtime_error("Expected a single-fermion monomial expression");
    }
    ParticleIndex getIndex() const { return Index; }
};

namespace detail {
template <typename... IndexTypes>
Operators::expression<RealType, IndexTypes...> field_expr(bool dag, std::tuple<IndexTypes...> const& info) {
    using E = Operators::expression<RealType, IndexTypes...>;
    return E(1.0, {typename E::op_type{dag, info}});
}
} // namespace detail

class CreationOperator : public FieldOperator {
public:
    template <typename... IndexTypes>
    CreationOperator(IndexClassification<IndexTypes...> const& IndexInfo, HilbertSpace<IndexTypes...> const& HS,
                     StatesClassification const& S, Hamiltonian const& H, ParticleIndex Index)
        : FieldOperator(detail::field_expr(true, IndexInfo.getInfo(Index)), HS, S, H, Index) {}
};

class AnnihilationOperator : public FieldOperator {
public:
    template <typename... IndexTypes>
    AnnihilationOperator(IndexClassification<IndexTypes...> const& IndexInfo, HilbertSpace<IndexTypes...> const& HS,
                         StatesClassification const& S, Hamiltonian const& H, ParticleIndex Index)
        : FieldOperator(detail::field_expr(false, IndexInfo.getInfo(Index)), HS, S, H, Index) {}
};

/// Product of two field operators, c^+_1 c_2 by default (MonomialOperator.hpp:288-346).
class QuadraticOperator : public MonomialOperator {
    ParticleIndex Index1, Index2;
    std::tuple<bool, bool> Dagger;

public:
    template <typename... IndexTypes>
    QuadraticOperator(IndexClassification<IndexTypes...> const& IndexInfo, HilbertSpace<IndexTypes...> const& HS,
                      StatesClassification const& S, Hamiltonian const& H, ParticleIndex Index1, ParticleIndex Index2,
                      std::tuple<bool, bool> const& Dagger = std::tuple<bool, bool>(true, false))
        : MonomialOperator(detail::field_expr(std::get<0>(Dagger), IndexInfo.getInfo(Index1)) *
                               detail::field_expr(std::get<1>(Dagger), IndexInfo.getInfo(Index2)),
                           HS, S, H),
          Index1(Index1), Index2(Index2), Dagger(Dagger) {}
    ParticleIndex getIndex1() const { return Index1; }
    ParticleIndex getIndex2() const { return Index2; }
    std::tuple<bool, bool> const& getDagger() const { return Dagger; }
};

/// Product of four field operators, c^+_1 c^+_2 c_3 c_4 by default (MonomialOperator.hpp:351-439).
class QuarticOperator : public MonomialOperator {
    ParticleIndex Index1, Index2, Index3, Index4;
    std::tuple<bool, bool, bool, bool> Dagger;

public:
    template <typename... IndexTypes>
    QuarticOperator(IndexClassification<IndexTypes...> const& IndexInfo, HilbertSpace<IndexTypes...> const& HS,
                    StatesClassification const& S, Hamiltonian const& H, ParticleIndex Index1, ParticleIndex Index2,
                    ParticleIndex Index3, ParticleIndex Index4,
                    std::tuple<bool, bool, bool, bool> const& Dagger = std::tuple<bool, bool, bool, bool>(true, true, false, false))
        : MonomialOperator(detail::field_expr(std::get<0>(Dagger), IndexInfo.getInfo(Index1)) *
                               detail::field_expr(std::get<1>(Dagger), IndexInfo.getInfo(Index2)) *
                               detail::field_expr(std::get<2>(Dagger), IndexInfo.getInfo(Index3)) *
                               detail::field_expr(std::get<3>(Dagger), IndexInfo.getInfo(Index4)),
                           HS, S, H),
          Index1(Index1), Index2(Index2), Index3(Index3), Index4(Index4), Dagger(Dagger) {}
};

/// c and c^+ for a set of indices; c is obtained as the adjoint of c^+ (FieldOperatorContainer.cpp:22-37).
class FieldOperatorContainer {
    std::map<ParticleIndex, CreationOperator> mapCreationOperators;
    std::map<ParticleIndex, AnnihilationOperator> mapAnnihilationOperators;

public:
    template <typename... IndexTypes>
    FieldOperatorContainer(IndexClassification<IndexTypes...> const& IndexInfo, HilbertSpace<IndexTypes...> const& HS,
                           StatesClassification const& S, Hamiltonian const& H, std::set<ParticleIndex> in = {}) {
        if(in.empty())
            for(ParticleIndex i = 0; i < IndexInfo.getIndexSize(); ++i)
                in.insert(i);
        for(ParticleIndex i : in) {
            mapCreationOperators.emplace(std::piecewise_construct, std::forward_as_tuple(i),
                                         std::forward_as_tuple(IndexInfo, HS, S, H, i));
            mapAnnihilationOperators.emplace(std::piecewise_construct, std::forward_as_tuple(i),
                                             std::forward_as_tuple(IndexInfo, HS, S, H, i));
        }
    }
    template <typename... IndexTypes> void prepareAll(HilbertSpace<IndexTypes...> const& HS) {
        for(auto& o : mapCreationOperators)
            o.second.prepare(HS);
        for(auto& o : mapAnnihilationOperators)
            o.second.prepare(HS);
    }
    void computeAll(RealType Tolerance = 1e-8);
    CreationOperator const& getCreationOperator(ParticleIndex in) const;
    AnnihilationOperator const& getAnnihilationOperator(ParticleIndex in) const;
};

/// Thermal average Tr[rho A] over the diagonal blocks of A (EnsembleAverage.cpp:27-67).
class EnsembleAverage : public Thermal, public ComputableObject {
    MonomialOperator const& A;
    DensityMatrix const& DM;
    ComplexType Result = 0;

public:
    EnsembleAverage(MonomialOperator const& A, DensityMatrix const& DM);
    void compute();
    ComplexType operator()() const { return Result; }
};

} // namespace Pomerol
