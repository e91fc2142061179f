// Index classification, Fock space, sector partition and state classification.
//
// References: IndexClassification.hpp:40-151 (sorted index tuples -> ParticleIndex),
// HilbertSpace.hpp:40-200 (Fock space of the Hamiltonian's generators; compute() = Phase I
// connected components of H plus Phase II merging so that every c_i, c^dag_i maps a sector into
// at most one sector), StatesClassification.hpp:30-127.
//
// Fock state bit i is ParticleIndex i.  A c_i / c^dag_i acting on a Fock state picks up the
// sign (-1)^(number of occupied modes j < i).  Sectors are numbered by their smallest Fock
// state and list their states in ascending order.
#pragma once

#include <ostream>
#include <tuple>

#include "Misc.hpp"
#include "Operators.hpp"

#include <map>
#include <memory>
#include <set>
#include <tuple>
#include <utility>
#include <vector>

namespace Pomerol {

/// A second-quantised operator over integer modes (the role of libcommute's loperator).
class FermionOperator {
public:
    struct Term {
        RealType coeff;
        std::vector<std::pair<bool, int>> ops; // (dagger, mode), rightmost acts first
    };
    std::vector<Term> terms;
    bool complex = false;

    /// Apply a single monomial: returns the sign (0 if the state is annihilated).
    static int apply_monomial(std::vector<std::pair<bool, int>> const& ops, QuantumState s, QuantumState& out);
    /// H|s> as a list of (target, amplitude) with equal targets combined.
    void apply(QuantumState s, std::vector<std::pair<QuantumState, RealType>>& out) const;
    bool is_monomial() const { return terms.size() == 1; }
};

/// Partition of the Fock space into invariant subspaces.
class SpacePartition {
public:
    std::vector<std::vector<QuantumState>> blocks;
    std::vector<BlockNumber> state_block;
    std::vector<InnerQuantumState> inner;
    QuantumState dim() const { return state_block.size(); }
    std::size_t n_subspaces() const { return blocks.size(); }
    /// (from, to) block pairs connected by a monomial operator (libcommute find_connections).
    std::set<std::pair<BlockNumber, BlockNumber>> find_connections(FermionOperator const& op) const;
};

/// Phase I + Phase II auto-partition (HilbertSpace.hpp:144-169).
SpacePartition make_space_partition(int n_modes, FermionOperator const& H);

template <typename... IndexTypes> class IndexClassification {
public:
    using IndexInfo = std::tuple<IndexTypes...>;

private:
    std::vector<IndexInfo> info_;
    std::map<IndexInfo, ParticleIndex> map_;

public:
    IndexClassification() = default;
    template <typename ScalarType> explicit IndexClassification(Operators::expression<ScalarType, IndexTypes...> const& e) {
        std::set<IndexInfo> s;
        for(auto const& t : e.get_monomials())
            for(auto const& o : t.first)
                s.insert(o.index);
        for(auto const& t : s) {
            map_.emplace(t, ParticleIndex(info_.size()));
            info_.push_back(t);
        }
    }
    ParticleIndex getIndexSize() const { return ParticleIndex(info_.size()); }
    bool checkIndex(ParticleIndex in) const { return in < info_.size(); }
    ParticleIndex getIndex(IndexTypes... args) const { return getIndex(IndexInfo(args...)); }
    ParticleIndex getIndex(IndexInfo const& t) const {
        auto it = map_.find(t);
        if(it == map_.end())
            throw std::runtime_error("IndexClassification: wrong index");
        return it->second;
    }
    IndexInfo const& getInfo(ParticleIndex in) const {
        if(!checkIndex(in))
            throw std::runtime_error("IndexClassification: wrong index " + std::to_string(in));
        return info_[in];
    }
    /// "Index i = (label, orbital, spin)" per line (IndexClassification.hpp:122-129).
    friend std::ostream& operator<<(std::ostream& os, IndexClassification const& ic) {
        for(ParticleIndex i = 0; i < ic.info_.size(); ++i) {
            os << "Index " << i << " = (";
            std::apply(
                [&os](auto const&... x) {
                    int k = 0;
                    ((os << (k++ ? ", " : "") << x), ...);
                },
                ic.info_[i]);
            os << ")\n";
        }
        return os;
    }

    template <typename ScalarType> FermionOperator toOperator(Operators::expression<ScalarType, IndexTypes...> const& e) const {
        FermionOperator op;
        op.complex = std::is_same<ScalarType, ComplexType>::value;
        for(auto const& t : e.get_monomials()) {
            FermionOperator::Term term;
            term.coeff = std::real(t.second);
            if(std::imag(ComplexType(t.second)) != 0.0)
                throw std::runtime_error("complex-valued operators are not supported by the GPU path");
            for(auto const& o : t.first)
                term.ops.emplace_back(o.dag, int(getIndex(o.index)));
            op.terms.push_back(term);
        }
        return op;
    }
};

template <typename ScalarType, typename... IndexTypes>
IndexClassification<IndexTypes...> MakeIndexClassification(Operators::expression<ScalarType, IndexTypes...> const& e) {
    return IndexClassification<IndexTypes...>(e);
}

/// The full Fock space of a set of indices (the role of libcommute's hilbert_space): the index
/// classification that maps an expression's operators to modes.
template <typename... IndexTypes> class FullHilbertSpace {
    IndexClassification<IndexTypes...> const* info_;

public:
    explicit FullHilbertSpace(IndexClassification<IndexTypes...> const& info) : info_(&info) {}
    IndexClassification<IndexTypes...> const& getIndexInfo() const { return *info_; }
    QuantumState dim() const { return QuantumState(1) << info_->getIndexSize(); }
};

/// An expression as a linear operator on a full Fock space (the role of libcommute's loperator,
/// Misc.hpp's LOperatorType<ScalarType>); real coefficients on the GPU path.
template <typename ScalarType> class LOperatorType : public FermionOperator {
public:
    template <typename S2, typename... IndexTypes>
    LOperatorType(Operators::expression<S2, IndexTypes...> const& e, FullHilbertSpace<IndexTypes...> const& hs)
        : FermionOperator(hs.getIndexInfo().toOperator(e)) {}
};

template <typename... IndexTypes> class HilbertSpace : public ComputableObject {
    IndexClassification<IndexTypes...> const& IndexInfo;
    FermionOperator HOp;
    std::unique_ptr<SpacePartition> Partition;

public:
    template <typename ScalarType>
    HilbertSpace(IndexClassification<IndexTypes...> const& IndexInfo, Operators::expression<ScalarType, IndexTypes...> const& H)
        : IndexInfo(IndexInfo), HOp(IndexInfo.toOperator(H)) {
        if(IndexInfo.getIndexSize() > 30)
            throw std::runtime_error("HilbertSpace: more than 30 modes are not supported");
    }

    void compute() {
        if(getStatus() >= Computed)
            return;
        Partition.reset(new SpacePartition(make_space_partition(int(IndexInfo.getIndexSize()), HOp)));
        setStatus(Computed);
    }

    IndexClassification<IndexTypes...> const& getIndexInfo() const { return IndexInfo; }
    FullHilbertSpace<IndexTypes...> getFullHilbertSpace() const { return FullHilbertSpace<IndexTypes...>(IndexInfo); }
    int getNumModes() const { return int(IndexInfo.getIndexSize()); }
    QuantumState getDim() const { return QuantumState(1) << IndexInfo.getIndexSize(); }
    FermionOperator const& getHamiltonianOperator() const { return HOp; }
    SpacePartition const& getSpacePartition() const {
        if(getStatus() < Computed)
            throw std::runtime_error("Hilbert space partition has not been computed");
        return *Partition;
    }
};

template <typename ScalarType, typename... IndexTypes>
HilbertSpace<IndexTypes...> MakeHilbertSpace(IndexClassification<IndexTypes...> const& IndexInfo,
                                             Operators::expression<ScalarType, IndexTypes...> const& H) {
    return HilbertSpace<IndexTypes...>(IndexInfo, H);
}

/// Block structure of the Fock space (StatesClassification.hpp:30-127).
class StatesClassification : public ComputableObject {
    std::vector<std::vector<QuantumState>> StatesContainer;
    std::vector<BlockNumber> StateBlockIndex;
    std::vector<InnerQuantumState> StateInnerIndex;

public:
    StatesClassification() = default;

    template <typename... IndexTypes> void compute(HilbertSpace<IndexTypes...> const& HS) {
        if(getStatus() == Computed)
            return;
        if(HS.getStatus() == Computed)
            initMultipleBlocks(HS.getSpacePartition());
        else
            initSingleBlock(HS.getDim());
        setStatus(Computed);
    }

    QuantumState getNumberOfStates() const { return StateBlockIndex.size(); }
    BlockNumber getNumberOfBlocks() const { return BlockNumber(StatesContainer.size()); }
    InnerQuantumState getBlockSize(BlockNumber in) const;
    std::vector<QuantumState> const& getFockStates(BlockNumber in) const;
    QuantumState getFockState(BlockNumber in, InnerQuantumState i) const;
    BlockNumber getBlockNumber(QuantumState in) const;
    InnerQuantumState getInnerState(QuantumState in) const;

private:
    void initSingleBlock(QuantumState Dim);
    void initMultipleBlocks(SpacePartition const& p);
    void checkComputed() const;
};

} // namespace Pomerol
