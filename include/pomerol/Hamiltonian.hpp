// Hamiltonian blocks, their eigendecomposition and the many-body density matrix.
//
// References: HamiltonianPart.hpp/.cpp (block assembly :56-76, eigensolve :90-102, accessors),
// Hamiltonian.hpp/.cpp (:27-150), DensityMatrixPart.hpp/.cpp (:23-50), DensityMatrix.hpp/.cpp
// (:25-92).  All blocks of a Hamiltonian live in one device arena; Hamiltonian::compute solves
// every block in one batched gfx950 launch (pomx_eigh_batched).  Host copies of eigenvalues and
// weights are kept for the accessors; eigenvectors stay on the device and are copied to the host
// only when getMatrix() asks for them.
#pragma once

#include "Device.hpp"
#include "HilbertSpace.hpp"

#include <memory>
#include <vector>

namespace Pomerol {

/// Device storage of a set of square blocks: matrices (row-major) and eigenvalue vectors.
struct BlockStore {
    detail::DeviceArray<double> mat;   // sum of n_b^2
    detail::DeviceArray<double> evals; // sum of n_b
};

class HamiltonianPart : public ComputableObject {
    std::shared_ptr<FermionOperator const> owned; // the operator, when the part was given a temporary
    FermionOperator const* HOp;
    StatesClassification const& S;
    BlockNumber Block;
    std::shared_ptr<BlockStore> store;
    int64_t mat_off = 0, val_off = 0;
    RealVectorType Eigenvalues;
    mutable std::shared_ptr<MatrixType<false>> host_matrix;

    friend class Hamiltonian;

public:
    HamiltonianPart(FermionOperator const& HOp, StatesClassification const& S, BlockNumber Block);
    /// From a temporary operator (e.g. LOperatorType<RealType>(expr, HS.getFullHilbertSpace())): the
    /// part keeps its own copy.
    HamiltonianPart(FermionOperator&& HOp, StatesClassification const& S, BlockNumber Block);

    /// Assemble the dense block on the device (HamiltonianPart.cpp:56-76).
    void prepare();
    /// Eigendecompose the block on the device (HamiltonianPart.cpp:78-102).
    void compute();

    bool isComplex() const { return false; }
    BlockNumber getBlockNumber() const { return Block; }
    InnerQuantumState getSize() const;
    RealType getEigenValue(InnerQuantumState state) const;
    RealVectorType const& getEigenValues() const;
    RealType getMinimumEigenvalue() const;
    /// The block before compute(), its eigenvectors (columns) after.
    template <bool C> MatrixType<C> const& getMatrix() const;
    template <bool C> VectorType<C> getEigenState(InnerQuantumState state) const;
    /// Not available on the GPU path (eigenvector truncation); throws std::logic_error.
    bool reduce(RealType Cutoff);

    // device access for the other GPU-backed objects
    double const* deviceMatrix() const { return store->mat.data() + mat_off; }
    double const* deviceEigenvalues() const { return store->evals.data() + val_off; }
    void invalidateHostCache() const { host_matrix.reset(); }

    /// The block (before compute()) or its eigenvectors (after), as HamiltonianPart.hpp:132-139.
    friend std::ostream& operator<<(std::ostream& os, HamiltonianPart const& part);
};

template <> MatrixType<false> const& HamiltonianPart::getMatrix<false>() const;
template <> MatrixType<true> const& HamiltonianPart::getMatrix<true>() const;
template <> VectorType<false> HamiltonianPart::getEigenState<false>(InnerQuantumState state) const;
template <> VectorType<true> HamiltonianPart::getEigenState<true>(InnerQuantumState state) const;
inline std::ostream& operator<<(std::ostream& os, HamiltonianPart const& part) {
    return os << part.getMatrix<false>() << '\n';
}

class Hamiltonian : public ComputableObject {
    StatesClassification const& S;
    FermionOperator HOp;
    std::vector<HamiltonianPart> parts;
    std::shared_ptr<BlockStore> store;
    RealType GroundEnergy = 0;

public:
    explicit Hamiltonian(StatesClassification const& S) : S(S) {}

    template <typename ScalarType, typename... IndexTypes>
    void prepare(Operators::expression<ScalarType, IndexTypes...> const& H, HilbertSpace<IndexTypes...> const& HS,
                 MPI_Comm const& comm = MPI_COMM_WORLD) {
        (void)comm;
        if(getStatus() >= Prepared)
            return;
        HOp = HS.getIndexInfo().toOperator(H);
        prepareImpl();
    }
    void compute(MPI_Comm const& comm = MPI_COMM_WORLD);
    /// Not available on the GPU path; throws std::logic_error.
    void reduce(RealType Cutoff);

    bool isComplex() const { return false; }
    HamiltonianPart const& getPart(BlockNumber Block) const { return parts[Block]; }
    HamiltonianPart& getPart(BlockNumber Block) { return parts[Block]; }
    InnerQuantumState getBlockSize(BlockNumber Block) const;
    RealType getEigenValue(QuantumState state) const;
    RealVectorType const& getEigenValues(BlockNumber Block) const;
    RealVectorType getEigenValues() const;
    RealType getGroundEnergy() const;
    FermionOperator const& getOperator() const { return HOp; }
    /// Device arena of all eigenvalues (block order, the layout DensityMatrix relies on).
    double const* deviceEigenvalues() const { return store->evals.data(); }
    int64_t eigenvalueOffset(BlockNumber b) const { return parts[b].val_off; }

private:
    void prepareImpl();
    void computeGroundEnergy();
};

class DensityMatrixPart : public Thermal {
    HamiltonianPart const& H;
    RealType GroundEnergy;
    RealType Z_part = 0; // never set by the reference either (DensityMatrixPart.hpp:54)
    RealVectorType weights;
    double const* d_weights = nullptr;
    bool Retained = true;

    friend class DensityMatrix;

public:
    DensityMatrixPart(HamiltonianPart const& H, RealType beta, RealType GroundEnergy);
    RealType getWeight(InnerQuantumState s) const { return weights(long(s)); }
    RealVectorType const& getWeights() const { return weights; }
    RealType getAverageEnergy() const;
    RealType getPartialZ() const { return Z_part; }
    void truncate(RealType Tolerance);
    bool isRetained() const { return Retained; }
    double const* deviceWeights() const { return d_weights; }
};

class DensityMatrix : public Thermal, public ComputableObject {
    StatesClassification const& S;
    Hamiltonian const& H;
    std::vector<DensityMatrixPart> parts;
    detail::DeviceArray<double> d_w;
    RealType Z = 0;

public:
    DensityMatrix(StatesClassification const& S, Hamiltonian const& H, RealType beta);
    void prepare();
    void compute();
    RealType getWeight(QuantumState state) const;
    DensityMatrixPart const& getPart(BlockNumber in) const { return parts[in]; }
    RealType getAverageEnergy() const;
    void truncateBlocks(RealType Tolerance, bool verbose = true);
    bool isRetained(BlockNumber B) const { return parts[B].isRetained(); }
    RealType getPartitionFunction() const { return Z; }
};

} // namespace Pomerol
