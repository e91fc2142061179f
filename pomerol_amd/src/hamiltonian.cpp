// Hamiltonian blocks (assembly + batched eigensolve on the device) and the density matrix.
#include "pomerol/Hamiltonian.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace Pomerol {

namespace {

// Host side of K1: the non-zero entries of every block as absolute positions in the arena
// (HamiltonianPart.cpp:56-76 applies H to each basis ket; this enumerates the same columns).
void block_entries(FermionOperator const& H, StatesClassification const& S, BlockNumber b, int64_t mat_off,
                   std::vector<int64_t>& pos, std::vector<double>& val) {
    auto const& states = S.getFockStates(b);
    int64_t n = int64_t(states.size());
    std::vector<std::pair<QuantumState, RealType>> buf;
    for(int64_t col = 0; col < n; ++col) {
        H.apply(states[col], buf);
        for(auto const& p : buf) {
            if(S.getBlockNumber(p.first) != b)
                throw std::runtime_error("Hamiltonian: block " + std::to_string(b) +
                                         " is not invariant (Hilbert space partition missing?)");
            int64_t row = int64_t(S.getInnerState(p.first));
            pos.push_back(mat_off + row * n + col);
            val.push_back(p.second);
        }
    }
}

void assemble(FermionOperator const& H, StatesClassification const& S, std::vector<BlockNumber> const& blocks,
              std::vector<int64_t> const& mat_offs, BlockStore& store) {
    std::vector<int64_t> pos;
    std::vector<double> val;
    for(std::size_t k = 0; k < blocks.size(); ++k)
        block_entries(H, S, blocks[k], mat_offs[k], pos, val);
    detail::DeviceArray<int64_t> d_pos(pos.size());
    detail::DeviceArray<double> d_val(val.size());
    d_pos.upload(pos);
    d_val.upload(val);
    detail::check(pomx_assemble(detail::gpu(), int64_t(pos.size()), d_pos.data(), d_val.data(), store.mat.data(),
                                int64_t(store.mat.size())),
                  "block assembly");
}

} // namespace

// ---- HamiltonianPart ------------------------------------------------------------------------
HamiltonianPart::HamiltonianPart(FermionOperator const& HOp, StatesClassification const& S, BlockNumber Block)
    : HOp(&HOp), S(S), Block(Block) {}

HamiltonianPart::HamiltonianPart(FermionOperator&& HOp, StatesClassification const& S, BlockNumber Block)
    : owned(std::make_shared<FermionOperator const>(std::move(HOp))), HOp(owned.get()), S(S), Block(Block) {}

InnerQuantumState HamiltonianPart::getSize() const { return S.getBlockSize(Block); }

void HamiltonianPart::prepare() {
    if(getStatus() >= Prepared)
        return;
    int64_t n = int64_t(getSize());
    store = std::make_shared<BlockStore>();
    store->mat = detail::DeviceArray<double>(std::size_t(n * n));
    store->evals = detail::DeviceArray<double>(std::size_t(n));
    mat_off = 0;
    val_off = 0;
    assemble(*HOp, S, {Block}, {0}, *store);
    host_matrix.reset();
    setStatus(Prepared);
}

void HamiltonianPart::compute() {
    if(getStatus() >= Computed)
        return;
    if(getStatus() < Prepared)
        prepare();
    int n = int(getSize());
    detail::check(pomx_eigh_batched(detail::gpu(), 1, &n, &mat_off, &val_off, store->mat.data(), store->evals.data()),
                  "eigensolve");
    Eigenvalues.assign(std::size_t(n), 0.0);
    store->evals.download(Eigenvalues.data(), std::size_t(n), std::size_t(val_off));
    host_matrix.reset();
    setStatus(Computed);
}

RealType HamiltonianPart::getEigenValue(InnerQuantumState state) const {
    if(getStatus() < Computed)
        throw StatusMismatch("HamiltonianPart is not computed yet.");
    return Eigenvalues(long(state));
}

RealVectorType const& HamiltonianPart::getEigenValues() const {
    if(getStatus() < Computed)
        throw StatusMismatch("HamiltonianPart is not computed yet.");
    return Eigenvalues;
}

RealType HamiltonianPart::getMinimumEigenvalue() const { return getEigenValues().minCoeff(); }

template <> MatrixType<false> const& HamiltonianPart::getMatrix<false>() const {
    if(getStatus() < Prepared)
        throw StatusMismatch("HamiltonianPart is not prepared yet.");
    if(!host_matrix) {
        long n = long(getSize());
        auto m = std::make_shared<MatrixType<false>>(n, n);
        store->mat.download(m->data(), std::size_t(n * n), std::size_t(mat_off));
        host_matrix = m;
    }
    return *host_matrix;
}
template <> MatrixType<true> const& HamiltonianPart::getMatrix<true>() const {
    throw std::runtime_error("Stored matrix type mismatch (real/complex)");
}

template <> VectorType<false> HamiltonianPart::getEigenState<false>(InnerQuantumState state) const {
    if(getStatus() < Computed)
        throw StatusMismatch("HamiltonianPart is not computed yet.");
    return getMatrix<false>().col(long(state));
}
template <> VectorType<true> HamiltonianPart::getEigenState<true>(InnerQuantumState) const {
    throw std::runtime_error("Stored matrix type mismatch (real/complex)");
}

bool HamiltonianPart::reduce(RealType) {
    throw std::logic_error("HamiltonianPart::reduce is not available on the GPU path");
}

// ---- Hamiltonian ----------------------------------------------------------------------------
void Hamiltonian::prepareImpl() {
    BlockNumber nb = S.getNumberOfBlocks();
    parts.clear();
    parts.reserve(std::size_t(nb));
    std::vector<BlockNumber> blocks(nb);
    std::vector<int64_t> mat_offs(nb);
    int64_t mo = 0, vo = 0;
    for(BlockNumber b = 0; b < nb; ++b) {
        parts.emplace_back(HOp, S, b);
        int64_t n = int64_t(S.getBlockSize(b));
        parts.back().mat_off = mo;
        parts.back().val_off = vo;
        blocks[b] = b;
        mat_offs[b] = mo;
        mo += n * n;
        vo += n;
    }
    store = std::make_shared<BlockStore>();
    store->mat = detail::DeviceArray<double>(std::size_t(mo));
    store->evals = detail::DeviceArray<double>(std::size_t(vo));
    assemble(HOp, S, blocks, mat_offs, *store);
    for(auto& p : parts) {
        p.store = store;
        p.setStatus(Prepared);
    }
    setStatus(Prepared);
}

void Hamiltonian::compute(MPI_Comm const&) {
    if(getStatus() >= Computed)
        return;
    if(getStatus() < Prepared)
        throw StatusMismatch("Hamiltonian is not prepared yet.");
    int nb = int(parts.size());
    std::vector<int> n(nb);
    std::vector<int64_t> mo(nb), vo(nb);
    for(int b = 0; b < nb; ++b) {
        n[b] = int(parts[b].getSize());
        mo[b] = parts[b].mat_off;
        vo[b] = parts[b].val_off;
    }
    detail::check(pomx_eigh_batched(detail::gpu(), nb, n.data(), mo.data(), vo.data(), store->mat.data(), store->evals.data()),
                  "batched eigensolve");
    std::vector<double> ev = store->evals.download();
    for(int b = 0; b < nb; ++b) {
        parts[b].Eigenvalues.assign(ev.begin() + vo[b], ev.begin() + vo[b] + n[b]);
        parts[b].host_matrix.reset();
        parts[b].setStatus(Computed);
    }
    computeGroundEnergy();
    setStatus(Computed);
}

void Hamiltonian::reduce(RealType) { throw std::logic_error("Hamiltonian::reduce is not available on the GPU path"); }

void Hamiltonian::computeGroundEnergy() {
    GroundEnergy = 0;
    bool first = true;
    for(auto const& p : parts) {
        RealType m = p.getMinimumEigenvalue();
        if(first || m < GroundEnergy)
            GroundEnergy = m;
        first = false;
    }
}

InnerQuantumState Hamiltonian::getBlockSize(BlockNumber Block) const { return parts[Block].getSize(); }

RealType Hamiltonian::getEigenValue(QuantumState state) const {
    return parts[S.getBlockNumber(state)].getEigenValue(S.getInnerState(state));
}

RealVectorType const& Hamiltonian::getEigenValues(BlockNumber Block) const { return parts[Block].getEigenValues(); }

RealVectorType Hamiltonian::getEigenValues() const {
    RealVectorType out;
    for(auto const& p : parts) {
        auto const& ev = p.getEigenValues();
        out.insert(out.end(), ev.begin(), ev.end());
    }
    return out;
}

RealType Hamiltonian::getGroundEnergy() const {
    if(getStatus() < Computed)
        throw StatusMismatch("Hamiltonian is not computed yet.");
    return GroundEnergy;
}

// ---- density matrix ---------------------------------------------------------------------------
DensityMatrixPart::DensityMatrixPart(HamiltonianPart const& H, RealType beta, RealType GroundEnergy)
    : Thermal(beta), H(H), GroundEnergy(GroundEnergy), weights(H.getSize()) {}

RealType DensityMatrixPart::getAverageEnergy() const {
    RealType s = 0;
    auto const& e = H.getEigenValues();
    for(std::size_t k = 0; k < weights.size(); ++k)
        s += weights[k] * e[k];
    return s;
}

void DensityMatrixPart::truncate(RealType Tolerance) {
    Retained = false;
    for(RealType w : weights)
        if(w > Tolerance) {
            Retained = true;
            break;
        }
}

DensityMatrix::DensityMatrix(StatesClassification const& S, Hamiltonian const& H, RealType beta)
    : Thermal(beta), S(S), H(H) {}

void DensityMatrix::prepare() {
    if(getStatus() >= Prepared)
        return;
    BlockNumber nb = S.getNumberOfBlocks();
    RealType E0 = H.getGroundEnergy();
    parts.reserve(std::size_t(nb));
    for(BlockNumber b = 0; b < nb; ++b)
        parts.emplace_back(H.getPart(b), beta, E0);
    setStatus(Prepared);
}

void DensityMatrix::compute() {
    if(getStatus() >= Computed)
        return;
    if(getStatus() < Prepared)
        prepare();
    int64_t nstates = int64_t(S.getNumberOfStates());
    d_w = detail::DeviceArray<double>(std::size_t(nstates));
    double E0 = 0;
    detail::check(pomx_boltzmann(detail::gpu(), nstates, H.deviceEigenvalues(), beta, d_w.data(), &E0, &Z),
                  "Boltzmann weights");
    std::vector<double> w = d_w.download();
    for(BlockNumber b = 0; b < BlockNumber(parts.size()); ++b) {
        int64_t off = H.eigenvalueOffset(b);
        auto& p = parts[b];
        p.weights.assign(w.begin() + off, w.begin() + off + int64_t(p.weights.size()));
        p.d_weights = d_w.data() + off;
    }
    setStatus(Computed);
}

RealType DensityMatrix::getWeight(QuantumState state) const {
    if(getStatus() < Computed)
        throw StatusMismatch("DensityMatrix is not computed yet.");
    return parts[S.getBlockNumber(state)].getWeight(S.getInnerState(state));
}

RealType DensityMatrix::getAverageEnergy() const {
    if(getStatus() < Computed)
        throw StatusMismatch("DensityMatrix is not computed yet.");
    RealType e = 0;
    for(auto const& p : parts)
        e += p.getAverageEnergy();
    return e;
}

void DensityMatrix::truncateBlocks(RealType Tolerance, bool verbose) {
    for(auto& p : parts)
        p.truncate(Tolerance);
    if(verbose) {
        QuantumState nb = 0, ns = 0;
        for(BlockNumber i = 0; i < S.getNumberOfBlocks(); ++i)
            if(isRetained(i)) {
                ++nb;
                ns += S.getBlockSize(i);
            }
        INFO("Number of blocks retained: " << nb);
        INFO("Number of states retained: " << ns);
    }
}

} // namespace Pomerol
