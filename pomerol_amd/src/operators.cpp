// Monomial operators in the eigenbasis: block connectivity, batched device rotation, adjoints.
#include "pomerol/MonomialOperator.hpp"

#include <algorithm>

namespace Pomerol {

// ---- MonomialOperatorPart -------------------------------------------------------------------
MonomialOperatorPart::MonomialOperatorPart(FermionOperator const& MOp, StatesClassification const& S,
                                           HamiltonianPart const& HFrom, HamiltonianPart const& HTo)
    : MOp(&MOp), S(S), HFrom(HFrom), HTo(HTo) {}

MonomialOperatorPart::MonomialOperatorPart(FermionOperator&& MOp, StatesClassification const& S,
                                           HamiltonianPart const& HFrom, HamiltonianPart const& HTo)
    : owned(std::make_shared<FermionOperator const>(std::move(MOp))), MOp(owned.get()), S(S), HFrom(HFrom), HTo(HTo) {}

namespace {

struct PendingJob {
    MonomialOperatorPart* part;
    std::vector<int> src;
    int n_to, n_from;
    int64_t u_to_off, u_from_off;
};

// Signed row gather of M U_R: for every target state t, +-(1 + source inner index) or 0.
std::vector<int> gather_map(FermionOperator const& M, StatesClassification const& S, BlockNumber from, BlockNumber to) {
    auto const& fs = S.getFockStates(from);
    std::vector<int> src(S.getBlockSize(to), 0);
    auto const& ops = M.terms.at(0).ops;
    double coeff = M.terms.at(0).coeff;
    if(coeff != 1.0 && coeff != -1.0)
        throw std::runtime_error("MonomialOperator: coefficient must be +-1");
    for(std::size_t a = 0; a < fs.size(); ++a) {
        QuantumState t;
        int sg = FermionOperator::apply_monomial(ops, fs[a], t);
        if(!sg)
            continue;
        if(S.getBlockNumber(t) != to)
            throw std::runtime_error("MonomialOperator: operator does not map block to block");
        int& slot = src[S.getInnerState(t)];
        if(slot != 0)
            throw std::runtime_error("MonomialOperator: operator is not injective on Fock states");
        slot = (sg * (coeff < 0 ? -1 : 1)) * int(a + 1);
    }
    return src;
}

} // namespace

void MonomialOperator::computeBatch(std::vector<MonomialOperator*> const& ops, RealType Tolerance) {
    std::vector<MonomialOperatorPart*> todo;
    Hamiltonian const* H0 = nullptr;
    for(auto* op : ops) {
        op->checkPrepared();
        if(op->getStatus() >= Computed)
            continue;
        if(H0 && &op->H != H0)
            throw std::runtime_error("MonomialOperator::computeBatch: operators of different Hamiltonians");
        H0 = &op->H;
        for(auto& p : op->parts)
            if(p.getStatus() < Computed)
                todo.push_back(&p);
    }
    if(!todo.empty()) {
        std::vector<pomx_rot_job> jobs;
        std::vector<int> src;
        int64_t out_total = 0;
        for(auto* p : todo) {
            pomx_rot_job j;
            j.n_to = int(p->HTo.getSize());
            j.n_from = int(p->HFrom.getSize());
            j.u_to_off = p->HTo.deviceMatrix() - H0->getPart(0).deviceMatrix();
            j.u_from_off = p->HFrom.deviceMatrix() - H0->getPart(0).deviceMatrix();
            j.src_off = int64_t(src.size());
            j.out_off = out_total;
            auto s = gather_map(*p->MOp, p->S, p->HFrom.getBlockNumber(), p->HTo.getBlockNumber());
            src.insert(src.end(), s.begin(), s.end());
            out_total += int64_t(j.n_to) * j.n_from;
            jobs.push_back(j);
        }
        auto store = std::make_shared<detail::DeviceArray<double>>(std::size_t(out_total));
        detail::DeviceArray<int> d_src(src.size());
        d_src.upload(src);
        detail::check(pomx_rotate_batched(detail::gpu(), int(jobs.size()), jobs.data(), H0->getPart(0).deviceMatrix(),
                                          d_src.data(), store->data(), Tolerance * 1e-12),
                      "operator rotation");
        for(std::size_t k = 0; k < todo.size(); ++k) {
            auto* p = todo[k];
            p->store = store;
            p->off = jobs[k].out_off;
            p->rs = jobs[k].n_from;
            p->cs = 1;
            p->host.reset();
            p->setStatus(Computed);
        }
    }
    for(auto* op : ops)
        op->setStatus(Computed);
}

void MonomialOperatorPart::compute(RealType Tolerance) {
    if(getStatus() >= Computed)
        return;
    pomx_rot_job j;
    j.n_to = int(HTo.getSize());
    j.n_from = int(HFrom.getSize());
    // the two blocks may live in different arenas: stage both eigenvector blocks in one buffer
    int64_t nt2 = int64_t(j.n_to) * j.n_to, nf2 = int64_t(j.n_from) * j.n_from;
    detail::DeviceArray<double> U(std::size_t(nt2 + nf2));
    auto const& UT = HTo.getMatrix<false>();
    auto const& UF = HFrom.getMatrix<false>();
    U.upload(UT.data(), std::size_t(nt2), 0);
    U.upload(UF.data(), std::size_t(nf2), std::size_t(nt2));
    j.u_to_off = 0;
    j.u_from_off = nt2;
    j.src_off = 0;
    j.out_off = 0;
    auto s = gather_map(*MOp, S, HFrom.getBlockNumber(), HTo.getBlockNumber());
    detail::DeviceArray<int> d_src(s.size());
    d_src.upload(s);
    store = std::make_shared<detail::DeviceArray<double>>(std::size_t(int64_t(j.n_to) * j.n_from));
    detail::check(pomx_rotate_batched(detail::gpu(), 1, &j, U.data(), d_src.data(), store->data(), Tolerance * 1e-12),
                  "operator rotation");
    off = 0;
    rs = j.n_from;
    cs = 1;
    host.reset();
    setStatus(Computed);
}

void MonomialOperatorPart::setFromAdjoint(MonomialOperatorPart const& part) {
    if(getStatus() >= Computed)
        return;
    if(getLeftIndex() != part.getRightIndex() || getRightIndex() != part.getLeftIndex())
        throw std::runtime_error("setFromAdjoint: block mismatch");
    store = part.store;
    off = part.off;
    rs = part.cs;
    cs = part.rs;
    host.reset();
    setStatus(Computed);
}

template <> MatrixType<false> const& MonomialOperatorPart::getRowMajorValue<false>() const {
    if(getStatus() < Computed)
        throw StatusMismatch("MonomialOperatorPart is not computed yet.");
    if(!host) {
        long r = rows(), c = cols();
        // storage of the (possibly transposed) view: a dense block of r*c entries at off
        std::vector<double> raw(std::size_t(r * c));
        store->download(raw.data(), raw.size(), std::size_t(off));
        auto m = std::make_shared<MatrixType<false>>(r, c);
        for(long i = 0; i < r; ++i)
            for(long k = 0; k < c; ++k)
                (*m)(i, k) = raw[std::size_t(i * rs + k * cs)];
        host = m;
    }
    return *host;
}
template <> MatrixType<true> const& MonomialOperatorPart::getRowMajorValue<true>() const {
    throw std::runtime_error("Stored matrix type mismatch (real/complex)");
}

// ---- MonomialOperator -----------------------------------------------------------------------
void MonomialOperator::prepareImpl(SpacePartition const* partition) {
    if(!partition) { // unpartitioned space: a single block
        mapPartsFromRight.emplace(0, 0);
        mapPartsFromLeft.emplace(0, 0);
        LeftRightBlocks.insert(BlockMapping(0, 0));
        parts.emplace_back(MOp, S, H.getPart(0), H.getPart(0));
        setStatus(Prepared);
        return;
    }
    auto conns = partition->find_connections(MOp);
    parts.reserve(conns.size());
    for(auto const& c : conns) {
        mapPartsFromRight.emplace(c.first, parts.size());
        mapPartsFromLeft.emplace(c.second, parts.size());
        LeftRightBlocks.insert(BlockMapping(c.second, c.first));
        parts.emplace_back(MOp, S, H.getPart(c.first), H.getPart(c.second));
    }
    setStatus(Prepared);
}

void MonomialOperator::checkPrepared() const {
    if(getStatus() < Prepared)
        throw StatusMismatch("MonomialOperator is not prepared yet.");
}

void MonomialOperator::compute(RealType Tolerance, MPI_Comm const&) {
    checkPrepared();
    if(getStatus() >= Computed)
        return;
    computeBatch({this}, Tolerance);
}

BlocksBimap const& MonomialOperator::getBlockMapping() const {
    checkPrepared();
    return LeftRightBlocks;
}

MonomialOperatorPart& MonomialOperator::getPartFromLeftIndex(BlockNumber in) {
    checkPrepared();
    return parts.at(mapPartsFromLeft.at(in));
}
MonomialOperatorPart const& MonomialOperator::getPartFromLeftIndex(BlockNumber in) const {
    checkPrepared();
    return parts.at(mapPartsFromLeft.at(in));
}
MonomialOperatorPart& MonomialOperator::getPartFromRightIndex(BlockNumber out) {
    checkPrepared();
    return parts.at(mapPartsFromRight.at(out));
}
MonomialOperatorPart const& MonomialOperator::getPartFromRightIndex(BlockNumber out) const {
    checkPrepared();
    return parts.at(mapPartsFromRight.at(out));
}

BlockNumber MonomialOperator::getRightIndex(BlockNumber LeftIndex) const {
    checkPrepared();
    auto it = LeftRightBlocks.left.find(LeftIndex);
    return it != LeftRightBlocks.left.end() ? it->second : INVALID_BLOCK_NUMBER;
}

BlockNumber MonomialOperator::getLeftIndex(BlockNumber RightIndex) const {
    checkPrepared();
    auto it = LeftRightBlocks.right.find(RightIndex);
    return it != LeftRightBlocks.right.end() ? it->second : INVALID_BLOCK_NUMBER;
}

// ---- FieldOperatorContainer -------------------------------------------------------------------
void FieldOperatorContainer::computeAll(RealType Tolerance) {
    std::vector<MonomialOperator*> cdags;
    for(auto& kv : mapCreationOperators)
        cdags.push_back(&kv.second);
    MonomialOperator::computeBatch(cdags, Tolerance);
    for(auto& kv : mapCreationOperators) {
        auto& cdag = kv.second;
        auto& c = mapAnnihilationOperators.find(kv.first)->second;
        for(auto const& m : cdag.getBlockMapping().right) { // (from, to) of c^+
            auto& cPart = c.getPartFromRightIndex(m.second);
            auto& cdagPart = cdag.getPartFromRightIndex(m.first);
            cPart.setFromAdjoint(cdagPart);
        }
        c.setStatus(ComputableObject::Computed);
    }
}

CreationOperator const& FieldOperatorContainer::getCreationOperator(ParticleIndex in) const {
    auto it = mapCreationOperators.find(in);
    if(it == mapCreationOperators.end())
        throw std::runtime_error("FieldOperatorContainer: wrong index " + std::to_string(in));
    return it->second;
}

AnnihilationOperator const& FieldOperatorContainer::getAnnihilationOperator(ParticleIndex in) const {
    auto it = mapAnnihilationOperators.find(in);
    if(it == mapAnnihilationOperators.end())
        throw std::runtime_error("FieldOperatorContainer: wrong index " + std::to_string(in));
    return it->second;
}

// ---- EnsembleAverage --------------------------------------------------------------------------
EnsembleAverage::EnsembleAverage(MonomialOperator const& A, DensityMatrix const& DM) : Thermal(DM.beta), A(A), DM(DM) {}

void EnsembleAverage::compute() {
    if(getStatus() >= Prepared)
        return;
    Result = 0;
    for(auto const& m : A.getBlockMapping().left) {
        if(m.first != m.second || !DM.isRetained(m.first))
            continue;
        auto const& M = A.getPartFromLeftIndex(m.first).getRowMajorValue<false>();
        auto const& W = DM.getPart(m.first);
        ComplexType part = 0;
        for(long k = 0; k < M.rows(); ++k)
            part += M(k, k) * W.getWeight(InnerQuantumState(k));
        Result += part;
    }
    setStatus(Prepared);
}

} // namespace Pomerol
