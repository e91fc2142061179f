// Host side of the batched solver: sector structure per distinct model structure, per-instance
// Hamiltonian entries, and the descriptor of the device batch.
#include "pomerol/Batch.hpp"

#include <algorithm>
#include <sstream>

namespace Pomerol {

struct BatchSolver::Structure {
    IndexClassification<std::string, unsigned short, LatticePresets::spin> info;
    SpacePartition partition;
    int mode = -1;                                  // ParticleIndex of the G index
    std::vector<std::pair<BlockNumber, BlockNumber>> conns; // c^+ (from, to), ascending
    std::vector<std::vector<int>> gather;           // per connection
};

namespace {

std::string structure_key(BatchSolver::Expr const& H) {
    std::ostringstream os;
    for(auto const& t : H.get_monomials()) {
        for(auto const& o : t.first)
            os << (o.dag ? '+' : '-') << std::get<0>(o.index) << ',' << std::get<1>(o.index) << ','
               << int(std::get<2>(o.index)) << ';';
        os << '|';
    }
    return os.str();
}

} // namespace

BatchSolver::BatchSolver(std::vector<Expr> Hs_, std::vector<RealType> betas_, Index index_, int nfreq_, RealType tol_)
    : betas(std::move(betas_)), index(std::move(index_)), nfreq(nfreq_), tol(tol_) {
    for(auto& H : Hs_) {
        Ops.push_back({std::move(H)});
        theta.push_back({1.0});
    }
    if(Ops.size() != betas.size() || Ops.empty())
        throw std::runtime_error("BatchSolver: one beta per Hamiltonian expected");
    if(nfreq < 1 || nfreq > 128)
        throw std::runtime_error("BatchSolver: 1 <= nfreq <= 128");
}

BatchSolver::BatchSolver(std::vector<std::vector<Expr>> Ops_, std::vector<std::vector<RealType>> theta_,
                         std::vector<RealType> betas_, Index index_, int nfreq_, RealType tol_)
    : Ops(std::move(Ops_)), theta(std::move(theta_)), betas(std::move(betas_)), index(std::move(index_)), nfreq(nfreq_),
      tol(tol_) {
    if(Ops.size() != betas.size() || Ops.empty() || theta.size() != Ops.size())
        throw std::runtime_error("BatchSolver: one operator list, parameter list and beta per instance expected");
    for(std::size_t k = 0; k < Ops.size(); ++k)
        if(Ops[k].empty() || Ops[k].size() != theta[k].size())
            throw std::runtime_error("BatchSolver: instance " + std::to_string(k) + " needs one parameter per operator");
    if(nfreq < 1 || nfreq > 128)
        throw std::runtime_error("BatchSolver: 1 <= nfreq <= 128");
}

BatchSolver::~BatchSolver() {
    if(batch)
        pomx_batch_destroy(detail::gpu(), batch);
}

void BatchSolver::assemble_values() {
    std::fill(values.begin(), values.end(), 0.0);
    for(std::size_t k = 0; k < Ops.size(); ++k)
        for(auto const& e : entries[k]) // entries of one value are contiguous, parameters ascending
            values[std::size_t(e.u)] += theta[k][std::size_t(e.param)] * e.val;
}

void BatchSolver::prepare() {
    if(batch)
        return;
    std::vector<int> n_block;
    std::vector<int64_t> mat_off, state_off, h_pos, inst_state_off, inst_nstates, job_from, job_to;
    std::vector<pomx_rot_job> jobs;
    std::vector<int> job_inst, src;
    int64_t mat_len = 0, nstates = 0, ops_len = 0;
    std::vector<std::pair<QuantumState, RealType>> buf;
    entries.assign(Ops.size(), {});
    for(std::size_t k = 0; k < Ops.size(); ++k) {
        // structure of sum_p c_p Ops[k][p]; the weights c_p keep monomials of different
        // parameters from cancelling
        Expr Hs = Ops[k][0];
        for(std::size_t p = 1; p < Ops[k].size(); ++p)
            Hs += (1.0 + 0.6180339887498949 * double(p)) * Ops[k][p];
        std::string key = structure_key(Hs);
        auto it = structures.find(key);
        if(it == structures.end()) {
            auto st = std::make_shared<Structure>();
            st->info = MakeIndexClassification(Hs);
            if(st->info.getIndexSize() > 30)
                throw std::runtime_error("BatchSolver: more than 30 modes");
            st->partition = make_space_partition(int(st->info.getIndexSize()), st->info.toOperator(Hs));
            st->mode = int(st->info.getIndex(index));
            FermionOperator cdag;
            cdag.terms.push_back(FermionOperator::Term{1.0, {{true, st->mode}}});
            auto c = st->partition.find_connections(cdag);
            st->conns.assign(c.begin(), c.end());
            for(auto const& cn : st->conns) {
                auto const& fs = st->partition.blocks[cn.first];
                std::vector<int> g(st->partition.blocks[cn.second].size(), 0);
                for(std::size_t a = 0; a < fs.size(); ++a) {
                    QuantumState t;
                    int sg = FermionOperator::apply_monomial(cdag.terms[0].ops, fs[a], t);
                    if(sg)
                        g[st->partition.inner[t]] = sg * int(a + 1);
                }
                st->gather.push_back(std::move(g));
            }
            it = structures.emplace(key, st).first;
        }
        Structure const& st = *it->second;
        auto const& P = st.partition;
        inst_state_off.push_back(nstates);
        inst_nstates.push_back(int64_t(P.dim()));
        std::vector<int64_t> bmat(P.blocks.size()), bstate(P.blocks.size());
        for(std::size_t b = 0; b < P.blocks.size(); ++b) {
            int64_t n = int64_t(P.blocks[b].size());
            bmat[b] = mat_len;
            bstate[b] = nstates;
            n_block.push_back(int(n));
            mat_off.push_back(mat_len);
            state_off.push_back(nstates);
            mat_len += n * n;
            nstates += n;
        }
        // entries of every parameter operator (HamiltonianPart.cpp:56-76), grouped by position
        struct Raw {
            int64_t pos;
            int param;
            RealType val;
        };
        std::vector<Raw> raw;
        for(std::size_t p = 0; p < Ops[k].size(); ++p) {
            FermionOperator Hop = st.info.toOperator(Ops[k][p]);
            for(std::size_t b = 0; b < P.blocks.size(); ++b) {
                int64_t n = int64_t(P.blocks[b].size());
                for(int64_t col = 0; col < n; ++col) {
                    Hop.apply(P.blocks[b][std::size_t(col)], buf);
                    for(auto const& e : buf) {
                        if(P.state_block[e.first] != BlockNumber(b))
                            throw std::runtime_error("BatchSolver: Hamiltonian is not block diagonal");
                        raw.push_back(Raw{bmat[b] + int64_t(P.inner[e.first]) * n + col, int(p), e.second});
                    }
                }
            }
        }
        std::stable_sort(raw.begin(), raw.end(), [](Raw const& x, Raw const& y) { return x.pos < y.pos; });
        for(std::size_t r = 0; r < raw.size(); ++r) {
            if(r == 0 || raw[r].pos != raw[r - 1].pos)
                h_pos.push_back(raw[r].pos);
            entries[k].push_back(Entry{int64_t(h_pos.size()) - 1, raw[r].param, raw[r].val});
        }
        for(std::size_t q = 0; q < st.conns.size(); ++q) {
            BlockNumber from = st.conns[q].first, to = st.conns[q].second;
            pomx_rot_job j;
            j.n_to = int(P.blocks[to].size());
            j.n_from = int(P.blocks[from].size());
            j.u_to_off = bmat[to];
            j.u_from_off = bmat[from];
            j.src_off = int64_t(src.size());
            j.out_off = ops_len;
            src.insert(src.end(), st.gather[q].begin(), st.gather[q].end());
            ops_len += int64_t(j.n_to) * j.n_from;
            jobs.push_back(j);
            job_inst.push_back(int(k));
            job_from.push_back(bstate[from]);
            job_to.push_back(bstate[to]);
        }
    }
    values.assign(h_pos.size(), 0.0);
    assemble_values();
    nblocks = int64_t(n_block.size());
    pomx_batch_desc d;
    d.ninst = int(Ops.size());
    d.nfreq = nfreq;
    d.nblocks = int(n_block.size());
    d.block_n = n_block.data();
    d.block_mat_off = mat_off.data();
    d.block_state_off = state_off.data();
    d.mat_len = mat_len;
    d.nstates = nstates;
    d.nnz = int64_t(h_pos.size());
    d.h_pos = h_pos.data();
    d.h_val = values.data();
    d.inst_beta = betas.data();
    d.inst_state_off = inst_state_off.data();
    d.inst_nstates = inst_nstates.data();
    d.njobs = int(jobs.size());
    d.jobs = jobs.data();
    d.job_inst = job_inst.data();
    d.job_from_state_off = job_from.data();
    d.job_to_state_off = job_to.data();
    d.src = src.data();
    d.src_len = int64_t(src.size());
    d.ops_len = ops_len;
    d.coeff_tol = tol;
    detail::check(pomx_batch_create(detail::gpu(), &d, &batch), "batch creation");
}

void BatchSolver::setParameters(std::vector<std::vector<RealType>> const& th) {
    if(th.size() != Ops.size())
        throw std::runtime_error("BatchSolver::setParameters: one parameter list per instance expected");
    for(std::size_t k = 0; k < th.size(); ++k)
        if(th[k].size() != Ops[k].size())
            throw std::runtime_error("BatchSolver::setParameters: instance " + std::to_string(k) +
                                     " needs one parameter per operator");
    theta = th;
    if(!batch)
        return; // prepare() assembles from theta
    assemble_values();
    detail::check(pomx_batch_set_values(detail::gpu(), batch, values.data()), "batch parameters");
}

std::vector<std::vector<RealType>> BatchSolver::gradient(std::vector<std::vector<ComplexType>> const& Gbar) {
    if(!batch)
        throw std::runtime_error("BatchSolver::gradient: not solved");
    if(Gbar.size() != Ops.size())
        throw std::runtime_error("BatchSolver::gradient: one Gbar row per instance expected");
    std::vector<double> g(Ops.size() * std::size_t(nfreq) * 2);
    for(std::size_t k = 0; k < Ops.size(); ++k) {
        if(Gbar[k].size() != std::size_t(nfreq))
            throw std::runtime_error("BatchSolver::gradient: nfreq values per instance expected");
        for(int n = 0; n < nfreq; ++n) {
            g[(k * nfreq + n) * 2] = Gbar[k][std::size_t(n)].real();
            g[(k * nfreq + n) * 2 + 1] = Gbar[k][std::size_t(n)].imag();
        }
    }
    std::vector<double> dval(values.size());
    detail::check(pomx_batch_adjoint(detail::gpu(), batch, g.data(), dval.data()), "batch adjoint");
    std::vector<std::vector<RealType>> out(Ops.size());
    for(std::size_t k = 0; k < Ops.size(); ++k) {
        out[k].assign(Ops[k].size(), 0.0);
        for(auto const& e : entries[k])
            out[k][std::size_t(e.param)] += e.val * dval[std::size_t(e.u)];
    }
    return out;
}

void BatchSolver::solve(bool use_graph) {
    prepare();
    detail::check(pomx_batch_solve(detail::gpu(), batch, use_graph ? 1 : 0), "batched solve");
}

std::string BatchSolver::autotune() {
    prepare();
    char rep[512] = {0};
    detail::check(pomx_batch_tune(detail::gpu(), batch, rep, sizeof(rep)), "batch autotune");
    return rep;
}

std::vector<std::vector<ComplexType>> BatchSolver::getG() const {
    if(!batch)
        throw std::runtime_error("BatchSolver: not solved");
    std::vector<double> h(Ops.size() * std::size_t(nfreq) * 2);
    detail::check(pomx_batch_read(detail::gpu(), batch, h.data(), nullptr), "batch results");
    std::vector<std::vector<ComplexType>> out(Ops.size(), std::vector<ComplexType>(std::size_t(nfreq)));
    for(std::size_t k = 0; k < Ops.size(); ++k)
        for(int n = 0; n < nfreq; ++n)
            out[k][std::size_t(n)] = ComplexType(h[(k * nfreq + n) * 2], h[(k * nfreq + n) * 2 + 1]);
    return out;
}

std::vector<RealType> BatchSolver::getGroundEnergies() const {
    if(!batch)
        throw std::runtime_error("BatchSolver: not solved");
    std::vector<RealType> e(Ops.size());
    detail::check(pomx_batch_read(detail::gpu(), batch, nullptr, e.data()), "batch results");
    return e;
}

} // namespace Pomerol
