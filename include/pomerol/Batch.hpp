// Batched full-ED + G(i w_n) solves of many independent Hamiltonians (no reference counterpart;
// each instance is what Hamiltonian::compute + FieldOperatorContainer::computeAll +
// GreensFunction::compute + GreensFunction::operator() produce for one model).
//
// The sector structure (partition, blocks, connectivity, gather maps) depends only on which
// monomials a Hamiltonian contains, so it is computed once per distinct structure; the numeric
// Hamiltonian entries are assembled per instance.  All instances are then solved by one
// captured hipGraph on the device (pomx_batch_*).  G is evaluated from the unmerged term set
// with the reference's per-term threshold |R| > CoefficientTolerance.
#pragma once

#include "Device.hpp"
#include "HilbertSpace.hpp"
#include "LatticePresets.hpp"

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

namespace Pomerol {

class BatchSolver {
public:
    using Expr = LatticePresets::RealExpr;
    using Index = std::tuple<std::string, unsigned short, LatticePresets::spin>;

    /// G_ii(i w_n), n = 0..nfreq-1 (nfreq <= 128), of every Hamiltonian, i = index.
    BatchSolver(std::vector<Expr> Hs, std::vector<RealType> betas, Index index, int nfreq,
                RealType CoefficientTolerance = 1e-8);
    /// Parametric instances: H_k = sum_p theta[k][p] * Ops[k][p].  The sector structure is that
    /// of sum_p Ops[k][p] (every monomial present), so it does not change with theta.
    BatchSolver(std::vector<std::vector<Expr>> Ops, std::vector<std::vector<RealType>> theta,
                std::vector<RealType> betas, Index index, int nfreq, RealType CoefficientTolerance = 1e-8);
    ~BatchSolver();
    BatchSolver(BatchSolver const&) = delete;
    BatchSolver& operator=(BatchSolver const&) = delete;

    /// Host-side structure and device descriptor (outside any timed region).
    void prepare();
    /// One batched solve on the device; use_graph replays the captured hipGraph.
    void solve(bool use_graph = true);
    /// Autotune workgroup and tile sizes on this batch; returns a report of the choices.
    std::string autotune();
    std::vector<std::vector<ComplexType>> getG() const;
    /// New parameters for the next solve (same shape as the constructor's theta).
    void setParameters(std::vector<std::vector<RealType>> const& theta);
    /// Adjoint of the last solve: Gbar[k][n] = dL/dRe G_k(i w_n) + i dL/dIm G_k(i w_n);
    /// returns dL/dtheta[k][p].
    std::vector<std::vector<RealType>> gradient(std::vector<std::vector<ComplexType>> const& Gbar);
    std::vector<RealType> getGroundEnergies() const;
    std::size_t size() const { return Ops.size(); }
    std::size_t numStructures() const { return structures.size(); }
    int64_t numBlocks() const { return nblocks; }

private:
    struct Structure;
    struct Entry {      // theta[k][param] * val contributes to Hamiltonian value u
        int64_t u;
        int param;
        RealType val;
    };
    std::vector<std::vector<Expr>> Ops;
    std::vector<std::vector<RealType>> theta;
    std::vector<std::vector<Entry>> entries; // per instance, sorted by u
    std::vector<RealType> values;            // current Hamiltonian values (all instances)
    void assemble_values();
    std::vector<RealType> betas;
    Index index;
    int nfreq;
    RealType tol;
    std::map<std::string, std::shared_ptr<Structure>> structures;
    pomx_batch* batch = nullptr;
    int64_t nblocks = 0;
};

} // namespace Pomerol
