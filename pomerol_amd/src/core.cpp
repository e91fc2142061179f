// Basic utilities, the device context, Fock-space operators and the sector partition.
#include "pomerol/HilbertSpace.hpp"
#include "pomerol/Device.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <numeric>

namespace Pomerol {

// Misc.cpp:41-66
std::array<Permutation3, 6> const permutations3 = {
    {{{0, 1, 2}, 1}, {{0, 2, 1}, -1}, {{1, 0, 2}, -1}, {{1, 2, 0}, 1}, {{2, 0, 1}, 1}, {{2, 1, 0}, -1}}};

std::array<Permutation4, 24> const permutations4 = {
    {{{0, 1, 2, 3}, 1},  {{0, 1, 3, 2}, -1}, {{0, 2, 1, 3}, -1}, {{0, 2, 3, 1}, 1},  {{0, 3, 1, 2}, 1},
     {{0, 3, 2, 1}, -1}, {{1, 0, 2, 3}, -1}, {{1, 0, 3, 2}, 1},  {{1, 2, 0, 3}, 1},  {{1, 2, 3, 0}, -1},
     {{1, 3, 0, 2}, -1}, {{1, 3, 2, 0}, 1},  {{2, 0, 1, 3}, 1},  {{2, 0, 3, 1}, -1}, {{2, 1, 0, 3}, -1},
     {{2, 1, 3, 0}, 1},  {{2, 3, 0, 1}, 1},  {{2, 3, 1, 0}, -1}, {{3, 0, 1, 2}, -1}, {{3, 0, 2, 1}, 1},
     {{3, 1, 0, 2}, 1},  {{3, 1, 2, 0}, -1}, {{3, 2, 0, 1}, -1}, {{3, 2, 1, 0}, 1}}};

std::ostream& operator<<(std::ostream& os, Permutation3 const& p) {
    if(p.sign == -1)
        os << "-";
    return os << p.perm[0] + 1 << p.perm[1] + 1 << p.perm[2] + 1;
}
std::ostream& operator<<(std::ostream& os, Permutation4 const& p) {
    if(p.sign == -1)
        os << "-";
    return os << p.perm[0] + 1 << p.perm[1] + 1 << p.perm[2] + 1 << p.perm[3] + 1;
}

std::ostream& operator<<(std::ostream& os, Channel channel) {
    switch(channel) {
    case Channel::PP: return os << "PP";
    case Channel::PH: return os << "PH";
    case Channel::xPH: return os << "xPH";
    }
    return os;
}

std::size_t hash_binned_real(double x, double bin_size) {
    return std::hash<std::int64_t>()(static_cast<std::int64_t>(std::round(x / bin_size)));
}

namespace detail {

pomx_ctx* gpu() {
    thread_local std::unique_ptr<pomx_ctx, void (*)(pomx_ctx*)> ctx(nullptr, pomx_ctx_destroy);
    if(!ctx) {
        pomx_ctx* c = nullptr;
        int rc = pomx_ctx_create(0, &c);
        if(rc != POMX_OK)
            throw std::runtime_error("pomerol_amd: no usable gfx950 device (pomx_ctx_create returned " +
                                     std::to_string(rc) + ")");
        ctx.reset(c);
    }
    return ctx.get();
}

void check(int rc, char const* what) {
    if(rc != POMX_OK)
        throw std::runtime_error(std::string("pomerol_amd: ") + what + ": " + pomx_last_error(gpu()));
}

} // namespace detail

// ---- FermionOperator ---------------------------------------------------------------------
int FermionOperator::apply_monomial(std::vector<std::pair<bool, int>> const& ops, QuantumState s, QuantumState& out) {
    int sign = 1;
    for(auto it = ops.rbegin(); it != ops.rend(); ++it) {
        QuantumState bit = QuantumState(1) << it->second;
        bool occ = (s & bit) != 0;
        if(it->first == occ)
            return 0;
        if(__builtin_popcountll(s & (bit - 1)) & 1)
            sign = -sign;
        s ^= bit;
    }
    out = s;
    return sign;
}

void FermionOperator::apply(QuantumState s, std::vector<std::pair<QuantumState, RealType>>& out) const {
    out.clear();
    for(auto const& t : terms) {
        QuantumState f;
        int sg = apply_monomial(t.ops, s, f);
        if(sg)
            out.emplace_back(f, sg * t.coeff);
    }
    std::sort(out.begin(), out.end(), [](auto const& a, auto const& b) { return a.first < b.first; });
    std::size_t w = 0;
    for(std::size_t r = 0; r < out.size(); ++r) {
        if(w > 0 && out[w - 1].first == out[r].first)
            out[w - 1].second += out[r].second;
        else
            out[w++] = out[r];
    }
    out.resize(w);
}

// ---- sector partition --------------------------------------------------------------------
namespace {
struct UnionFind {
    std::vector<QuantumState> parent;
    explicit UnionFind(QuantumState n) : parent(n) { std::iota(parent.begin(), parent.end(), QuantumState(0)); }
    QuantumState find(QuantumState x) {
        while(parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    }
    bool unite(QuantumState a, QuantumState b) {
        a = find(a);
        b = find(b);
        if(a == b)
            return false;
        if(b < a)
            std::swap(a, b);
        parent[b] = a;
        return true;
    }
};
} // namespace

SpacePartition make_space_partition(int n_modes, FermionOperator const& H) {
    QuantumState dim = QuantumState(1) << n_modes;
    UnionFind uf(dim);
    // Phase I: states connected by a non-zero matrix element of H share a sector
    std::vector<std::pair<QuantumState, RealType>> buf;
    for(QuantumState s = 0; s < dim; ++s) {
        H.apply(s, buf);
        for(auto const& p : buf)
            if(p.first != s && p.second != 0.0)
                uf.unite(s, p.first);
    }
    // Phase II: every c_i and c^+_i must map a sector into a single sector
    std::vector<std::vector<std::pair<bool, int>>> ops;
    for(int i = 0; i < n_modes; ++i) {
        ops.push_back({{true, i}});
        ops.push_back({{false, i}});
    }
    std::vector<int64_t> first(dim);
    bool changed = true;
    while(changed) {
        changed = false;
        for(auto const& op : ops) {
            std::fill(first.begin(), first.end(), -1);
            for(QuantumState s = 0; s < dim; ++s) {
                QuantumState t;
                if(!FermionOperator::apply_monomial(op, s, t))
                    continue;
                QuantumState r = uf.find(s);
                if(first[r] < 0)
                    first[r] = int64_t(t);
                else if(uf.unite(QuantumState(first[r]), t))
                    changed = true;
            }
        }
    }
    SpacePartition P;
    P.state_block.assign(dim, INVALID_BLOCK_NUMBER);
    P.inner.assign(dim, 0);
    std::vector<BlockNumber> root_block(dim, INVALID_BLOCK_NUMBER);
    for(QuantumState s = 0; s < dim; ++s) {
        QuantumState r = uf.find(s);
        if(root_block[r] == INVALID_BLOCK_NUMBER) {
            root_block[r] = BlockNumber(P.blocks.size());
            P.blocks.emplace_back();
        }
        BlockNumber b = root_block[r];
        P.state_block[s] = b;
        P.inner[s] = P.blocks[b].size();
        P.blocks[b].push_back(s);
    }
    return P;
}

std::set<std::pair<BlockNumber, BlockNumber>> SpacePartition::find_connections(FermionOperator const& op) const {
    std::set<std::pair<BlockNumber, BlockNumber>> conns;
    if(!op.is_monomial())
        throw std::runtime_error("find_connections: monomial operator expected");
    auto const& ops = op.terms[0].ops;
    for(BlockNumber b = 0; b < BlockNumber(blocks.size()); ++b)
        for(QuantumState s : blocks[b]) {
            QuantumState t;
            if(FermionOperator::apply_monomial(ops, s, t))
                conns.emplace(b, state_block[t]);
        }
    return conns;
}

// ---- StatesClassification ------------------------------------------------------------------
void StatesClassification::initSingleBlock(QuantumState Dim) {
    StateBlockIndex.assign(Dim, 0);
    StateInnerIndex.resize(Dim);
    std::iota(StateInnerIndex.begin(), StateInnerIndex.end(), InnerQuantumState(0));
    StatesContainer.emplace_back(Dim);
    std::iota(StatesContainer.back().begin(), StatesContainer.back().end(), QuantumState(0));
}

void StatesClassification::initMultipleBlocks(SpacePartition const& p) {
    StatesContainer = p.blocks;
    StateBlockIndex = p.state_block;
    StateInnerIndex = p.inner;
}

void StatesClassification::checkComputed() const {
    if(getStatus() < Computed)
        throw StatusMismatch("StatesClassification is not computed yet.");
}

InnerQuantumState StatesClassification::getBlockSize(BlockNumber in) const {
    checkComputed();
    return getFockStates(in).size();
}

std::vector<QuantumState> const& StatesClassification::getFockStates(BlockNumber in) const {
    checkComputed();
    if(in < 0 || in >= BlockNumber(StatesContainer.size()))
        throw std::runtime_error("Wrong block " + std::to_string(in));
    return StatesContainer[in];
}

QuantumState StatesClassification::getFockState(BlockNumber in, InnerQuantumState m) const {
    auto const& b = getFockStates(in);
    if(m < b.size())
        return b[m];
    throw std::runtime_error("Wrong inner state " + std::to_string(m));
}

BlockNumber StatesClassification::getBlockNumber(QuantumState in) const {
    checkComputed();
    if(in >= StateBlockIndex.size())
        throw std::runtime_error("Wrong state " + std::to_string(in));
    return StateBlockIndex[in];
}

InnerQuantumState StatesClassification::getInnerState(QuantumState in) const {
    checkComputed();
    if(in >= StateBlockIndex.size())
        throw std::runtime_error("Wrong state " + std::to_string(in));
    return StateInnerIndex[in];
}

} // namespace Pomerol
