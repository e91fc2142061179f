// Irreducible two-particle vertex part Gamma4 = chi - chi0 and its Matsubara container.
//
// References: Vertex4.hpp/.cpp (Vertex4::value: chi(n1,n2,n3) + beta G13(n1) G24(n2) [n1 == n3]
// - beta G14(n1) G23(n2) [n2 == n3]; compute(N) fills the container), MatsubaraContainers.hpp
// (MatsubaraContainer4: one (2N - |W+1|)^2 block per bosonic index W = n1 + n2 in
// [-2N, 2N - 2], values cached, uncached indices fall back to Source.value).
//
// Vertex4::compute(N) fills the whole container at once: the two-particle GF at every stored
// triplet in one device evaluation per part (TwoParticleGF::evaluate) and the Green's functions
// at every needed fermionic index in one device Lehmann sum each; the disconnected parts are
// then added elementwise.
#pragma once

#include "TwoParticleGF.hpp"

#include <cstdlib>
#include <vector>

namespace Pomerol {

template <typename SourceObject> class MatsubaraContainer4 {
    SourceObject const& Source;
    long NumberOfMatsubaras = 0;
    // per bosonic index: row-major (nu = n1, nu' = n3) block and the fermionic offset of its rows
    std::vector<std::vector<ComplexType>> Values;
    std::vector<long> Size, FermionicIndexOffset;

public:
    explicit MatsubaraContainer4(SourceObject const& Source) : Source(Source) {}

    /// Storage layout of N: per bosonic index W = n1 + n2 (stored as W + 2N), the block size
    /// 2N - |W + 1| and the smallest stored n1 (= n3).
    void layout(long N) {
        NumberOfMatsubaras = N;
        Values.assign(std::size_t(N > 0 ? 4 * N - 1 : 0), {});
        Size.assign(Values.size(), 0);
        FermionicIndexOffset.assign(Values.size(), 0);
        for(long V = 0; V < long(Values.size()); ++V) {
            long const W = V - 2 * N;
            Size[std::size_t(V)] = 2 * N - std::abs(W + 1);
            FermionicIndexOffset[std::size_t(V)] = (W < 0 ? 0 : W + 1) - N;
            Values[std::size_t(V)].assign(std::size_t(Size[std::size_t(V)] * Size[std::size_t(V)]), 0.0);
        }
    }
    /// Stored triplets (n1, n2, n3) in storage order.
    std::vector<std::array<long, 3>> triplets() const {
        std::vector<std::array<long, 3>> t;
        for(long V = 0; V < long(Values.size()); ++V) {
            long const W = V - 2 * NumberOfMatsubaras, n = Size[std::size_t(V)], o = FermionicIndexOffset[std::size_t(V)];
            for(long a = 0; a < n; ++a)
                for(long b = 0; b < n; ++b)
                    t.push_back({a + o, W - (a + o), b + o});
        }
        return t;
    }
    /// Store values given in the order of triplets().
    void store(std::vector<ComplexType> const& v) {
        std::size_t k = 0;
        for(auto& blk : Values)
            for(auto& x : blk)
                x = v.at(k++);
    }
    /// MatsubaraContainer4::fill: every stored value from Source.value (one call per triplet).
    void fill(long N) {
        layout(N);
        std::vector<ComplexType> v;
        for(auto const& t : triplets())
            v.push_back(Source.value(t[0], t[1], t[2]));
        store(v);
    }
    ComplexType operator()(long n1, long n2, long n3) const {
        long const V = n1 + n2 + 2 * NumberOfMatsubaras;
        if(V >= 0 && V < long(Values.size())) {
            long const a = n1 - FermionicIndexOffset[std::size_t(V)], b = n3 - FermionicIndexOffset[std::size_t(V)];
            long const n = Size[std::size_t(V)];
            if(a >= 0 && a < n && b >= 0 && b < n)
                return Values[std::size_t(V)][std::size_t(a * n + b)];
        }
        return Source.value(n1, n2, n3);
    }
    long getNumberOfMatsubaras() const { return NumberOfMatsubaras; }
};

class Vertex4 : public Thermal, public ComputableObject {
    TwoParticleGF const& Chi;
    GreensFunction const& G13;
    GreensFunction const& G24;
    GreensFunction const& G14;
    GreensFunction const& G23;
    mutable MatsubaraContainer4<Vertex4> Storage;

    friend class MatsubaraContainer4<Vertex4>;

public:
    Vertex4(TwoParticleGF const& Chi, GreensFunction const& G13, GreensFunction const& G24, GreensFunction const& G14,
            GreensFunction const& G23);

    /// Fill the container for |n1|, |n2|, |n3| below NumberOfMatsubaras (Vertex4.cpp:33-38).
    void compute(long NumberOfMatsubaras = 0);
    /// Stored value, or value() for triplets outside the container.
    ComplexType operator()(long MatsubaraNumber1, long MatsubaraNumber2, long MatsubaraNumber3) const;
    /// chi(n1,n2,n3) + beta G13(n1) G24(n2) [n1 == n3] - beta G14(n1) G23(n2) [n2 == n3]
    /// (Vertex4.cpp:40-49), evaluated on the device point by point.
    ComplexType value(long MatsubaraNumber1, long MatsubaraNumber2, long MatsubaraNumber3) const;
};

} // namespace Pomerol
