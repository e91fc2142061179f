// Vertex4 (Vertex4.cpp:21-49) over the device Lehmann sums.
#include "pomerol/Vertex4.hpp"

#include <algorithm>
#include <map>

namespace Pomerol {

Vertex4::Vertex4(TwoParticleGF const& Chi, GreensFunction const& G13, GreensFunction const& G24,
                 GreensFunction const& G14, GreensFunction const& G23)
    : Thermal(Chi.beta), Chi(Chi), G13(G13), G24(G24), G14(G14), G23(G23), Storage(*this) {}

void Vertex4::compute(long NumberOfMatsubaras) {
    if(getStatus() >= Computed)
        return;
    Storage.layout(NumberOfMatsubaras);
    auto const trip = Storage.triplets();
    auto matsubara = [this](long n) { return MatsubaraSpacing * RealType(2 * n + 1); };
    FreqVec3 f;
    f.reserve(trip.size());
    for(auto const& t : trip)
        f.emplace_back(matsubara(t[0]), matsubara(t[1]), matsubara(t[2]));
    std::vector<ComplexType> v = Chi.evaluate(f);

    // the Green's functions at every fermionic index of a disconnected term, one device sum each
    std::vector<long> idx;
    for(auto const& t : trip)
        if(t[0] == t[2] || t[1] == t[2]) {
            idx.push_back(t[0]);
            idx.push_back(t[1]);
        }
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
    std::vector<ComplexType> z;
    for(long n : idx)
        z.push_back(matsubara(n));
    std::map<GreensFunction const*, std::vector<ComplexType>> g;
    for(GreensFunction const* G : {&G13, &G24, &G14, &G23})
        if(!g.count(G))
            g[G] = G->evaluate(z);
    auto at = [&](GreensFunction const& G, long n) {
        return g[&G][std::size_t(std::lower_bound(idx.begin(), idx.end(), n) - idx.begin())];
    };
    for(std::size_t k = 0; k < trip.size(); ++k) {
        long const n1 = trip[k][0], n2 = trip[k][1], n3 = trip[k][2];
        if(n1 == n3)
            v[k] += beta * at(G13, n1) * at(G24, n2);
        if(n2 == n3)
            v[k] -= beta * at(G14, n1) * at(G23, n2);
    }
    Storage.store(v);
    setStatus(Computed);
}

ComplexType Vertex4::value(long n1, long n2, long n3) const {
    ComplexType Value = Chi(n1, n2, n3);
    if(n1 == n3)
        Value += beta * G13(n1) * G24(n2);
    if(n2 == n3)
        Value -= beta * G14(n1) * G23(n2);
    return Value;
}

ComplexType Vertex4::operator()(long n1, long n2, long n3) const { return Storage(n1, n2, n3); }

} // namespace Pomerol
