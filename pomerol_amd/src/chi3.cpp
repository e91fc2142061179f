// Three-point susceptibility, its parts and its container on top of the device term pipeline
// (pomx_chi3_part_compute / pomx_chi3_terms_eval).
#include "pomerol/ThreePointSusceptibility.hpp"

#include <algorithm>
#include <stdexcept>

namespace Pomerol {

namespace {

void free_chi3_terms(pomx_chi3_terms* t) {
    if(t) {
        if(t->data)
            pomx_free(detail::gpu(), t->data);
        delete t;
    }
}

// Device copy of (z1, z2) pairs: 4 doubles per pair (Re z1, Im z1, Re z2, Im z2).
detail::DeviceArray<double> upload_pairs(FreqVec2 const& freqs) {
    std::vector<double> h(4 * std::max<std::size_t>(freqs.size(), 1), 0.0);
    for(std::size_t w = 0; w < freqs.size(); ++w) {
        ComplexType z1 = std::get<0>(freqs[w]), z2 = std::get<1>(freqs[w]);
        h[4 * w] = z1.real();
        h[4 * w + 1] = z1.imag();
        h[4 * w + 2] = z2.real();
        h[4 * w + 3] = z2.imag();
    }
    detail::DeviceArray<double> d(h.size());
    d.upload(h);
    return d;
}

std::vector<ComplexType> download_values(detail::DeviceArray<double> const& acc, std::size_t n) {
    std::vector<ComplexType> out(n, 0.0);
    if(n == 0)
        return out;
    auto h = acc.download();
    for(std::size_t w = 0; w < n; ++w)
        out[w] = ComplexType(h[2 * w], h[2 * w + 1]);
    return out;
}

} // namespace

// ---- ThreePointSusceptibilityPart (ThreePointSusceptibilityPart.cpp:193-359) -------------------
ThreePointSusceptibilityPart::ThreePointSusceptibilityPart(
    MonomialOperatorPart const& F1, MonomialOperatorPart const& F2, MonomialOperatorPart const& B1,
    MonomialOperatorPart const& B2, HamiltonianPart const& Hpart1, HamiltonianPart const& Hpart2,
    HamiltonianPart const& Hpart3, DensityMatrixPart const& DMpart1, DensityMatrixPart const& DMpart2,
    DensityMatrixPart const& DMpart3, Channel channel, bool SwappedFermionOps, RealType PoleResolution,
    RealType CoefficientTolerance)
    : Thermal(DMpart1.beta), F1(F1), F2(F2), B1(B1), B2(B2), Hpart1(Hpart1), Hpart2(Hpart2), Hpart3(Hpart3),
      DMpart1(DMpart1), DMpart2(DMpart2), DMpart3(DMpart3), channel(channel), SwappedFermionOps(SwappedFermionOps),
      PoleResolution(PoleResolution), CoefficientTolerance(CoefficientTolerance) {}

pomx_chi3_part ThreePointSusceptibilityPart::devicePart() const {
    pomx_chi3_part d;
    d.n1 = int(Hpart1.getSize());
    d.n2 = int(Hpart2.getSize());
    d.n3 = int(Hpart3.getSize());
    d.n4 = int(B1.cols()); // block between B1 and B2
    d.Fa = F1.deviceView();
    d.Fb = F2.deviceView();
    d.B1 = B1.deviceView();
    d.B2 = B2.deviceView();
    d.E1 = Hpart1.deviceEigenvalues();
    d.E2 = Hpart2.deviceEigenvalues();
    d.E3 = Hpart3.deviceEigenvalues();
    d.w1 = DMpart1.deviceWeights();
    d.w2 = DMpart2.deviceWeights();
    d.w3 = DMpart3.deviceWeights();
    d.channel = int(channel);
    d.swapped = SwappedFermionOps ? 1 : 0;
    return d;
}

void ThreePointSusceptibilityPart::run(int nfreq, double const* d_freq, double* d_acc, bool keep) {
    pomx_chi3_part d = devicePart();
    std::shared_ptr<pomx_chi3_terms> t;
    if(keep)
        t = std::shared_ptr<pomx_chi3_terms>(new pomx_chi3_terms{{0, 0, 0}, 1, nullptr}, free_chi3_terms);
    detail::check(pomx_chi3_part_compute(detail::gpu(), &d, beta, PoleResolution, CoefficientTolerance, nfreq, d_freq,
                                         d_acc, t.get(), counts, &raw_terms),
                  "three-point susceptibility part");
    terms = t;
    setStatus(keep ? Computed : Constructed);
}

void ThreePointSusceptibilityPart::compute() {
    if(getStatus() >= Computed)
        return;
    run(0, nullptr, nullptr, true);
}

void ThreePointSusceptibilityPart::clear() {
    terms.reset();
    setStatus(Constructed);
}

ComplexType ThreePointSusceptibilityPart::operator()(ComplexType z1, ComplexType z2) const {
    if(getStatus() != Computed || !terms)
        throw StatusMismatch("3PSusceptibilityPart: Calling operator() on uncomputed container. Did you purge all the "
                             "terms when called compute()?");
    auto df = upload_pairs(FreqVec2{FreqTuple2(z1, z2)});
    detail::DeviceArray<double> acc(2);
    acc.zero();
    detail::check(pomx_chi3_terms_eval(detail::gpu(), terms.get(), PoleResolution, 1, df.data(), acc.data()),
                  "three-point susceptibility evaluation");
    return download_values(acc, 1)[0];
}

ComplexType ThreePointSusceptibilityPart::operator()(long n1, long n2) const {
    return (*this)(MatsubaraSpacing * RealType(2 * n1 + 1), MatsubaraSpacing * RealType(2 * n2 + 1));
}

std::vector<double> ThreePointSusceptibilityPart::hostTerms() const {
    std::vector<double> h;
    if(!terms)
        return h;
    h.resize(std::size_t(3 * counts[0] + 3 * counts[1] + 2 * counts[2]));
    if(!h.empty())
        detail::check(pomx_d2h(detail::gpu(), h.data(), terms->data, sizeof(double) * h.size()), "term copy");
    return h;
}

std::vector<ThreePointSusceptibilityPart::NonResonantFFTerm> ThreePointSusceptibilityPart::getNonResonantFFTerms() const {
    std::vector<NonResonantFFTerm> out;
    auto h = hostTerms();
    std::size_t n = std::size_t(counts[0]);
    for(std::size_t k = 0; k < n && !h.empty(); ++k) {
        NonResonantFFTerm t;
        t.Coeff = h[k];
        t.Poles = {h[n + k], h[2 * n + k]};
        out.push_back(t);
    }
    return out;
}

std::vector<ThreePointSusceptibilityPart::NonResonantFBTerm> ThreePointSusceptibilityPart::getNonResonantFBTerms() const {
    std::vector<NonResonantFBTerm> out;
    auto h = hostTerms();
    std::size_t b = std::size_t(3 * counts[0]), n = std::size_t(counts[1]);
    for(std::size_t k = 0; k < n && !h.empty(); ++k) {
        NonResonantFBTerm t;
        t.Coeff = h[b + k];
        t.P1 = h[b + n + k];
        t.P12 = h[b + 2 * n + k];
        t.xi = channel == PP ? -1 : 1;
        out.push_back(t);
    }
    return out;
}

std::vector<ThreePointSusceptibilityPart::ResonantTerm> ThreePointSusceptibilityPart::getResonantTerms() const {
    std::vector<ResonantTerm> out;
    auto h = hostTerms();
    std::size_t b = std::size_t(3 * counts[0] + 3 * counts[1]), n = std::size_t(counts[2]);
    for(std::size_t k = 0; k < n && !h.empty(); ++k) {
        ResonantTerm t;
        t.Coeff = h[b + k];
        t.P = h[b + n + k];
        t.xi = channel == PP ? -1 : 1;
        out.push_back(t);
    }
    return out;
}

// ---- ThreePointSusceptibility (ThreePointSusceptibility.cpp:25-243) ----------------------------
ThreePointSusceptibility::ThreePointSusceptibility(Channel channel, StatesClassification const& S, Hamiltonian const& H,
                                                   CreationOperator const& CX1, AnnihilationOperator const& C2,
                                                   CreationOperator const& CX3, AnnihilationOperator const& C4,
                                                   DensityMatrix const& DM)
    : Thermal(DM.beta), channel(channel), S(S), H(H), CX1(CX1), C2(C2), CX3(CX3), C4(C4), DM(DM) {}

CreationOperator const& ThreePointSusceptibility::getF1() const { return CX1; }

MonomialOperator const& ThreePointSusceptibility::getF2() const {
    switch(channel) {
    case PP: return CX3;
    case PH: return C2;
    case xPH: return C4;
    }
    throw std::runtime_error("ThreePointSusceptibility: Wrong channel");
}
MonomialOperator const& ThreePointSusceptibility::getB1() const {
    switch(channel) {
    case PP: return C2;
    case PH: return CX3;
    case xPH: return CX3;
    }
    throw std::runtime_error("ThreePointSusceptibility: Wrong channel");
}
MonomialOperator const& ThreePointSusceptibility::getB2() const {
    switch(channel) {
    case PP: return C4;
    case PH: return C4;
    case xPH: return C2;
    }
    throw std::runtime_error("ThreePointSusceptibility: Wrong channel");
}

void ThreePointSusceptibility::prepare() {
    if(getStatus() >= Prepared)
        return;
    CreationOperator const& F1 = getF1();
    MonomialOperator const& F2 = getF2();
    MonomialOperator const& B1 = getB1();
    MonomialOperator const& B2 = getB2();
    for(auto const& rl : B2.getBlockMapping().right) { // ascending right block of B2
        BlockNumber B2right = rl.first, B2left = rl.second;
        BlockNumber B1left = B1.getLeftIndex(B2left);
        if(B1left == INVALID_BLOCK_NUMBER)
            continue;
        // <B2right|F1|F1right><F2left|F2|B1left><B1left|B1|B2left><B2left|B2|B2right>
        BlockNumber F1right = F1.getRightIndex(B2right);
        BlockNumber F2left = F2.getLeftIndex(B1left);
        if(F1right == F2left && F1right != INVALID_BLOCK_NUMBER && DM.isRetained(B2right) && DM.isRetained(F1right) &&
           DM.isRetained(B1left))
            parts.emplace_back(F1.getPartFromLeftIndex(B2right), F2.getPartFromLeftIndex(F2left),
                               B1.getPartFromLeftIndex(B1left), B2.getPartFromLeftIndex(B2left), H.getPart(B2right),
                               H.getPart(F1right), H.getPart(B1left), DM.getPart(B2right), DM.getPart(F1right),
                               DM.getPart(B1left), channel, false, PoleResolution, CoefficientTolerance);
        // <B2right|F2|F2right><F1left|F1|B1left><B1left|B1|B2left><B2left|B2|B2right>
        BlockNumber F2right = F2.getRightIndex(B2right);
        BlockNumber F1left = F1.getLeftIndex(B1left);
        if(F2right == F1left && F2right != INVALID_BLOCK_NUMBER && DM.isRetained(B2right) && DM.isRetained(F2right) &&
           DM.isRetained(B1left))
            parts.emplace_back(F2.getPartFromLeftIndex(B2right), F1.getPartFromLeftIndex(F1left),
                               B1.getPartFromLeftIndex(B1left), B2.getPartFromLeftIndex(B2left), H.getPart(B2right),
                               H.getPart(F2right), H.getPart(B1left), DM.getPart(B2right), DM.getPart(F2right),
                               DM.getPart(B1left), channel, true, PoleResolution, CoefficientTolerance);
    }
    Vanishing = parts.empty();
    setStatus(Prepared);
}

std::vector<ComplexType> ThreePointSusceptibility::compute(bool clear, FreqVec2 const& freqs, MPI_Comm const&) {
    if(getStatus() < Prepared)
        throw StatusMismatch("ThreePointSusceptibility is not prepared yet.");
    std::vector<ComplexType> m_data;
    if(getStatus() >= Computed)
        return m_data;
    m_data.assign(freqs.size(), 0.0);
    if(!Vanishing) {
        int nf = int(freqs.size());
        auto df = upload_pairs(freqs);
        detail::DeviceArray<double> acc(std::size_t(2 * std::max(nf, 1)));
        acc.zero();
        for(auto& p : parts)
            p.run(nf, nf ? df.data() : nullptr, nf ? acc.data() : nullptr, !clear);
        m_data = download_values(acc, freqs.size());
    }
    setStatus(Computed);
    return m_data;
}

ParticleIndex ThreePointSusceptibility::getIndex(std::size_t Position) const {
    switch(Position) {
    case 0: return CX1.getIndex();
    case 1: return C2.getIndex();
    case 2: return CX3.getIndex();
    case 3: return C4.getIndex();
    default: throw std::runtime_error("ThreePointSusceptibility: Could not get operator index");
    }
}

ComplexType ThreePointSusceptibility::operator()(ComplexType z1, ComplexType z2) const {
    if(Vanishing)
        return 0;
    ComplexType s = 0;
    for(auto const& p : parts)
        s += p(z1, z2);
    return s;
}

ComplexType ThreePointSusceptibility::operator()(long n1, long n2) const {
    return (*this)(MatsubaraSpacing * RealType(2 * n1 + 1), MatsubaraSpacing * RealType(2 * n2 + 1));
}

std::vector<ComplexType> ThreePointSusceptibility::evaluate(FreqVec2 const& freqs) const {
    std::vector<ComplexType> out(freqs.size(), 0.0);
    if(Vanishing || freqs.empty())
        return out;
    int nf = int(freqs.size());
    auto df = upload_pairs(freqs);
    detail::DeviceArray<double> acc(std::size_t(2 * nf));
    acc.zero();
    for(auto const& p : parts) {
        if(p.getStatus() != ThreePointSusceptibilityPart::Computed || !p.terms)
            throw StatusMismatch("ThreePointSusceptibility::evaluate: the terms of a part were not kept by compute()");
        detail::check(pomx_chi3_terms_eval(detail::gpu(), p.terms.get(), p.PoleResolution, nf, df.data(), acc.data()),
                      "three-point susceptibility evaluation");
    }
    return download_values(acc, freqs.size());
}

// ---- ThreePointSusceptibilityContainer (ThreePointSusceptibilityContainer.cpp:19-48,
//      IndexContainer4.hpp:171-273) -------------------------------------------------------------
void ThreePointSusceptibilityContainer::fill(std::set<IndexCombination4> Indices) {
    ElementsMap.clear();
    NonTrivialElements.clear();
    if(Indices.empty())
        for(ParticleIndex i1 = 0; i1 < NumIndices; ++i1)
            for(ParticleIndex i2 = i1; i2 < NumIndices; ++i2)
                for(ParticleIndex i3 = 0; i3 < NumIndices; ++i3)
                    for(ParticleIndex i4 = i3; i4 < NumIndices; ++i4)
                        Indices.emplace(i1, i2, i3, i4);
    for(auto const& ic : Indices)
        if(!isInContainer(ic))
            create(ic);
}

ElementWithPermFreq<ThreePointSusceptibility>& ThreePointSusceptibilityContainer::create(IndexCombination4 const& I) {
    auto p = std::make_shared<ThreePointSusceptibility>(
        channel, S, H, Operators.getCreationOperator(I.Index1), Operators.getAnnihilationOperator(I.Index2),
        Operators.getCreationOperator(I.Index3), Operators.getAnnihilationOperator(I.Index4), DM);
    auto it = ElementsMap.emplace(I, ElementWithPermFreq<ThreePointSusceptibility>(p, permutations4[0])).first;
    NonTrivialElements.emplace(I, p);
    bool sameC = I.Index1 == I.Index2, sameCX = I.Index3 == I.Index4;
    if(!sameC) {
        IndexCombination4 J(I.Index2, I.Index1, I.Index3, I.Index4);
        if(!isInContainer(J))
            ElementsMap.emplace(J, ElementWithPermFreq<ThreePointSusceptibility>(p, permutations4[6]));
    }
    if(!sameCX) {
        IndexCombination4 J(I.Index1, I.Index2, I.Index4, I.Index3);
        if(!isInContainer(J))
            ElementsMap.emplace(J, ElementWithPermFreq<ThreePointSusceptibility>(p, permutations4[1]));
    }
    if(!sameC && !sameCX) {
        IndexCombination4 J(I.Index2, I.Index1, I.Index4, I.Index3);
        if(!isInContainer(J))
            ElementsMap.emplace(J, ElementWithPermFreq<ThreePointSusceptibility>(p, permutations4[7]));
    }
    return it->second;
}

void ThreePointSusceptibilityContainer::prepareAll(std::set<IndexCombination4> const& Indices) {
    fill(Indices);
    for(auto& kv : ElementsMap) {
        ThreePointSusceptibility& chi3 = kv.second;
        chi3.PoleResolution = PoleResolution;
        chi3.CoefficientTolerance = CoefficientTolerance;
        chi3.prepare();
    }
}

// As the reference, every entry of the map (permutation images included) calls compute() on its
// element: the first entry of an element computes it, later entries of the same element receive
// the empty list compute() returns for an element already computed.
std::map<IndexCombination4, std::vector<ComplexType>>
ThreePointSusceptibilityContainer::computeAll(bool clearTerms, FreqVec2 const& freqs, MPI_Comm const& comm) {
    std::map<IndexCombination4, std::vector<ComplexType>> out;
    for(auto& kv : ElementsMap) {
        ThreePointSusceptibility& chi3 = kv.second;
        out.emplace(kv.first, chi3.compute(clearTerms, freqs, comm));
    }
    return out;
}

ElementWithPermFreq<ThreePointSusceptibility>&
ThreePointSusceptibilityContainer::operator()(IndexCombination4 const& Indices) {
    auto it = ElementsMap.find(Indices);
    if(it == ElementsMap.end())
        return create(Indices);
    return it->second;
}

} // namespace Pomerol
