// Green's function, susceptibility and two-particle GF on top of the device term pipelines.
#include "pomerol/TwoParticleGF.hpp"
#include "pomerol/Tangent.hpp"

#include <algorithm>
#include <limits>
#include <numeric>

namespace Pomerol {

namespace {

// Device copy of a list of complex points, interleaved (re, im).
detail::DeviceArray<double> upload_points(std::vector<ComplexType> const& z) {
    std::vector<double> h(2 * z.size());
    for(std::size_t k = 0; k < z.size(); ++k) {
        h[2 * k] = z[k].real();
        h[2 * k + 1] = z[k].imag();
    }
    detail::DeviceArray<double> d(h.size());
    d.upload(h);
    return d;
}

std::vector<ComplexType> pair_eval(int kind, double const* res, double const* pole, int64_t nterms,
                                   std::vector<ComplexType> const& z) {
    std::vector<ComplexType> out(z.size(), 0.0);
    if(nterms == 0 || z.empty())
        return out;
    auto dz = upload_points(z);
    detail::DeviceArray<double> acc(2 * z.size());
    acc.zero();
    detail::check(pomx_pair_eval(detail::gpu(), kind, nterms, res, pole, int(z.size()), dz.data(), acc.data()),
                  "Lehmann sum");
    std::vector<double> h = acc.download();
    for(std::size_t k = 0; k < z.size(); ++k)
        out[k] = ComplexType(h[2 * k], h[2 * k + 1]);
    return out;
}

std::vector<ComplexType> pair_eval_tau(int kind, double const* res, double const* pole, int64_t nterms,
                                       std::vector<RealType> const& tau, RealType beta) {
    std::vector<ComplexType> out(tau.size(), 0.0);
    for(RealType t : tau)
        if(!(t >= 0 && t <= beta))
            throw std::runtime_error("imaginary time outside [0, beta]");
    if(nterms == 0 || tau.empty())
        return out;
    detail::DeviceArray<double> dt(tau.size()), acc(2 * tau.size());
    dt.upload(std::vector<double>(tau.begin(), tau.end()));
    acc.zero();
    detail::check(pomx_pair_eval_tau(detail::gpu(), kind, nterms, res, pole, int(tau.size()), dt.data(), beta, acc.data()),
                  "imaginary-time Lehmann sum");
    std::vector<double> h = acc.download();
    for(std::size_t k = 0; k < tau.size(); ++k)
        out[k] = ComplexType(h[2 * k], h[2 * k + 1]);
    return out;
}

std::vector<PoleTerm> download_terms(double const* res, double const* pole, int64_t n) {
    std::vector<double> r(static_cast<std::size_t>(n)), p(static_cast<std::size_t>(n));
    if(n) {
        detail::check(pomx_d2h(detail::gpu(), r.data(), res, sizeof(double) * std::size_t(n)), "term copy");
        detail::check(pomx_d2h(detail::gpu(), p.data(), pole, sizeof(double) * std::size_t(n)), "term copy");
    }
    std::vector<PoleTerm> t(static_cast<std::size_t>(n));
    for(std::size_t k = 0; k < t.size(); ++k)
        t[k] = PoleTerm{r[k], p[k]};
    return t;
}

} // namespace

// ---- GreensFunctionPart / GreensFunction ------------------------------------------------------
GreensFunctionPart::GreensFunctionPart(MonomialOperatorPart const& F1, MonomialOperatorPart const& F2,
                                       HamiltonianPart const& HpartInner, HamiltonianPart const& HpartOuter,
                                       DensityMatrixPart const& DMpartInner, DensityMatrixPart const& DMpartOuter,
                                       RealType PoleResolution, RealType CoefficientTolerance)
    : Thermal(DMpartInner.beta), HpartInner(HpartInner), HpartOuter(HpartOuter), DMpartInner(DMpartInner),
      DMpartOuter(DMpartOuter), F1(F1), F2(F2), PoleResolution(PoleResolution),
      CoefficientTolerance(CoefficientTolerance) {}

pomx_pair_part GreensFunctionPart::devicePart() const {
    pomx_pair_part p;
    p.n_out = int(HpartOuter.getSize());
    p.n_in = int(HpartInner.getSize());
    p.A = F1.deviceView();
    p.B = F2.deviceView();
    p.E_out = HpartOuter.deviceEigenvalues();
    p.E_in = HpartInner.deviceEigenvalues();
    p.w_out = DMpartOuter.deviceWeights();
    p.w_in = DMpartInner.deviceWeights();
    return p;
}

ComplexType GreensFunctionPart::operator()(ComplexType z) const {
    return pair_eval(POMX_KIND_GF, d_res, d_pole, n_terms, {z})[0];
}

std::vector<PoleTerm> GreensFunctionPart::getTerms() const { return download_terms(d_res, d_pole, n_terms); }

ComplexType GreensFunctionPart::of_tau(RealType tau) const {
    return pair_eval_tau(POMX_KIND_GF, d_res, d_pole, n_terms, {tau}, beta)[0];
}

GreensFunction::GreensFunction(StatesClassification const& S, Hamiltonian const& H, FieldOperator const& F1,
                               FieldOperator const& F2, DensityMatrix const& DM)
    : Thermal(DM.beta), S(S), H(H), F1(F1), F2(F2), DM(DM) {}

GreensFunction::GreensFunction(GreensFunction const& GF)
    : Thermal(GF.beta), ComputableObject(GF), S(GF.S), H(GF.H), F1(GF.F1), F2(GF.F2), DM(GF.DM), Vanishing(GF.Vanishing),
      parts(GF.parts), d_res(GF.d_res), d_pole(GF.d_pole), n_terms(GF.n_terms), PoleResolution(GF.PoleResolution),
      CoefficientTolerance(GF.CoefficientTolerance) {}

// GreensFunction.cpp:41-86: merge-join of F1's left blocks with F2's right blocks
void GreensFunction::prepare() {
    if(getStatus() >= Prepared)
        return;
    auto const& m1 = F1.getBlockMapping().left;  // left -> right of F1
    auto const& m2 = F2.getBlockMapping().right; // right -> left of F2
    auto i1 = m1.begin();
    auto i2 = m2.begin();
    while(i1 != m1.end() && i2 != m2.end()) {
        BlockNumber F1left = i1->first, F1right = i1->second;
        BlockNumber F2right = i2->first, F2left = i2->second;
        if(F1left == F2right && F1right == F2left && (DM.isRetained(F1left) || DM.isRetained(F1right)))
            parts.emplace_back(F1.getPartFromLeftIndex(F1left), F2.getPartFromRightIndex(F2right), H.getPart(F1right),
                               H.getPart(F1left), DM.getPart(F1right), DM.getPart(F1left), PoleResolution,
                               CoefficientTolerance);
        unsigned long a = (unsigned long)F1left, b = (unsigned long)F2right;
        if(a <= b)
            ++i1;
        if(a >= b)
            ++i2;
    }
    Vanishing = parts.empty();
    setStatus(Prepared);
}

void GreensFunction::compute() {
    if(getStatus() >= Computed)
        return;
    if(getStatus() < Prepared)
        prepare();
    if(!parts.empty()) {
        std::vector<pomx_pair_part> dp;
        for(auto const& p : parts)
            dp.push_back(p.devicePart());
        std::vector<int64_t> cnt(parts.size());
        double *r = nullptr, *p = nullptr;
        // F1 and F2 are single fermion operators: their rotated blocks have elements of magnitude
        // at most 1, so pairs of states with negligible weights are skipped unread
        detail::check(pomx_pair_terms(detail::gpu(), POMX_KIND_GF | POMX_PAIR_UNIT_ELEMENTS, int(parts.size()), dp.data(),
                                      PoleResolution,
                                      CoefficientTolerance, &r, &p, cnt.data(), nullptr),
                      "Green's function terms");
        n_terms = std::accumulate(cnt.begin(), cnt.end(), int64_t(0));
        d_res = std::make_shared<detail::DeviceArray<double>>(detail::DeviceArray<double>::adopt(r, std::size_t(n_terms)));
        d_pole = std::make_shared<detail::DeviceArray<double>>(detail::DeviceArray<double>::adopt(p, std::size_t(n_terms)));
        int64_t off = 0;
        for(std::size_t k = 0; k < parts.size(); ++k) {
            parts[k].d_res = r + off;
            parts[k].d_pole = p + off;
            parts[k].n_terms = cnt[k];
            off += cnt[k];
        }
    }
    setStatus(Computed);
}

ParticleIndex GreensFunction::getIndex(std::size_t Position) const {
    switch(Position) {
    case 0: return F1.getIndex();
    case 1: return F2.getIndex();
    default: throw std::runtime_error("GreensFunction: wrong operator position");
    }
}

std::vector<ComplexType> GreensFunction::evaluate(std::vector<ComplexType> const& z) const {
    if(Vanishing)
        return std::vector<ComplexType>(z.size(), 0.0);
    if(getStatus() < Computed)
        throw StatusMismatch("GreensFunction is not computed yet.");
    return pair_eval(POMX_KIND_GF, d_res->data(), d_pole->data(), n_terms, z);
}

ComplexType GreensFunction::operator()(ComplexType z) const { return evaluate({z})[0]; }

std::vector<ComplexType> GreensFunction::evaluateTau(std::vector<RealType> const& tau) const {
    if(Vanishing)
        return std::vector<ComplexType>(tau.size(), 0.0);
    if(getStatus() < Computed)
        throw StatusMismatch("GreensFunction is not computed yet.");
    return pair_eval_tau(POMX_KIND_GF, d_res->data(), d_pole->data(), n_terms, tau, beta);
}

ComplexType GreensFunction::of_tau(RealType tau) const { return evaluateTau({tau})[0]; }

// ---- GFContainer ------------------------------------------------------------------------------
void GFContainer::prepareAll(std::set<IndexCombination2> const& Indices) {
    std::set<IndexCombination2> II = Indices;
    if(II.empty())
        for(ParticleIndex i = 0; i < NumIndices; ++i)
            for(ParticleIndex j = 0; j < NumIndices; ++j)
                II.emplace(i, j);
    for(auto const& ic : II)
        (*this)(ic);
    for(auto& kv : ElementsMap) {
        kv.second->PoleResolution = PoleResolution;
        kv.second->CoefficientTolerance = CoefficientTolerance;
        kv.second->prepare();
    }
}

void GFContainer::computeAll() {
    for(auto& kv : ElementsMap)
        kv.second->compute();
}

GreensFunction& GFContainer::operator()(IndexCombination2 const& Indices) {
    auto it = ElementsMap.find(Indices);
    if(it == ElementsMap.end()) {
        FieldOperator const& F2 = Anomalous ? static_cast<FieldOperator const&>(Operators.getAnnihilationOperator(Indices.Index2))
                                            : static_cast<FieldOperator const&>(Operators.getCreationOperator(Indices.Index2));
        auto g = std::make_shared<GreensFunction>(S, H, Operators.getAnnihilationOperator(Indices.Index1), F2, DM);
        it = ElementsMap.emplace(Indices, g).first;
    }
    return *it->second;
}

// ---- Susceptibility -----------------------------------------------------------------------------
SusceptibilityPart::SusceptibilityPart(MonomialOperatorPart const& A, MonomialOperatorPart const& B,
                                       HamiltonianPart const& HpartInner, HamiltonianPart const& HpartOuter,
                                       DensityMatrixPart const& DMpartInner, DensityMatrixPart const& DMpartOuter,
                                       RealType PoleResolution, RealType CoefficientTolerance)
    : Thermal(DMpartInner.beta), HpartInner(HpartInner), HpartOuter(HpartOuter), DMpartInner(DMpartInner),
      DMpartOuter(DMpartOuter), A(A), B(B), PoleResolution(PoleResolution), CoefficientTolerance(CoefficientTolerance) {}

pomx_pair_part SusceptibilityPart::devicePart() const {
    pomx_pair_part p;
    p.n_out = int(HpartOuter.getSize());
    p.n_in = int(HpartInner.getSize());
    p.A = A.deviceView();
    p.B = B.deviceView();
    p.E_out = HpartOuter.deviceEigenvalues();
    p.E_in = HpartInner.deviceEigenvalues();
    p.w_out = DMpartOuter.deviceWeights();
    p.w_in = DMpartInner.deviceWeights();
    return p;
}

ComplexType SusceptibilityPart::operator()(ComplexType z) const {
    ComplexType zp = std::abs(z) <= std::numeric_limits<RealType>::epsilon() ? ZeroPoleWeight * beta : 0;
    return pair_eval(POMX_KIND_CHI, d_res, d_pole, n_terms, {z})[0] + zp;
}

std::vector<PoleTerm> SusceptibilityPart::getTerms() const { return download_terms(d_res, d_pole, n_terms); }

ComplexType SusceptibilityPart::of_tau(RealType tau) const {
    return pair_eval_tau(POMX_KIND_CHI, d_res, d_pole, n_terms, {tau}, beta)[0] + ZeroPoleWeight;
}

Susceptibility::Susceptibility(StatesClassification const& S, Hamiltonian const& H, MonomialOperator const& A,
                               MonomialOperator const& B, DensityMatrix const& DM)
    : Thermal(DM.beta), S(S), H(H), A(A), B(B), DM(DM) {}

Susceptibility::Susceptibility(Susceptibility const& Chi)
    : Thermal(Chi.beta), ComputableObject(Chi), S(Chi.S), H(Chi.H), A(Chi.A), B(Chi.B), DM(Chi.DM),
      Vanishing(Chi.Vanishing), parts(Chi.parts), d_res(Chi.d_res), d_pole(Chi.d_pole), n_terms(Chi.n_terms),
      zero_pole_total(Chi.zero_pole_total), SubtractDisconnected(Chi.SubtractDisconnected), ave_A(Chi.ave_A),
      ave_B(Chi.ave_B), PoleResolution(Chi.PoleResolution), CoefficientTolerance(Chi.CoefficientTolerance) {}

// Susceptibility.cpp:42-88
void Susceptibility::prepare() {
    if(getStatus() >= Prepared)
        return;
    auto const& ma = A.getBlockMapping().left;
    auto const& mb = B.getBlockMapping().right;
    auto ia = ma.begin();
    auto ib = mb.begin();
    while(ia != ma.end() && ib != mb.end()) {
        BlockNumber Aleft = ia->first, Aright = ia->second;
        BlockNumber Bright = ib->first, Bleft = ib->second;
        if(Aleft == Bright && Aright == Bleft && (DM.isRetained(Aleft) || DM.isRetained(Aright)))
            parts.emplace_back(A.getPartFromLeftIndex(Aleft), B.getPartFromRightIndex(Bright), H.getPart(Aright),
                               H.getPart(Aleft), DM.getPart(Aright), DM.getPart(Aleft), PoleResolution,
                               CoefficientTolerance);
        unsigned long a = (unsigned long)Aleft, b = (unsigned long)Bright;
        if(a <= b)
            ++ia;
        if(a >= b)
            ++ib;
    }
    Vanishing = parts.empty();
    setStatus(Prepared);
}

void Susceptibility::compute() {
    if(getStatus() >= Computed)
        return;
    if(getStatus() < Prepared)
        prepare();
    if(!parts.empty()) {
        std::vector<pomx_pair_part> dp;
        for(auto const& p : parts)
            dp.push_back(p.devicePart());
        std::vector<int64_t> cnt(parts.size());
        std::vector<double> zp(parts.size());
        double *r = nullptr, *p = nullptr;
        detail::check(pomx_pair_terms(detail::gpu(), POMX_KIND_CHI, int(parts.size()), dp.data(), PoleResolution,
                                      CoefficientTolerance, &r, &p, cnt.data(), zp.data()),
                      "susceptibility terms");
        n_terms = std::accumulate(cnt.begin(), cnt.end(), int64_t(0));
        d_res = std::make_shared<detail::DeviceArray<double>>(detail::DeviceArray<double>::adopt(r, std::size_t(n_terms)));
        d_pole = std::make_shared<detail::DeviceArray<double>>(detail::DeviceArray<double>::adopt(p, std::size_t(n_terms)));
        int64_t off = 0;
        zero_pole_total = 0;
        for(std::size_t k = 0; k < parts.size(); ++k) {
            parts[k].d_res = r + off;
            parts[k].d_pole = p + off;
            parts[k].n_terms = cnt[k];
            parts[k].ZeroPoleWeight = zp[k];
            zero_pole_total += zp[k];
            off += cnt[k];
        }
    }
    setStatus(Computed);
}

void Susceptibility::subtractDisconnected() {
    EnsembleAverage EA_A(A, DM);
    EnsembleAverage EA_B(B, DM);
    subtractDisconnected(EA_A, EA_B);
}
void Susceptibility::subtractDisconnected(ComplexType a, ComplexType b) {
    SubtractDisconnected = true;
    ave_A = a;
    ave_B = b;
}
void Susceptibility::subtractDisconnected(EnsembleAverage& EA_A, EnsembleAverage& EA_B) {
    EA_A.compute();
    EA_B.compute();
    subtractDisconnected(EA_A(), EA_B());
}

std::vector<ComplexType> Susceptibility::evaluate(std::vector<ComplexType> const& z) const {
    std::vector<ComplexType> v(z.size(), 0.0);
    if(!Vanishing) {
        if(getStatus() < Computed)
            throw StatusMismatch("Susceptibility is not computed yet.");
        v = pair_eval(POMX_KIND_CHI, d_res->data(), d_pole->data(), n_terms, z);
        for(std::size_t k = 0; k < z.size(); ++k)
            if(std::abs(z[k]) <= std::numeric_limits<RealType>::epsilon())
                v[k] += zero_pole_total * beta;
    }
    for(std::size_t k = 0; k < z.size(); ++k)
        if(SubtractDisconnected && std::abs(z[k]) <= std::numeric_limits<double>::epsilon())
            v[k] -= ave_A * ave_B * beta;
    return v;
}

ComplexType Susceptibility::operator()(ComplexType z) const { return evaluate({z})[0]; }

std::vector<ComplexType> Susceptibility::evaluateTau(std::vector<RealType> const& tau) const {
    std::vector<ComplexType> v(tau.size(), 0.0);
    if(!Vanishing) {
        if(getStatus() < Computed)
            throw StatusMismatch("Susceptibility is not computed yet.");
        v = pair_eval_tau(POMX_KIND_CHI, d_res->data(), d_pole->data(), n_terms, tau, beta);
        for(auto& x : v)
            x += zero_pole_total;
    }
    if(SubtractDisconnected)
        for(auto& x : v)
            x -= ave_A * ave_B;
    return v;
}

ComplexType Susceptibility::of_tau(RealType tau) const { return evaluateTau({tau})[0]; }

// ---- TwoParticleGFPart -------------------------------------------------------------------------
TwoParticleGFPart::TwoParticleGFPart(MonomialOperatorPart const& O1, MonomialOperatorPart const& O2,
                                     MonomialOperatorPart const& O3, MonomialOperatorPart const& CX4,
                                     HamiltonianPart const& Hpart1, HamiltonianPart const& Hpart2,
                                     HamiltonianPart const& Hpart3, HamiltonianPart const& Hpart4,
                                     DensityMatrixPart const& DMpart1, DensityMatrixPart const& DMpart2,
                                     DensityMatrixPart const& DMpart3, DensityMatrixPart const& DMpart4,
                                     Permutation3 Permutation, RealType PoleResolution, RealType CoefficientTolerance)
    : Thermal(DMpart1.beta), O1(O1), O2(O2), O3(O3), CX4(CX4), Hpart1(Hpart1), Hpart2(Hpart2), Hpart3(Hpart3),
      Hpart4(Hpart4), DMpart1(DMpart1), DMpart2(DMpart2), DMpart3(DMpart3), DMpart4(DMpart4), Permutation(Permutation),
      PoleResolution(PoleResolution), CoefficientTolerance(CoefficientTolerance) {
    for(int p = 0; p < 6; ++p)
        if(permutations3[p] == Permutation)
            perm_index = p;
}

pomx_g2_part TwoParticleGFPart::devicePart() const {
    pomx_g2_part d;
    d.n1 = int(Hpart1.getSize());
    d.n2 = int(Hpart2.getSize());
    d.n3 = int(Hpart3.getSize());
    d.n4 = int(Hpart4.getSize());
    d.O1 = O1.deviceView();
    d.O2 = O2.deviceView();
    d.O3 = O3.deviceView();
    d.CX4 = CX4.deviceView();
    d.E1 = Hpart1.deviceEigenvalues();
    d.E2 = Hpart2.deviceEigenvalues();
    d.E3 = Hpart3.deviceEigenvalues();
    d.E4 = Hpart4.deviceEigenvalues();
    d.w1 = DMpart1.deviceWeights();
    d.w2 = DMpart2.deviceWeights();
    d.w3 = DMpart3.deviceWeights();
    d.w4 = DMpart4.deviceWeights();
    d.perm = perm_index;
    d.sign = Permutation.sign;
    return d;
}

namespace {
void free_terms(pomx_g2_terms* t) {
    if(t) {
        if(t->data)
            pomx_free(detail::gpu(), t->data);
        delete t;
    }
}
} // namespace

void TwoParticleGFPart::run(int nfreq, double const* d_freq, double* d_acc, bool keep) {
    pomx_g2_part d = devicePart();
    std::shared_ptr<pomx_g2_terms> t;
    if(keep)
        t = std::shared_ptr<pomx_g2_terms>(new pomx_g2_terms{{0, 0, 0, 0}, nullptr}, free_terms);
    detail::check(pomx_g2_part_compute(detail::gpu(), &d, beta, PoleResolution, CoefficientTolerance, nfreq, d_freq, d_acc,
                                       t.get(), counts, &raw_terms),
                  "two-particle GF part");
    terms = t;
    setStatus(keep ? Computed : Constructed);
}

void TwoParticleGFPart::runTangent(SpectrumTangent& T, int nfreq, double const* d_freq, double* d_acc, double* d_dacc) {
    pomx_g2_part d = devicePart();
    pomx_g2_tangent t;
    t.dO1 = T.operatorTangent(O1);
    t.dO2 = T.operatorTangent(O2);
    t.dO3 = T.operatorTangent(O3);
    t.dCX4 = T.operatorTangent(CX4);
    t.dE1 = T.deviceEigenvalueTangent(Hpart1.getBlockNumber());
    t.dE2 = T.deviceEigenvalueTangent(Hpart2.getBlockNumber());
    t.dE3 = T.deviceEigenvalueTangent(Hpart3.getBlockNumber());
    t.dE4 = T.deviceEigenvalueTangent(Hpart4.getBlockNumber());
    t.dw1 = T.deviceWeightTangent(Hpart1.getBlockNumber());
    t.dw2 = T.deviceWeightTangent(Hpart2.getBlockNumber());
    t.dw3 = T.deviceWeightTangent(Hpart3.getBlockNumber());
    t.dw4 = T.deviceWeightTangent(Hpart4.getBlockNumber());
    detail::check(pomx_g2_part_tangent(detail::gpu(), &d, &t, beta, PoleResolution, CoefficientTolerance, nfreq, d_freq,
                                       d_acc, d_dacc, counts, &raw_terms),
                  "two-particle GF part tangent");
}

void TwoParticleGFPart::compute() {
    if(getStatus() >= Computed)
        return;
    run(0, nullptr, nullptr, true);
}

void TwoParticleGFPart::clear() {
    terms.reset();
    setStatus(Constructed);
}

ComplexType TwoParticleGFPart::operator()(ComplexType z1, ComplexType z2, ComplexType z3) const {
    if(getStatus() != Computed || !terms)
        throw StatusMismatch(
            "2PGFPart: Calling operator() on uncomputed container. Did you purge all the terms when called compute()?");
    std::vector<double> f = {z1.real(), z1.imag(), z2.real(), z2.imag(), z3.real(), z3.imag()};
    detail::DeviceArray<double> df(6), acc(2);
    df.upload(f);
    acc.zero();
    detail::check(pomx_g2_terms_eval(detail::gpu(), terms.get(), perm_index, PoleResolution, 1, df.data(), acc.data()),
                  "two-particle GF evaluation");
    auto h = acc.download();
    return ComplexType(h[0], h[1]);
}

ComplexType TwoParticleGFPart::operator()(long n1, long n2, long n3) const {
    return (*this)(MatsubaraSpacing * RealType(2 * n1 + 1), MatsubaraSpacing * RealType(2 * n2 + 1),
                   MatsubaraSpacing * RealType(2 * n3 + 1));
}

std::vector<TwoParticleGFPart::NonResonantTerm> TwoParticleGFPart::getNonResonantTerms() const {
    std::vector<NonResonantTerm> out;
    if(!terms)
        return out;
    int64_t n = counts[0] + counts[1];
    std::vector<double> h(std::size_t(4 * n));
    if(n)
        detail::check(pomx_d2h(detail::gpu(), h.data(), terms->data, sizeof(double) * h.size()), "term copy");
    for(int64_t k = 0; k < n; ++k) {
        NonResonantTerm t;
        t.Coeff = h[std::size_t(k)];
        t.Poles = {h[std::size_t(n + k)], h[std::size_t(2 * n + k)], h[std::size_t(3 * n + k)]};
        t.isz4 = k >= counts[0];
        t.Weight = 1;
        out.push_back(t);
    }
    return out;
}

std::vector<TwoParticleGFPart::ResonantTerm> TwoParticleGFPart::getResonantTerms() const {
    std::vector<ResonantTerm> out;
    if(!terms)
        return out;
    int64_t nnr = counts[0] + counts[1], n = counts[2] + counts[3];
    std::vector<double> h(std::size_t(5 * n));
    if(n)
        detail::check(pomx_d2h(detail::gpu(), h.data(), terms->data + 4 * nnr, sizeof(double) * h.size()), "term copy");
    for(int64_t k = 0; k < n; ++k) {
        ResonantTerm t;
        t.ResCoeff = h[std::size_t(k)];
        t.NonResCoeff = h[std::size_t(n + k)];
        t.Poles = {h[std::size_t(2 * n + k)], h[std::size_t(3 * n + k)], h[std::size_t(4 * n + k)]};
        t.isz1z2 = k >= counts[2];
        t.Weight = 1;
        out.push_back(t);
    }
    return out;
}

// ---- TwoParticleGF -------------------------------------------------------------------------------
TwoParticleGF::TwoParticleGF(StatesClassification const& S, Hamiltonian const& H, AnnihilationOperator const& C1,
                             AnnihilationOperator const& C2, CreationOperator const& CX3, CreationOperator const& CX4,
                             DensityMatrix const& DM)
    : Thermal(DM.beta), S(S), H(H), C1(C1), C2(C2), CX3(CX3), CX4(CX4), DM(DM) {}

BlockNumber TwoParticleGF::getLeftIndex(std::size_t p, std::size_t pos, BlockNumber RightIndex) const {
    switch(permutations3[p].perm[pos]) {
    case 0: return C1.getLeftIndex(RightIndex);
    case 1: return C2.getLeftIndex(RightIndex);
    case 2: return CX3.getLeftIndex(RightIndex);
    default: return INVALID_BLOCK_NUMBER;
    }
}

BlockNumber TwoParticleGF::getRightIndex(std::size_t p, std::size_t pos, BlockNumber LeftIndex) const {
    switch(permutations3[p].perm[pos]) {
    case 0: return C1.getRightIndex(LeftIndex);
    case 1: return C2.getRightIndex(LeftIndex);
    case 2: return CX3.getRightIndex(LeftIndex);
    default: return INVALID_BLOCK_NUMBER;
    }
}

MonomialOperatorPart const& TwoParticleGF::OperatorPartAtPosition(std::size_t p, std::size_t pos, BlockNumber LeftIndex) const {
    switch(permutations3[p].perm[pos]) {
    case 0: return C1.getPartFromLeftIndex(LeftIndex);
    case 1: return C2.getPartFromLeftIndex(LeftIndex);
    case 2: return CX3.getPartFromLeftIndex(LeftIndex);
    default: throw std::runtime_error("TwoParticleGF: Could not find operator part");
    }
}

// TwoParticleGF.cpp:67-120
void TwoParticleGF::prepare() {
    if(getStatus() >= Prepared)
        return;
    for(auto const& outer : CX4.getBlockMapping().right) { // ascending right (source) block of CX4
        for(std::size_t p = 0; p < 6; ++p) {
            std::array<BlockNumber, 4> L{};
            L[0] = outer.first;
            L[3] = outer.second;
            L[2] = getLeftIndex(p, 2, L[3]);
            L[1] = getRightIndex(p, 0, L[0]);
            if(getRightIndex(p, 1, L[1]) == L[2] && L[1] != INVALID_BLOCK_NUMBER && L[2] != INVALID_BLOCK_NUMBER) {
                bool retained = false;
                for(int k = 0; k < 4; ++k)
                    retained = retained || DM.isRetained(L[k]);
                if(!retained)
                    continue;
                parts.emplace_back(OperatorPartAtPosition(p, 0, L[0]), OperatorPartAtPosition(p, 1, L[1]),
                                   OperatorPartAtPosition(p, 2, L[2]), CX4.getPartFromLeftIndex(L[3]), H.getPart(L[0]),
                                   H.getPart(L[1]), H.getPart(L[2]), H.getPart(L[3]), DM.getPart(L[0]), DM.getPart(L[1]),
                                   DM.getPart(L[2]), DM.getPart(L[3]), permutations3[p], PoleResolution,
                                   CoefficientTolerance);
            }
        }
    }
    Vanishing = parts.empty();
    setStatus(Prepared);
}

std::vector<ComplexType> TwoParticleGF::computeParts(std::vector<std::size_t> const& which, bool clear,
                                                     FreqVec3 const& freqs) {
    std::vector<ComplexType> m_data(freqs.size(), 0.0);
    if(Vanishing)
        return m_data;
    int nf = int(freqs.size());
    detail::DeviceArray<double> df(std::size_t(6 * std::max(nf, 1))), acc(std::size_t(2 * std::max(nf, 1)));
    if(nf) {
        std::vector<double> h(std::size_t(6 * nf));
        for(int w = 0; w < nf; ++w) {
            ComplexType z[3] = {std::get<0>(freqs[w]), std::get<1>(freqs[w]), std::get<2>(freqs[w])};
            for(int k = 0; k < 3; ++k) {
                h[std::size_t(6 * w + 2 * k)] = z[k].real();
                h[std::size_t(6 * w + 2 * k + 1)] = z[k].imag();
            }
        }
        df.upload(h);
    }
    acc.zero();
    for(std::size_t k : which)
        parts.at(k).run(nf, nf ? df.data() : nullptr, nf ? acc.data() : nullptr, !clear);
    if(nf) {
        auto h = acc.download();
        for(int w = 0; w < nf; ++w)
            m_data[std::size_t(w)] = ComplexType(h[std::size_t(2 * w)], h[std::size_t(2 * w + 1)]);
    }
    return m_data;
}

std::vector<ComplexType> TwoParticleGF::computeTangent(SpectrumTangent& T, FreqVec3 const& freqs,
                                                       std::vector<ComplexType>* values) {
    if(getStatus() < Prepared)
        throw StatusMismatch("TwoParticleGF is not prepared yet.");
    std::vector<ComplexType> d_data(freqs.size(), 0.0);
    if(values)
        values->assign(freqs.size(), 0.0);
    if(Vanishing || freqs.empty())
        return d_data;
    int nf = int(freqs.size());
    std::vector<double> h(std::size_t(6 * nf));
    for(int w = 0; w < nf; ++w) {
        ComplexType z[3] = {std::get<0>(freqs[w]), std::get<1>(freqs[w]), std::get<2>(freqs[w])};
        for(int k = 0; k < 3; ++k) {
            h[std::size_t(6 * w + 2 * k)] = z[k].real();
            h[std::size_t(6 * w + 2 * k + 1)] = z[k].imag();
        }
    }
    detail::DeviceArray<double> df(h.size()), acc(std::size_t(2 * nf)), dacc(std::size_t(2 * nf));
    df.upload(h);
    acc.zero();
    dacc.zero();
    for(auto& p : parts)
        p.runTangent(T, nf, df.data(), acc.data(), dacc.data());
    auto ha = acc.download(), hd = dacc.download();
    for(int w = 0; w < nf; ++w) {
        d_data[std::size_t(w)] = ComplexType(hd[std::size_t(2 * w)], hd[std::size_t(2 * w + 1)]);
        if(values)
            (*values)[std::size_t(w)] = ComplexType(ha[std::size_t(2 * w)], ha[std::size_t(2 * w + 1)]);
    }
    return d_data;
}

// TwoParticleGF.cpp:161-204 (one GPU: the MPI job farm and Allreduce reduce to a loop)
std::vector<ComplexType> TwoParticleGF::compute(bool clear, FreqVec3 const& freqs, MPI_Comm const&) {
    if(getStatus() < Prepared)
        throw StatusMismatch("TwoParticleGF is not prepared yet.");
    std::vector<ComplexType> m_data;
    if(getStatus() >= Computed)
        return m_data;
    if(!Vanishing) { // a vanishing function returns an empty vector (TwoParticleGF.cpp:169-201)
        std::vector<std::size_t> all(parts.size());
        std::iota(all.begin(), all.end(), std::size_t(0));
        m_data = computeParts(all, clear, freqs);
    }
    setStatus(Computed);
    return m_data;
}

ParticleIndex TwoParticleGF::getIndex(std::size_t Position) const {
    switch(Position) {
    case 0: return C1.getIndex();
    case 1: return C2.getIndex();
    case 2: return CX3.getIndex();
    case 3: return CX4.getIndex();
    default: throw std::runtime_error("TwoParticleGF: Could not get operator index");
    }
}

ComplexType TwoParticleGF::operator()(ComplexType z1, ComplexType z2, ComplexType z3) const {
    if(Vanishing)
        return 0;
    ComplexType s = 0;
    for(auto const& p : parts)
        s += p(z1, z2, z3);
    return s;
}

std::vector<ComplexType> TwoParticleGF::evaluate(FreqVec3 const& freqs) const {
    std::vector<ComplexType> out(freqs.size(), 0.0);
    if(Vanishing || freqs.empty())
        return out;
    int nf = int(freqs.size());
    std::vector<double> h(std::size_t(6 * nf));
    for(int w = 0; w < nf; ++w) {
        ComplexType z[3] = {std::get<0>(freqs[w]), std::get<1>(freqs[w]), std::get<2>(freqs[w])};
        for(int k = 0; k < 3; ++k) {
            h[std::size_t(6 * w + 2 * k)] = z[k].real();
            h[std::size_t(6 * w + 2 * k + 1)] = z[k].imag();
        }
    }
    detail::DeviceArray<double> df(h.size()), acc(std::size_t(2 * nf));
    df.upload(h);
    acc.zero();
    for(auto const& p : parts) {
        if(p.getStatus() != Computed || !p.terms)
            throw StatusMismatch("TwoParticleGF::evaluate: the terms of a part were not kept by compute()");
        detail::check(pomx_g2_terms_eval(detail::gpu(), p.terms.get(), p.perm_index, p.PoleResolution, nf, df.data(),
                                         acc.data()),
                      "two-particle GF evaluation");
    }
    auto ha = acc.download();
    for(int w = 0; w < nf; ++w)
        out[std::size_t(w)] = ComplexType(ha[std::size_t(2 * w)], ha[std::size_t(2 * w + 1)]);
    return out;
}

ComplexType TwoParticleGF::operator()(long n1, long n2, long n3) const {
    return (*this)(MatsubaraSpacing * RealType(2 * n1 + 1), MatsubaraSpacing * RealType(2 * n2 + 1),
                   MatsubaraSpacing * RealType(2 * n3 + 1));
}

// ---- TwoParticleGFContainer (IndexContainer4.hpp:171-273, TwoParticleGFContainer.cpp:23-136) --------
void TwoParticleGFContainer::fill(std::set<IndexCombination4> Indices) {
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

ElementWithPermFreq<TwoParticleGF>& TwoParticleGFContainer::create(IndexCombination4 const& I) {
    auto p = std::make_shared<TwoParticleGF>(S, H, Operators.getAnnihilationOperator(I.Index1),
                                             Operators.getAnnihilationOperator(I.Index2),
                                             Operators.getCreationOperator(I.Index3), Operators.getCreationOperator(I.Index4),
                                             DM);
    auto it = ElementsMap.emplace(I, ElementWithPermFreq<TwoParticleGF>(p, permutations4[0])).first;
    NonTrivialElements.emplace(I, p);
    bool sameC = I.Index1 == I.Index2, sameCX = I.Index3 == I.Index4;
    if(!sameC) {
        IndexCombination4 J(I.Index2, I.Index1, I.Index3, I.Index4);
        if(!isInContainer(J))
            ElementsMap.emplace(J, ElementWithPermFreq<TwoParticleGF>(p, permutations4[6]));
    }
    if(!sameCX) {
        IndexCombination4 J(I.Index1, I.Index2, I.Index4, I.Index3);
        if(!isInContainer(J))
            ElementsMap.emplace(J, ElementWithPermFreq<TwoParticleGF>(p, permutations4[1]));
    }
    if(!sameC && !sameCX) {
        IndexCombination4 J(I.Index2, I.Index1, I.Index4, I.Index3);
        if(!isInContainer(J))
            ElementsMap.emplace(J, ElementWithPermFreq<TwoParticleGF>(p, permutations4[7]));
    }
    return it->second;
}

void TwoParticleGFContainer::prepareAll(std::set<IndexCombination4> const& Indices) {
    fill(Indices);
    for(auto& kv : NonTrivialElements) {
        kv.second->PoleResolution = PoleResolution;
        kv.second->CoefficientTolerance = CoefficientTolerance;
        kv.second->prepare();
    }
}

// TwoParticleGFContainer.cpp:33-136.  split (the default): one entry per computed element (on one
// rank the reference's communicator split gives every element to this rank).  Without split,
// every element and image in index order: the first of an element's entries computes it, the
// others find it computed and get an empty vector, as TwoParticleGF::compute returns then.
std::map<IndexCombination4, std::vector<ComplexType>>
TwoParticleGFContainer::computeAll(bool clearTerms, FreqVec3 const& freqs, MPI_Comm const& comm, bool split) {
    std::map<IndexCombination4, std::vector<ComplexType>> out;
    if(split)
        for(auto& kv : NonTrivialElements)
            out.emplace(kv.first, kv.second->compute(clearTerms, freqs, comm));
    else
        for(auto& kv : ElementsMap)
            out.emplace(kv.first, static_cast<TwoParticleGF&>(kv.second).compute(clearTerms, freqs, comm));
    return out;
}

ElementWithPermFreq<TwoParticleGF>& TwoParticleGFContainer::operator()(IndexCombination4 const& Indices) {
    auto it = ElementsMap.find(Indices);
    if(it == ElementsMap.end())
        return create(Indices);
    return it->second;
}

} // namespace Pomerol
