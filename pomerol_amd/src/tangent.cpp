// SpectrumTangent: first-order response of the eigendecomposition, the Boltzmann weights and
// the rotated operators to one Hamiltonian direction (no counterpart in the reference).
#include "pomerol/Tangent.hpp"

#include <stdexcept>

namespace Pomerol {

void SpectrumTangent::compute() {
    if(getStatus() >= Computed)
        return;
    if(getStatus() < Prepared)
        throw StatusMismatch("SpectrumTangent is not prepared yet.");
    if(H.getStatus() < Computed || DM.getStatus() < Computed)
        throw StatusMismatch("SpectrumTangent: the Hamiltonian and the density matrix must be computed first.");
    int nb = int(S.getNumberOfBlocks());
    std::vector<int> n(std::size_t(nb), 0);
    std::vector<int64_t> vo(std::size_t(nb), 0);
    mat_off.assign(std::size_t(nb), 0);
    int64_t mat_total = 0, nstates = 0;
    double const* U0 = H.getPart(0).deviceMatrix();
    double const* D0 = dH.getPart(0).deviceMatrix();
    for(int b = 0; b < nb; ++b) {
        auto const& p = H.getPart(BlockNumber(b));
        n[std::size_t(b)] = int(p.getSize());
        mat_off[std::size_t(b)] = p.deviceMatrix() - U0;
        vo[std::size_t(b)] = H.eigenvalueOffset(BlockNumber(b));
        if(dH.getPart(BlockNumber(b)).deviceMatrix() - D0 != mat_off[std::size_t(b)])
            throw std::logic_error("SpectrumTangent: direction and Hamiltonian blocks are laid out differently");
        mat_total = std::max(mat_total, mat_off[std::size_t(b)] + int64_t(n[std::size_t(b)]) * n[std::size_t(b)]);
        nstates = std::max(nstates, vo[std::size_t(b)] + int64_t(n[std::size_t(b)]));
    }
    K = detail::DeviceArray<double>(std::size_t(mat_total));
    dE = detail::DeviceArray<double>(std::size_t(nstates));
    dw = detail::DeviceArray<double>(std::size_t(nstates));
    detail::check(pomx_eigh_tangent(detail::gpu(), nb, n.data(), mat_off.data(), vo.data(), U0, H.deviceEigenvalues(), D0,
                                    K.data(), dE.data()),
                  "eigendecomposition tangent");
    detail::check(pomx_boltzmann_tangent(detail::gpu(), nstates, DM.getPart(0).deviceWeights(), dE.data(), DM.beta,
                                         dw.data()),
                  "Boltzmann weight tangent");
    op_cache.clear();
    setStatus(Computed);
}

double const* SpectrumTangent::deviceEigenvalueTangent(BlockNumber b) const {
    if(getStatus() < Computed)
        throw StatusMismatch("SpectrumTangent is not computed yet.");
    return dE.data() + H.eigenvalueOffset(b);
}

double const* SpectrumTangent::deviceWeightTangent(BlockNumber b) const {
    if(getStatus() < Computed)
        throw StatusMismatch("SpectrumTangent is not computed yet.");
    return dw.data() + H.eigenvalueOffset(b);
}

double const* SpectrumTangent::deviceGenerator(BlockNumber b) const {
    if(getStatus() < Computed)
        throw StatusMismatch("SpectrumTangent is not computed yet.");
    return K.data() + mat_off.at(std::size_t(b));
}

RealVectorType SpectrumTangent::getEigenValueTangents(BlockNumber b) const {
    RealVectorType v(std::size_t(H.getBlockSize(b)));
    dE.download(v.data(), v.size(), std::size_t(H.eigenvalueOffset(b)));
    return v;
}

RealVectorType SpectrumTangent::getWeightTangents(BlockNumber b) const {
    RealVectorType v(std::size_t(H.getBlockSize(b)));
    dw.download(v.data(), v.size(), std::size_t(H.eigenvalueOffset(b)));
    return v;
}

pomx_view SpectrumTangent::operatorTangent(MonomialOperatorPart const& part) {
    if(getStatus() < Computed)
        throw StatusMismatch("SpectrumTangent is not computed yet.");
    pomx_view v = part.deviceView();
    int const nL = int(part.rows()), nR = int(part.cols());
    auto key = std::make_tuple(v.base, v.rs, v.cs);
    auto it = op_cache.find(key);
    if(it == op_cache.end()) {
        // the view as a GEMM operand: row-major (ld = rs) or the transpose of a row-major matrix
        // (ld = cs); a single row or column has no meaningful stride of its own
        int tO, ldO;
        if(v.cs == 1 && (v.rs >= nR || nL == 1)) {
            tO = 0;
            ldO = nL == 1 ? nR : int(v.rs);
        } else if(v.rs == 1 && (v.cs >= nL || nR == 1)) {
            tO = 1;
            ldO = nR == 1 ? nL : int(v.cs);
        } else
            throw std::logic_error("SpectrumTangent: operator part is neither row- nor column-major");
        auto out = std::make_shared<detail::DeviceArray<double>>(std::size_t(int64_t(nL) * nR));
        int64_t const kL = mat_off.at(std::size_t(part.getLeftIndex())), kR = mat_off.at(std::size_t(part.getRightIndex()));
        pomx_gemm_job right{nL, nR, nR, tO, 0, ldO, nR, nR, 0, kR, 0, 1.0, 0.0}; // dO = O K_R
        pomx_gemm_job left{nL, nR, nL, 0, tO, nL, ldO, nR, kL, 0, 0, -1.0, 1.0}; // dO -= K_L O
        detail::check(pomx_gemm_batched(detail::gpu(), 1, &right, v.base, K.data(), out->data()), "operator tangent");
        detail::check(pomx_gemm_batched(detail::gpu(), 1, &left, K.data(), v.base, out->data()), "operator tangent");
        it = op_cache.emplace(key, out).first;
    }
    return pomx_view{it->second->data(), nR, 1};
}

} // namespace Pomerol
