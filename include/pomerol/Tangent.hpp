// First-order response of a solved model to one direction of its Hamiltonian.
//
// No counterpart in the reference: forward-mode differentiation of the correlators with respect
// to one Hamiltonian parameter t, H(t) = H + t dH at t = 0 (for instance dH = dH/dU of a Hubbard
// interaction).  SpectrumTangent::compute moves, on the device,
//   the block eigenvalues    dE = diag(U^T dH U),
//   the eigenvectors         dU = U K, K_ij = (U^T dH U)_ij / (E_j - E_i) (antisymmetric; zero
//                            inside degenerate eigenspaces),
//   the Boltzmann weights    dw = -beta w (dE - <dE>),
// and operatorTangent gives the motion of a rotated operator part O = U_L^T M U_R,
//   dO = O K_R - K_L O   (two MFMA GEMMs; cached per part).
// TwoParticleGF::computeTangent evaluates the derivative of the two-particle GF from these.
#pragma once

#include "MonomialOperator.hpp"

#include <map>
#include <memory>
#include <tuple>

namespace Pomerol {

class SpectrumTangent : public ComputableObject {
    StatesClassification const& S;
    Hamiltonian const& H;
    DensityMatrix const& DM;
    Hamiltonian dH; // the direction, assembled block by block in the layout of H
    detail::DeviceArray<double> K, dE, dw;
    std::vector<int64_t> mat_off;
    std::map<std::tuple<double const*, int64_t, int64_t>, std::shared_ptr<detail::DeviceArray<double>>> op_cache;

public:
    SpectrumTangent(StatesClassification const& S, Hamiltonian const& H, DensityMatrix const& DM)
        : S(S), H(H), DM(DM), dH(S) {}

    /// Assemble the direction dH (an expression over the Hilbert space of H, conserving its
    /// symmetry blocks).
    template <typename ScalarType, typename... IndexTypes>
    void prepare(Operators::expression<ScalarType, IndexTypes...> const& Direction, HilbertSpace<IndexTypes...> const& HS) {
        if(getStatus() >= Prepared)
            return;
        dH.prepare(Direction, HS);
        setStatus(Prepared);
    }
    /// dE, K and dw (H and the density matrix must be computed).
    void compute();

    RealVectorType getEigenValueTangents(BlockNumber Block) const;
    RealVectorType getWeightTangents(BlockNumber Block) const;
    double const* deviceEigenvalueTangent(BlockNumber Block) const;
    double const* deviceWeightTangent(BlockNumber Block) const;
    /// K of a block (row-major, n x n), device.
    double const* deviceGenerator(BlockNumber Block) const;
    /// dO of a rotated operator part (row-major rows() x cols(), device; computed once per part).
    pomx_view operatorTangent(MonomialOperatorPart const& part);
};

} // namespace Pomerol
