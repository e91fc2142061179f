// oracle/ref_cpu.hpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of pomerol's full-ED path (sector partition, block eigensolve, operator
// rotation, Lehmann-term generation with TermList merging, and frequency evaluation), used as
// the parity checker for the HIP product path and as the timed CPU baseline.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  It is never part of
// the product: pomerol_amd/ must not include, link or import anything under oracle/.
//
// Each routine cites the reference file:line it restates.  Deliberate choices:
//  * Eigen's SelfAdjointEigenSolver (HamiltonianPart.cpp:98) is replaced by a cyclic Jacobi
//    solver, an algorithm independent of the product's Householder+QL/D&C path.  Output order
//    matches Eigen's: eigenvalues ascending, eigenvectors as columns of a row-major matrix.
//  * libcommute's space_partition (HilbertSpace.hpp:144-169) is restated as BFS connected
//    components (Phase I) plus a fixpoint merge (Phase II).  Blocks are numbered by their
//    smallest Fock state, states ascending inside a block (parity unpinned: libcommute's own
//    numbering is not in the reference tree).
//  * Fermionic sign convention: c_i / c_i^dag pick up (-1)^(number of occupied modes j < i).
#pragma once
#include <complex>
#include <cstdint>
#include <functional>
#include <string>
#include <unordered_set>
#include <vector>

namespace oracle {

using cplx = std::complex<double>;

struct FOp {
    bool dag;
    int mode;
};
using Mono = std::vector<FOp>; // ordered product; the rightmost operator acts first
struct HTerm {
    double coeff;
    Mono ops;
};

int apply_mono(Mono const& m, uint64_t s, uint64_t& out);

struct Sectors {
    std::vector<std::vector<uint64_t>> blocks; // block -> ascending Fock states
    std::vector<int> state_block;             // Fock state -> block
    std::vector<int> inner;                   // Fock state -> position inside its block
};

// HilbertSpace.hpp:144-169 + StatesClassification.hpp:110-118
Sectors partition(int n_modes, std::vector<HTerm> const& H);

// cyclic Jacobi; A row-major n x n (destroyed); V receives eigenvectors as columns; w ascending
void jacobi_eigh(int n, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w);
void tridiag_qr_eigh(int n, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w);

// Rotated monomial operator, one entry per connected (from -> to) block pair.
struct OpPart {
    int from, to;            // right (source) block, left (target) block
    int n_to, n_from;        // matrix is n_to x n_from, row-major
    std::vector<double> m;   // entries with |x| <= Tol*1e-12 stored as exact 0 (sparseView)
};
struct RotOp {
    std::vector<OpPart> parts;               // ascending 'from' (find_connections order)
    std::vector<int> left_of_right, right_of_left; // -1 where absent
};

struct GFResult {
    std::vector<cplx> values;
    std::vector<cplx> tau_values; // G(tau) on the requested imaginary times
    long n_parts = 0, n_terms = 0;
    std::vector<double> poles; std::vector<cplx> residues; // merged terms, part by part
};
struct ChiResult {
    std::vector<cplx> values;
    std::vector<cplx> tau_values; // chi(tau) on the requested imaginary times
    long n_parts = 0, n_terms = 0;
    cplx zero_pole_weight = 0;
};
struct G2Result {
    std::vector<cplx> values;
    long n_parts = 0, n_nonres = 0, n_res = 0;
    long n_quadruples = 0; // weighted index quadruples visited (sum of weights above tolerance)
    long n_erased = 0;     // merged terms dropped by TermList's tol/(size+1) rule
    double gen_seconds = 0, eval_seconds = 0; // term generation + merge; frequency fill
    std::vector<long> part_nonres, part_res;
    std::vector<int> part_blocks; // 5 ints per part: S1 S2 S3 S4 perm
};
struct Chi3Result {
    std::vector<cplx> values;
    long n_parts = 0, n_ff = 0, n_fb = 0, n_res = 0;
    long n_triples = 0; // index triples with a non-zero matrix element
    std::vector<long> part_ff, part_fb, part_res;
    std::vector<int> part_blocks; // 4 ints per part: S1 S2 S3 swapped
};

class ED {
public:
    // eigensolver 0: cyclic Jacobi sweeps; 1: Householder tridiagonalisation + implicit QR
    ED(int n_modes, std::vector<HTerm> H, double beta, int eigensolver = 0);

    int n_modes;
    double beta;
    std::vector<HTerm> H;
    Sectors S;
    std::vector<std::vector<double>> hmat;   // assembled blocks (row-major), kept for checks
    std::vector<std::vector<double>> evecs;  // eigenvectors as columns (row-major)
    std::vector<std::vector<double>> evals;  // ascending
    std::vector<std::vector<double>> weights;
    std::vector<char> retained;
    double E0 = 0, Z = 0;

    void truncate_blocks(double tol); // DensityMatrix.cpp:77-92
    // Replace block b's spectrum (for parity checks of the term pipeline on identical input).
    void set_spectrum(int b, std::vector<double> const& evals, std::vector<double> const& evecs,
                      std::vector<double> const& weights);

    RotOp rotate(Mono const& op, double tol = 1e-8) const;  // MonomialOperatorPart.cpp:48-92
    static RotOp adjoint(RotOp const& cdag);                  // MonomialOperatorPart.cpp:94-111

    GFResult gf(int i, int j, std::vector<cplx> const& z, double pole_res, double coeff_tol,
                std::vector<double> const& tau = {}) const;
    ChiResult chi(Mono const& A, Mono const& B, std::vector<cplx> const& z, double pole_res,
                  double coeff_tol, bool subtract_disconnected, std::vector<double> const& tau = {}) const;
    cplx ensemble_average(Mono const& A) const; // EnsembleAverage.cpp:27-67
    // 2PGF <c_i c_j c^dag_k c^dag_l> at the frequency triplets (z1, z2, z3) of zs (3 per point).
    G2Result g2(int i, int j, int k, int l, std::vector<cplx> const& zs, double pole_res,
                double coeff_tol, int max_parts = -1, std::vector<int> const* part_subset = nullptr) const;
    // 3-point susceptibility of (c^dag_i, c_j, c^dag_k, c_l) in channel 0 PP, 1 PH, 2 xPH at the
    // frequency pairs (z1, z2) of zs (2 per point).
    Chi3Result chi3(int channel, int i, int j, int k, int l, std::vector<cplx> const& zs, double pole_res,
                    double coeff_tol) const;

private:
    mutable std::vector<RotOp> cdag_cache_, c_cache_;
    mutable std::vector<char> cache_ok_;
    RotOp const& cdag_op(int i) const;
    RotOp const& c_op(int i) const;
};

} // namespace oracle
