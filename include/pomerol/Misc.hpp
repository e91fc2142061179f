// Basic types of the pomerol-compatible API (reference: include/pomerol/Misc.hpp:70-179,
// ComputableObject.hpp:31-58, Thermal.hpp:29-56).
//
// Dense matrices are this library's own row-major container (the reference uses Eigen);
// rotated operator blocks are returned dense (the reference's sparseView keeps every entry
// with |x| > Tol*1e-12, so its sparse blocks are dense in practice).
#pragma once

#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace Pomerol {

#define INFO(MSG) std::cout << MSG << '\n'
#define INFO_NONEWLINE(MSG) std::cout << MSG
#define ERROR(MSG) std::cerr << __FILE__ << ":" << __LINE__ << ": " << MSG << '\n'

using RealType = double;
using ComplexType = std::complex<double>;
using ParticleIndex = unsigned int;
using QuantumState = std::uint64_t;
using InnerQuantumState = std::uint64_t;
using BlockNumber = int;
constexpr BlockNumber INVALID_BLOCK_NUMBER = -1;

template <bool Complex> using MelemType = typename std::conditional<Complex, ComplexType, RealType>::type;

static ComplexType const I = ComplexType(0.0, 1.0);

/// Row-major dense matrix.
template <typename T> class DenseMatrix {
    std::vector<T> data_;
    long rows_ = 0, cols_ = 0;

public:
    DenseMatrix() = default;
    DenseMatrix(long r, long c) : data_(std::size_t(r) * std::size_t(c), T(0)), rows_(r), cols_(c) {}
    long rows() const { return rows_; }
    long cols() const { return cols_; }
    long size() const { return rows_ * cols_; }
    T& operator()(long r, long c) { return data_[std::size_t(r) * cols_ + c]; }
    T const& operator()(long r, long c) const { return data_[std::size_t(r) * cols_ + c]; }
    T coeff(long r, long c) const { return (*this)(r, c); }
    T* data() { return data_.data(); }
    T const* data() const { return data_.data(); }
    std::vector<T> col(long c) const {
        std::vector<T> v(rows_);
        for(long r = 0; r < rows_; ++r)
            v[r] = (*this)(r, c);
        return v;
    }
    DenseMatrix transpose() const {
        DenseMatrix t(cols_, rows_);
        for(long r = 0; r < rows_; ++r)
            for(long c = 0; c < cols_; ++c)
                t(c, r) = (*this)(r, c);
        return t;
    }
};

template <bool Complex> using MatrixType = DenseMatrix<MelemType<Complex>>;
template <bool Complex> using RowMajorMatrixType = DenseMatrix<MelemType<Complex>>;
template <bool Complex> using ColMajorMatrixType = DenseMatrix<MelemType<Complex>>;

/// Dense real vector with Eigen-style element access.
class RealVectorType : public std::vector<RealType> {
public:
    using std::vector<RealType>::vector;
    RealType& operator()(long i) { return (*this)[std::size_t(i)]; }
    RealType operator()(long i) const { return (*this)[std::size_t(i)]; }
    RealType minCoeff() const {
        RealType m = empty() ? 0.0 : front();
        for(RealType x : *this)
            m = x < m ? x : m;
        return m;
    }
};

class ComplexVectorType : public std::vector<ComplexType> {
public:
    using std::vector<ComplexType>::vector;
    ComplexType& operator()(long i) { return (*this)[std::size_t(i)]; }
    ComplexType operator()(long i) const { return (*this)[std::size_t(i)]; }
};

/// Permutation of 3 elements with its signature (Misc.hpp:125-139, Misc.cpp:41-42).
struct Permutation3 {
    std::array<std::size_t, 3> perm;
    int sign;
    bool operator==(Permutation3 const& rhs) const { return sign == rhs.sign && perm == rhs.perm; }
    bool operator!=(Permutation3 const& rhs) const { return !(*this == rhs); }
};
extern std::array<Permutation3, 6> const permutations3;

/// Permutation of 4 elements with its signature (Misc.hpp:141-157, Misc.cpp:57-66).
struct Permutation4 {
    std::array<std::size_t, 4> perm;
    int sign;
    bool operator==(Permutation4 const& rhs) const { return sign == rhs.sign && perm == rhs.perm; }
    bool operator!=(Permutation4 const& rhs) const { return !(*this == rhs); }
};
extern std::array<Permutation4, 24> const permutations4;

std::ostream& operator<<(std::ostream& os, Permutation3 const& p);
std::ostream& operator<<(std::ostream& os, Permutation4 const& p);

/// Channel in which a susceptibility function is defined (Misc.hpp:159-165, Misc.cpp:69-79).
enum Channel : int {
    PP,  ///< particle-particle
    PH,  ///< particle-hole
    xPH  ///< crossed particle-hole
};
std::ostream& operator<<(std::ostream& os, Channel channel);

/// Bin index of a real number (Misc.cpp:83-86).
std::size_t hash_binned_real(double x, double bin_size);

/// Status machine of computable objects (ComputableObject.hpp:31-58).
class ComputableObject {
public:
    enum StatusEnum { Constructed, Prepared, Computed };

    class StatusMismatch : public std::runtime_error {
    public:
        explicit StatusMismatch(std::string const& str) : std::runtime_error(str) {}
    };

protected:
    StatusEnum Status = Constructed;

public:
    StatusEnum getStatus() const { return Status; }
    void setStatus(StatusEnum Status_in) {
        if(Status_in > Computed)
            throw std::logic_error("Wrong status");
        Status = Status_in;
    }
};

/// Thermal object: inverse temperature and Matsubara spacing (Thermal.hpp:29-56).
struct Thermal {
    RealType const beta;
    ComplexType const MatsubaraSpacing;
    explicit Thermal(RealType beta) : beta(beta), MatsubaraSpacing(I * M_PI / beta) {}
};

/// Single-process stand-in for an MPI communicator: the library runs on one GPU, every
/// `comm` argument of the reference API is accepted and ignored.
struct MPI_Comm_t {
    int id = 0;
};
using MPI_Comm = MPI_Comm_t;
static MPI_Comm const MPI_COMM_WORLD{};

} // namespace Pomerol
