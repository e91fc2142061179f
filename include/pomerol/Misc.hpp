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

// MPI: the communicator type and MPI_COMM_WORLD of <mpi.h> where it is available (so callers that
// include it and call MPI themselves, as the reference's tests and programs do, see one type);
// otherwise a single-process stand-in with MPICH's representation, so the library's interface is
// the same either way.  The library itself never calls MPI: it runs on one GPU.
#if __has_include(<mpi.h>)
#include <mpi.h>
#else
typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#endif

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

namespace detail {
/// Comma initializer of a pre-sized vector or matrix, as Eigen's `x << a, b, c;`: coefficients in
/// storage order (row by row for matrices), from the first.
template <typename Obj, typename T> class CommaInit {
    Obj& obj_;
    std::size_t at_ = 0;
    void put(T const& x) {
        if(at_ >= std::size_t(obj_.size()))
            throw std::out_of_range("comma initializer: too many coefficients");
        obj_.data()[at_++] = x;
    }

public:
    CommaInit(Obj& obj, T const& first) : obj_(obj) { put(first); }
    CommaInit& operator,(T const& x) {
        put(x);
        return *this;
    }
};
} // namespace detail

/// Dense vector (Eigen-style access, arithmetic and printing on a std::vector).
template <typename T> class DenseVector : public std::vector<T> {
    using base = std::vector<T>;

public:
    using Scalar = T;
    using base::base;
    DenseVector() = default;
    explicit DenseVector(long n) : base(std::size_t(n), T(0)) {}
    explicit DenseVector(int n) : base(std::size_t(n), T(0)) {}
    DenseVector(base v) : base(std::move(v)) {}
    T& operator()(long i) { return (*this)[std::size_t(i)]; }
    T const& operator()(long i) const { return (*this)[std::size_t(i)]; }
    long rows() const { return long(base::size()); }
    long cols() const { return 1; }
    T sum() const {
        T s = T(0);
        for(T const& x : *this)
            s += x;
        return s;
    }
    template <typename U = T> U minCoeff() const {
        U m = base::empty() ? U(0) : base::front();
        for(T const& x : *this)
            m = x < m ? x : m;
        return m;
    }
    template <typename U = T> U maxCoeff() const {
        U m = base::empty() ? U(0) : base::front();
        for(T const& x : *this)
            m = x > m ? x : m;
        return m;
    }
    DenseVector<RealType> cwiseAbs() const {
        DenseVector<RealType> a(long(base::size()));
        for(std::size_t k = 0; k < base::size(); ++k)
            a[k] = std::abs((*this)[k]);
        return a;
    }
    DenseVector const& eval() const { return *this; }
    DenseVector& operator*=(T const& x) {
        for(T& v : *this)
            v *= x;
        return *this;
    }
    DenseVector& operator/=(T const& x) {
        for(T& v : *this)
            v /= x;
        return *this;
    }
    DenseVector& operator+=(DenseVector const& o) {
        for(std::size_t k = 0; k < base::size(); ++k)
            (*this)[k] += o[k];
        return *this;
    }
    DenseVector& operator-=(DenseVector const& o) {
        for(std::size_t k = 0; k < base::size(); ++k)
            (*this)[k] -= o[k];
        return *this;
    }
    detail::CommaInit<DenseVector, T> operator<<(T const& x) { return detail::CommaInit<DenseVector, T>(*this, x); }
    friend DenseVector operator+(DenseVector a, DenseVector const& b) { return a += b; }
    friend DenseVector operator-(DenseVector a, DenseVector const& b) { return a -= b; }
    friend DenseVector operator*(T const& x, DenseVector a) { return a *= x; }
    friend DenseVector operator*(DenseVector a, T const& x) { return a *= x; }
    friend DenseVector operator/(DenseVector a, T const& x) { return a /= x; }
    /// One coefficient per line, as Eigen prints a column vector.
    friend std::ostream& operator<<(std::ostream& os, DenseVector const& v) {
        for(std::size_t k = 0; k < v.size(); ++k)
            os << v[k] << (k + 1 < v.size() ? "\n" : "");
        return os;
    }
};

/// Row-major dense matrix.
template <typename T> class DenseMatrix {
    std::vector<T> data_;
    long rows_ = 0, cols_ = 0;

public:
    using Scalar = T;
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
    DenseVector<T> col(long c) const {
        DenseVector<T> v(rows_);
        for(long r = 0; r < rows_; ++r)
            v[std::size_t(r)] = (*this)(r, c);
        return v;
    }
    DenseVector<T> row(long r) const {
        DenseVector<T> v(cols_);
        for(long c = 0; c < cols_; ++c)
            v[std::size_t(c)] = (*this)(r, c);
        return v;
    }
    DenseMatrix transpose() const {
        DenseMatrix t(cols_, rows_);
        for(long r = 0; r < rows_; ++r)
            for(long c = 0; c < cols_; ++c)
                t(c, r) = (*this)(r, c);
        return t;
    }
    DenseMatrix const& eval() const { return *this; }
    /// Zero every entry with |x| <= |reference| * epsilon (Eigen's SparseMatrix::prune).
    void prune(T const& reference, RealType epsilon = 1e-12) {
        for(T& x : data_)
            if(std::abs(x) <= std::abs(reference) * epsilon)
                x = T(0);
    }
    long nonZeros() const {
        long n = 0;
        for(T const& x : data_)
            n += (x != T(0)) ? 1 : 0;
        return n;
    }
    DenseMatrix<RealType> cwiseAbs() const {
        DenseMatrix<RealType> a(rows_, cols_);
        for(long k = 0; k < size(); ++k)
            a.data()[k] = std::abs(data_[std::size_t(k)]);
        return a;
    }
    RealType maxCoeff() const {
        RealType m = data_.empty() ? 0.0 : std::real(data_.front());
        for(T const& x : data_)
            m = std::real(x) > m ? std::real(x) : m;
        return m;
    }
    DenseMatrix& operator-=(DenseMatrix const& o) {
        for(std::size_t k = 0; k < data_.size(); ++k)
            data_[k] -= o.data_[k];
        return *this;
    }
    DenseMatrix& operator+=(DenseMatrix const& o) {
        for(std::size_t k = 0; k < data_.size(); ++k)
            data_[k] += o.data_[k];
        return *this;
    }
    friend DenseMatrix operator-(DenseMatrix a, DenseMatrix const& b) { return a -= b; }
    friend DenseMatrix operator+(DenseMatrix a, DenseMatrix const& b) { return a += b; }
    friend DenseVector<T> operator*(DenseMatrix const& m, DenseVector<T> const& v) {
        DenseVector<T> out(m.rows_);
        for(long r = 0; r < m.rows_; ++r) {
            T s = T(0);
            for(long c = 0; c < m.cols_; ++c)
                s += m(r, c) * v[std::size_t(c)];
            out[std::size_t(r)] = s;
        }
        return out;
    }
    detail::CommaInit<DenseMatrix, T> operator<<(T const& x) { return detail::CommaInit<DenseMatrix, T>(*this, x); }
    /// Rows on separate lines, coefficients separated by spaces, as Eigen prints a matrix.
    friend std::ostream& operator<<(std::ostream& os, DenseMatrix const& m) {
        for(long r = 0; r < m.rows_; ++r) {
            for(long c = 0; c < m.cols_; ++c)
                os << m(r, c) << (c + 1 < m.cols_ ? " " : "");
            if(r + 1 < m.rows_)
                os << '\n';
        }
        return os;
    }
};

using RealVectorType = DenseVector<RealType>;
using ComplexVectorType = DenseVector<ComplexType>;
using RealMatrixType = DenseMatrix<RealType>;
using ComplexMatrixType = DenseMatrix<ComplexType>;
template <bool Complex> using MatrixType = DenseMatrix<MelemType<Complex>>;
template <bool Complex> using RowMajorMatrixType = DenseMatrix<MelemType<Complex>>;
template <bool Complex> using ColMajorMatrixType = DenseMatrix<MelemType<Complex>>;
template <bool Complex> using VectorType = DenseVector<MelemType<Complex>>;

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

} // namespace Pomerol
