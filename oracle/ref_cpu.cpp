// oracle/ref_cpu.cpp -- TEST INFRASTRUCTURE ONLY (see ref_cpu.hpp).
#include "ref_cpu.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <limits>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

int apply_mono(Mono const& m, uint64_t s, uint64_t& out) {
    int sign = 1;
    for(int k = int(m.size()) - 1; k >= 0; --k) {
        uint64_t bit = uint64_t(1) << m[k].mode;
        bool occ = (s & bit) != 0;
        if(m[k].dag == occ)
            return 0;
        if(__builtin_popcountll(s & (bit - 1)) & 1)
            sign = -sign;
        s ^= bit;
    }
    out = s;
    return sign;
}

// H|s> as a combined list of (target, amplitude), amplitudes summed over terms
static void apply_H(std::vector<HTerm> const& H, uint64_t s, std::vector<std::pair<uint64_t, double>>& out) {
    out.clear();
    for(auto const& t : H) {
        uint64_t f;
        int sg = apply_mono(t.ops, s, f);
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

// HilbertSpace.hpp:144-169 (Phase I: connected components of H; Phase II: merge so that each
// c_i, c^dag_i maps a sector into at most one sector), StatesClassification.hpp:110-118.
Sectors partition(int n_modes, std::vector<HTerm> const& H) {
    uint64_t dim = uint64_t(1) << n_modes;
    std::vector<int> lab(dim, -1);
    std::vector<std::vector<uint64_t>> adj(dim);
    std::vector<std::pair<uint64_t, double>> buf;
    for(uint64_t s = 0; s < dim; ++s) {
        apply_H(H, s, buf);
        for(auto const& p : buf)
            if(p.second != 0.0 && p.first != s) {
                adj[s].push_back(p.first);
                adj[p.first].push_back(s);
            }
    }
    int nlab = 0;
    std::vector<uint64_t> stack;
    for(uint64_t s = 0; s < dim; ++s) {
        if(lab[s] >= 0)
            continue;
        lab[s] = nlab;
        stack.assign(1, s);
        while(!stack.empty()) {
            uint64_t x = stack.back();
            stack.pop_back();
            for(uint64_t y : adj[x])
                if(lab[y] < 0) {
                    lab[y] = nlab;
                    stack.push_back(y);
                }
        }
        ++nlab;
    }
    adj.clear();
    adj.shrink_to_fit();
    // Phase II: relabel until every field operator maps a label to a single label
    bool changed = true;
    while(changed) {
        changed = false;
        for(int i = 0; i < n_modes; ++i)
            for(int d = 0; d < 2; ++d) {
                Mono op{{d == 1, i}};
                std::map<int, int> first_target;
                std::map<int, int> relabel; // label -> label to merge into
                for(uint64_t s = 0; s < dim; ++s) {
                    uint64_t f;
                    if(!apply_mono(op, s, f))
                        continue;
                    auto it = first_target.find(lab[s]);
                    if(it == first_target.end())
                        first_target.emplace(lab[s], lab[f]);
                    else if(it->second != lab[f]) {
                        int a = it->second, b = lab[f];
                        if(a != b) {
                            int lo = std::min(a, b), hi = std::max(a, b);
                            for(uint64_t x = 0; x < dim; ++x)
                                if(lab[x] == hi)
                                    lab[x] = lo;
                            for(auto& kv : first_target)
                                if(kv.second == hi)
                                    kv.second = lo;
                            changed = true;
                        }
                    }
                }
            }
    }
    // number blocks by their smallest Fock state
    std::map<int, int> order;
    Sectors S;
    S.state_block.assign(dim, -1);
    S.inner.assign(dim, -1);
    for(uint64_t s = 0; s < dim; ++s) {
        auto it = order.find(lab[s]);
        int b;
        if(it == order.end()) {
            b = int(S.blocks.size());
            order.emplace(lab[s], b);
            S.blocks.emplace_back();
        } else
            b = it->second;
        S.state_block[s] = b;
        S.inner[s] = int(S.blocks[b].size());
        S.blocks[b].push_back(s);
    }
    return S;
}

void jacobi_eigh(int n, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w) {
    V.assign(std::size_t(n) * n, 0.0);
    for(int i = 0; i < n; ++i)
        V[std::size_t(i) * n + i] = 1.0;
    auto a = [&](int r, int c) -> double& { return A[std::size_t(r) * n + c]; };
    double fro = 0;
    for(double x : A)
        fro += x * x;
    for(int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for(int p = 0; p < n; ++p)
            for(int q = p + 1; q < n; ++q)
                off += a(p, q) * a(p, q);
        if(off <= 1e-36 * fro || off == 0.0)
            break;
        for(int p = 0; p < n - 1; ++p)
            for(int q = p + 1; q < n; ++q) {
                double apq = a(p, q);
                if(apq == 0.0)
                    continue;
                double app = a(p, p), aqq = a(q, q);
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for(int k = 0; k < n; ++k) { // columns p, q
                    double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for(int k = 0; k < n; ++k) { // rows p, q
                    double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                a(p, q) = a(q, p) = 0.0;
                for(int k = 0; k < n; ++k) {
                    double vkp = V[std::size_t(k) * n + p], vkq = V[std::size_t(k) * n + q];
                    V[std::size_t(k) * n + p] = c * vkp - s * vkq;
                    V[std::size_t(k) * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
    w.resize(n);
    std::vector<double> V2(V.size());
    for(int j = 0; j < n; ++j) {
        w[j] = a(idx[j], idx[j]);
        for(int k = 0; k < n; ++k)
            V2[std::size_t(k) * n + j] = V[std::size_t(k) * n + idx[j]];
    }
    V.swap(V2);
}

// The algorithm of Eigen's SelfAdjointEigenSolver, which HamiltonianPart.cpp:98 calls (Eigen is
// not in the reference tree; restated from Golub & Van Loan, Matrix Computations, 4th ed.,
// algorithms 8.3.1 and 8.3.3): Householder tridiagonalisation Q^T A Q = T with Q accumulated,
// then implicit symmetric QR steps with the Wilkinson shift on T, the Givens rotations applied to
// Q's columns, deflating where |e_k| <= eps (|d_k| + |d_k+1|); eigenvalues ascending with their
// columns.  O(n^3), used where Jacobi sweeps would be too slow (blocks of hundreds and more).
void tridiag_qr_eigh(int n, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w) {
    auto a = [&](int r, int c) -> double& { return A[std::size_t(r) * n + c]; };
    V.assign(std::size_t(n) * n, 0.0);
    for(int i = 0; i < n; ++i)
        V[std::size_t(i) * n + i] = 1.0;
    std::vector<double> v(n), p(n), q(n);
    // 1. A = Q T Q^T, Householder reflector P_k = I - beta v v^T zeroing column k below row k+1
    for(int k = 0; k + 2 < n; ++k) {
        int const m = n - k - 1; // trailing size
        double sigma = 0;
        for(int i = 1; i < m; ++i)
            sigma += a(k + 1 + i, k) * a(k + 1 + i, k);
        if(sigma == 0.0)
            continue;
        double const x0 = a(k + 1, k), mu = std::sqrt(x0 * x0 + sigma);
        double const v0 = x0 <= 0 ? x0 - mu : -sigma / (x0 + mu);
        double const beta = 2.0 * v0 * v0 / (sigma + v0 * v0);
        v[0] = 1.0;
        for(int i = 1; i < m; ++i)
            v[i] = a(k + 1 + i, k) / v0;
        // trailing update A22 -= v w^T + w v^T, p = beta A22 v, w = p - (beta/2)(p^T v) v
        for(int i = 0; i < m; ++i) {
            double t = 0;
            for(int j = 0; j < m; ++j)
                t += a(k + 1 + i, k + 1 + j) * v[j];
            p[i] = beta * t;
        }
        double pv = 0;
        for(int i = 0; i < m; ++i)
            pv += p[i] * v[i];
        for(int i = 0; i < m; ++i)
            q[i] = p[i] - 0.5 * beta * pv * v[i];
        for(int i = 0; i < m; ++i)
            for(int j = 0; j < m; ++j)
                a(k + 1 + i, k + 1 + j) -= v[i] * q[j] + q[i] * v[j];
        a(k + 1, k) = a(k, k + 1) = mu; // P x = |x| e_1
        for(int i = 2; i <= m; ++i)
            a(k + i, k) = a(k, k + i) = 0.0;
        // Q <- Q P_k (columns k+1 .. n-1)
        for(int r = 0; r < n; ++r) {
            double t = 0;
            for(int j = 0; j < m; ++j)
                t += V[std::size_t(r) * n + k + 1 + j] * v[j];
            t *= beta;
            for(int j = 0; j < m; ++j)
                V[std::size_t(r) * n + k + 1 + j] -= t * v[j];
        }
    }
    std::vector<double> d(n), e(n > 0 ? n : 1, 0.0);
    for(int i = 0; i < n; ++i)
        d[i] = a(i, i);
    for(int i = 0; i + 1 < n; ++i)
        e[i] = a(i + 1, i);
    // 2. implicit symmetric QR steps with the Wilkinson shift on the unreduced block [lo, hi]
    double const eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, iter = 0;
    while(hi > 0) {
        if(std::fabs(e[hi - 1]) <= eps * (std::fabs(d[hi - 1]) + std::fabs(d[hi]))) {
            e[hi - 1] = 0.0;
            --hi;
            continue;
        }
        int lo = hi - 1;
        while(lo > 0 && std::fabs(e[lo - 1]) > eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo])))
            --lo;
        if(++iter > 60 * n)
            throw std::runtime_error("tridiag_qr_eigh: no convergence");
        double const td = 0.5 * (d[hi - 1] - d[hi]), e2 = e[hi - 1] * e[hi - 1];
        double const shift = d[hi] - e2 / (td + (td >= 0 ? 1.0 : -1.0) * std::hypot(td, e[hi - 1]));
        double x = d[lo] - shift, z = e[lo];
        for(int k = lo; k < hi; ++k) {
            // Givens (c, s) with [c -s; s c]^T [x; z] = [r; 0]
            double c, s;
            if(z == 0.0) {
                c = 1.0;
                s = 0.0;
            } else if(std::fabs(z) > std::fabs(x)) {
                double const t = -x / z;
                s = 1.0 / std::sqrt(1.0 + t * t);
                c = s * t;
            } else {
                double const t = -z / x;
                c = 1.0 / std::sqrt(1.0 + t * t);
                s = c * t;
            }
            // T <- G^T T G on rows / columns k, k+1
            double const dk = d[k], dk1 = d[k + 1], ek = e[k];
            if(k > lo)
                e[k - 1] = c * x - s * z;
            d[k] = c * c * dk - 2.0 * c * s * ek + s * s * dk1;
            d[k + 1] = s * s * dk + 2.0 * c * s * ek + c * c * dk1;
            e[k] = c * s * (dk - dk1) + (c * c - s * s) * ek;
            if(k + 1 < hi) {
                x = e[k];
                z = -s * e[k + 1];
                e[k + 1] *= c;
            }
            for(int r = 0; r < n; ++r) {
                double const vk = V[std::size_t(r) * n + k], vk1 = V[std::size_t(r) * n + k + 1];
                V[std::size_t(r) * n + k] = c * vk - s * vk1;
                V[std::size_t(r) * n + k + 1] = s * vk + c * vk1;
            }
        }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x_, int y_) { return d[x_] < d[y_]; });
    w.resize(n);
    std::vector<double> V2(V.size());
    for(int j = 0; j < n; ++j) {
        w[j] = d[idx[j]];
        for(int k = 0; k < n; ++k)
            V2[std::size_t(k) * n + j] = V[std::size_t(k) * n + idx[j]];
    }
    V.swap(V2);
}

ED::ED(int n_modes_, std::vector<HTerm> H_, double beta_, int eigensolver)
    : n_modes(n_modes_), beta(beta_), H(std::move(H_)) {
    S = partition(n_modes, H);
    int nb = int(S.blocks.size());
    hmat.resize(nb);
    evecs.resize(nb);
    evals.resize(nb);
    // HamiltonianPart.cpp:56-76: H_b[:, col] = H |col>
    std::vector<std::pair<uint64_t, double>> buf;
    for(int b = 0; b < nb; ++b) {
        int n = int(S.blocks[b].size());
        hmat[b].assign(std::size_t(n) * n, 0.0);
        for(int col = 0; col < n; ++col) {
            apply_H(H, S.blocks[b][col], buf);
            for(auto const& p : buf)
                hmat[b][std::size_t(S.inner[p.first]) * n + col] += p.second;
        }
    }
    // HamiltonianPart.cpp:90-102 (eigensolve per block; n = 1 special case)
#pragma omp parallel for schedule(dynamic, 1)
    for(int b = 0; b < nb; ++b) {
        int n = int(S.blocks[b].size());
        if(n == 1) {
            evals[b] = {hmat[b][0]};
            evecs[b] = {1.0};
        } else {
            std::vector<double> A = hmat[b];
            if(eigensolver == 1)
                tridiag_qr_eigh(n, A, evecs[b], evals[b]);
            else
                jacobi_eigh(n, A, evecs[b], evals[b]);
        }
    }
    // Hamiltonian.cpp:125-130
    E0 = 1e300;
    for(auto const& e : evals)
        for(double x : e)
            E0 = std::min(E0, x);
    // DensityMatrixPart.cpp:23-31, DensityMatrix.cpp:44-52
    weights.resize(nb);
    Z = 0;
    for(int b = 0; b < nb; ++b) {
        weights[b].resize(evals[b].size());
        double zp = 0;
        for(std::size_t k = 0; k < evals[b].size(); ++k) {
            weights[b][k] = std::exp(-beta * (evals[b][k] - E0));
            zp += weights[b][k];
        }
        Z += zp;
    }
    for(auto& wb : weights)
        for(double& x : wb)
            x /= Z;
    retained.assign(nb, 1);
    cdag_cache_.resize(n_modes);
    c_cache_.resize(n_modes);
    cache_ok_.assign(n_modes, 0);
}

void ED::set_spectrum(int b, std::vector<double> const& ev, std::vector<double> const& vec, std::vector<double> const& w) {
    std::size_t n = S.blocks.at(b).size();
    if(ev.size() != n || vec.size() != n * n || w.size() != n)
        throw std::runtime_error("set_spectrum: size mismatch");
    evals[b] = ev;
    evecs[b] = vec;
    weights[b] = w;
    std::fill(cache_ok_.begin(), cache_ok_.end(), 0);
}

void ED::truncate_blocks(double tol) {
    for(std::size_t b = 0; b < weights.size(); ++b) {
        retained[b] = 0;
        for(double x : weights[b])
            if(x > tol) {
                retained[b] = 1;
                break;
            }
    }
}

// MonomialOperator.hpp:160-195 (find_connections), MonomialOperatorPart.cpp:48-92 (rotation,
// sparseView(Tol) which drops |x| <= Tol * 1e-12)
RotOp ED::rotate(Mono const& op, double tol) const {
    int nb = int(S.blocks.size());
    std::set<std::pair<int, int>> conns;
    for(int b = 0; b < nb; ++b)
        for(uint64_t s : S.blocks[b]) {
            uint64_t f;
            if(apply_mono(op, s, f))
                conns.emplace(b, S.state_block[f]);
        }
    RotOp R;
    R.left_of_right.assign(nb, -1);
    R.right_of_left.assign(nb, -1);
    for(auto const& c : conns) {
        if(R.left_of_right[c.first] != -1)
            throw std::runtime_error("operator maps a block into more than one block");
        R.left_of_right[c.first] = c.second;
        R.right_of_left[c.second] = c.first;
        OpPart P;
        P.from = c.first;
        P.to = c.second;
        P.n_from = int(S.blocks[P.from].size());
        P.n_to = int(S.blocks[P.to].size());
        int nf = P.n_from, nt = P.n_to;
        std::vector<double> X(std::size_t(nt) * nf, 0.0);
        auto const& UR = evecs[P.from];
        for(int a = 0; a < nf; ++a) {
            uint64_t f;
            int sg = apply_mono(op, S.blocks[P.from][a], f);
            if(!sg)
                continue;
            int t = S.inner[f];
            for(int j = 0; j < nf; ++j)
                X[std::size_t(t) * nf + j] += sg * UR[std::size_t(a) * nf + j];
        }
        auto const& UL = evecs[P.to];
        P.m.assign(std::size_t(nt) * nf, 0.0);
        for(int t = 0; t < nt; ++t)
            for(int i = 0; i < nt; ++i) {
                double u = UL[std::size_t(t) * nt + i];
                if(u == 0.0)
                    continue;
                double const* xr = &X[std::size_t(t) * nf];
                double* orow = &P.m[std::size_t(i) * nf];
                for(int j = 0; j < nf; ++j)
                    orow[j] += u * xr[j];
            }
        double cut = tol * 1e-12;
        for(double& x : P.m)
            if(std::fabs(x) <= cut)
                x = 0.0;
        R.parts.push_back(std::move(P));
    }
    return R;
}

RotOp ED::adjoint(RotOp const& cd) {
    RotOp R;
    R.left_of_right = cd.right_of_left;
    R.right_of_left = cd.left_of_right;
    for(auto const& P : cd.parts) {
        OpPart Q;
        Q.from = P.to;
        Q.to = P.from;
        Q.n_from = P.n_to;
        Q.n_to = P.n_from;
        Q.m.assign(P.m.size(), 0.0);
        for(int r = 0; r < P.n_to; ++r)
            for(int c = 0; c < P.n_from; ++c)
                Q.m[std::size_t(c) * Q.n_from + r] = P.m[std::size_t(r) * P.n_from + c];
        R.parts.push_back(std::move(Q));
    }
    std::sort(R.parts.begin(), R.parts.end(), [](OpPart const& a, OpPart const& b) { return a.from < b.from; });
    return R;
}

RotOp const& ED::cdag_op(int i) const {
    if(!cache_ok_[i]) {
        cdag_cache_[i] = rotate(Mono{{true, i}});
        c_cache_[i] = adjoint(cdag_cache_[i]); // FieldOperatorContainer.cpp:22-37
        cache_ok_[i] = 1;
    }
    return cdag_cache_[i];
}
RotOp const& ED::c_op(int i) const {
    cdag_op(i);
    return c_cache_[i];
}

static OpPart const* part_from_left(RotOp const& R, int left) {
    for(auto const& p : R.parts)
        if(p.to == left)
            return &p;
    return nullptr;
}
static OpPart const* part_from_right(RotOp const& R, int right) {
    for(auto const& p : R.parts)
        if(p.from == right)
            return &p;
    return nullptr;
}

// Misc.cpp:83-86
static inline int64_t bin_of(double x, double bin) { return static_cast<int64_t>(std::round(x / bin)); }
static inline std::size_t mix(std::size_t h, std::size_t v) {
    return h ^ (v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2));
}

// ---- TermList (TermList.hpp:41-135): merge a term into a similar one (same hash and KeyEqual),
// erase the sum if it is negligible with tolerance divided by (size + 1).
template <class T> class TermSet {
    std::unordered_set<T, typename T::Hash, typename T::Eq> data_;
    double tol_;
    long erased_ = 0;

public:
    TermSet(double res, double tol) : data_(1, typename T::Hash{res}, typename T::Eq{res}), tol_(tol) {}
    void add(T const& t) {
        auto it = data_.find(t);
        if(it == data_.end()) {
            data_.insert(t);
            return;
        }
        T sum = *it;
        sum += t;
        data_.erase(it);
        if(!sum.negligible(tol_, data_.size() + 1))
            data_.insert(sum);
        else
            ++erased_;
    }
    std::size_t size() const { return data_.size(); }
    long erased() const { return erased_; }
    auto begin() const { return data_.begin(); }
    auto end() const { return data_.end(); }
};

// GreensFunctionPart.hpp:75-127, GreensFunctionPart.cpp:24-36 (pole kept, residues summed)
struct GTerm {
    cplx R;
    double P;
    struct Hash {
        double res;
        std::size_t operator()(GTerm const& t) const { return std::hash<int64_t>()(bin_of(t.P, res)); }
    };
    struct Eq {
        double res;
        bool operator()(GTerm const& a, GTerm const& b) const { return std::fabs(a.P - b.P) <= res; }
    };
    GTerm& operator+=(GTerm const& o) {
        R += o.R;
        return *this;
    }
    bool negligible(double tol, std::size_t div) const { return std::abs(R) <= tol / double(div); }
};

// TwoParticleGFPart.hpp:83-142 and .cpp:35-43
struct NRTerm {
    cplx C;
    std::array<double, 3> P;
    bool isz4;
    long W;
    struct Hash {
        double res;
        std::size_t operator()(NRTerm const& t) const {
            std::size_t h = std::size_t(t.isz4);
            for(int k = 0; k < 3; ++k)
                h = mix(h, std::hash<int64_t>()(bin_of(t.P[k], res)));
            return h;
        }
    };
    struct Eq {
        double res;
        bool operator()(NRTerm const& a, NRTerm const& b) const {
            return a.isz4 == b.isz4 && std::fabs(a.P[0] - b.P[0]) <= res && std::fabs(a.P[1] - b.P[1]) <= res &&
                   std::fabs(a.P[2] - b.P[2]) <= res;
        }
    };
    NRTerm& operator+=(NRTerm const& o) {
        long cw = W + o.W;
        for(int k = 0; k < 3; ++k)
            P[k] = (double(W) * P[k] + double(o.W) * o.P[k]) / double(cw);
        W = cw;
        C += o.C;
        return *this;
    }
    bool negligible(double tol, std::size_t div) const { return std::abs(C) <= tol / double(div); }
    // TwoParticleGFPart.hpp:475-479
    cplx operator()(cplx z1, cplx z2, cplx z3) const {
        return isz4 ? C / ((z1 - P[0]) * (z1 + z2 + z3 - P[0] - P[1] - P[2]) * (z3 - P[2]))
                    : C / ((z1 - P[0]) * (z2 - P[1]) * (z3 - P[2]));
    }
};

// TwoParticleGFPart.hpp:208-268 and .cpp:86-95
struct RTerm {
    cplx R, N;
    std::array<double, 3> P;
    bool isz1z2;
    long W;
    struct Hash {
        double res;
        std::size_t operator()(RTerm const& t) const {
            std::size_t h = std::size_t(t.isz1z2) + 2;
            for(int k = 0; k < 3; ++k)
                h = mix(h, std::hash<int64_t>()(bin_of(t.P[k], res)));
            return h;
        }
    };
    struct Eq {
        double res;
        bool operator()(RTerm const& a, RTerm const& b) const {
            return a.isz1z2 == b.isz1z2 && std::fabs(a.P[0] - b.P[0]) <= res &&
                   std::fabs(a.P[1] - b.P[1]) <= res && std::fabs(a.P[2] - b.P[2]) <= res;
        }
    };
    RTerm& operator+=(RTerm const& o) {
        long cw = W + o.W;
        for(int k = 0; k < 3; ++k)
            P[k] = (double(W) * P[k] + double(o.W) * o.P[k]) / double(cw);
        W = cw;
        R += o.R;
        N += o.N;
        return *this;
    }
    bool negligible(double tol, std::size_t div) const {
        return std::abs(R) <= tol / double(div) && std::abs(N) <= tol / double(div);
    }
    // TwoParticleGFPart.hpp:481-495
    cplx operator()(cplx z1, cplx z2, cplx z3, double dtol) const {
        cplx D = isz1z2 ? (z1 + z2 - P[0] - P[1]) : (z2 + z3 - P[1] - P[2]);
        return (std::abs(D) <= dtol ? R : (N / D)) / ((z1 - P[0]) * (z3 - P[2]));
    }
};

// GreensFunction.cpp:41-86 (parts), GreensFunctionPart.cpp:65-110 (terms), GreensFunction.hpp:123-136
// GreensFunctionPart.cpp:28-31: a term R/(z - P) at imaginary time 0 <= tau <= beta
static cplx gf_term_tau(cplx R, double P, double tau, double beta) {
    return P > 0 ? -R * std::exp(-tau * P) / (1.0 + std::exp(-beta * P))
                 : -R * std::exp((beta - tau) * P) / (std::exp(beta * P) + 1.0);
}

// SusceptibilityPart.cpp:31-35: a term -R/(z - P) at imaginary time 0 <= tau <= beta
static cplx chi_term_tau(cplx R, double P, double tau, double beta) {
    return P > 0 ? -R * std::exp(-tau * P) / std::expm1(-beta * P)
                 : R * std::exp((beta - tau) * P) / std::expm1(beta * P);
}

GFResult ED::gf(int i, int j, std::vector<cplx> const& z, double pole_res, double coeff_tol,
                std::vector<double> const& tau) const {
    RotOp const& F1 = c_op(i);
    RotOp const& F2 = cdag_op(j);
    GFResult out;
    out.values.assign(z.size(), 0.0);
    out.tau_values.assign(tau.size(), 0.0); // GreensFunction.hpp:138-146
    int nb = int(S.blocks.size());
    for(int L = 0; L < nb; ++L) { // merge-join over F1.left (ascending) and F2.right
        int R = F1.right_of_left[L];
        if(R < 0 || F2.left_of_right[L] != R)
            continue;
        if(!(retained[L] || retained[R]))
            continue;
        OpPart const* p1 = part_from_left(F1, L);  // <L|c_i|R>, n_L x n_R
        OpPart const* p2 = part_from_right(F2, L); // <R|c^dag_j|L>, n_R x n_L
        TermSet<GTerm> terms(pole_res, coeff_tol);
        int nL = p1->n_to, nR = p1->n_from;
        for(int i1 = 0; i1 < nL; ++i1)
            for(int i2 = 0; i2 < nR; ++i2) {
                double a = p1->m[std::size_t(i1) * nR + i2], b = p2->m[std::size_t(i2) * nL + i1];
                if(a == 0.0 || b == 0.0)
                    continue;
                cplx Res = cplx(a) * b * (weights[L][i1] + weights[R][i2]);
                if(std::abs(Res) > coeff_tol)
                    terms.add(GTerm{Res, evals[R][i2] - evals[L][i1]});
            }
        out.n_parts++;
        out.n_terms += long(terms.size());
        for(auto const& t : terms) {
            out.poles.push_back(t.P);
            out.residues.push_back(t.R);
        }
        for(std::size_t w = 0; w < z.size(); ++w) {
            cplx s = 0;
            for(auto const& t : terms)
                s += t.R / (z[w] - t.P);
            out.values[w] += s;
        }
        for(std::size_t w = 0; w < tau.size(); ++w) {
            cplx s = 0;
            for(auto const& t : terms)
                s += gf_term_tau(t.R, t.P, tau[w], beta);
            out.tau_values[w] += s;
        }
    }
    return out;
}

cplx ED::ensemble_average(Mono const& A) const {
    RotOp Ar = rotate(A);
    cplx res = 0;
    for(auto const& p : Ar.parts) {
        if(p.from != p.to || !retained[p.to])
            continue;
        cplx rp = 0;
        for(int k = 0; k < p.n_to; ++k)
            rp += p.m[std::size_t(k) * p.n_from + k] * weights[p.to][k];
        res += rp;
    }
    return res;
}

// Susceptibility.cpp:42-119, SusceptibilityPart.cpp:24-127, SusceptibilityPart.hpp:216-220
ChiResult ED::chi(Mono const& A, Mono const& B, std::vector<cplx> const& z, double pole_res, double coeff_tol,
                  bool subtract_disconnected, std::vector<double> const& tau) const {
    RotOp Ar = rotate(A), Br = rotate(B);
    ChiResult out;
    out.values.assign(z.size(), 0.0);
    out.tau_values.assign(tau.size(), 0.0); // Susceptibility.hpp:160-170, SusceptibilityPart.hpp:222-224
    int nb = int(S.blocks.size());
    for(int L = 0; L < nb; ++L) {
        int R = Ar.right_of_left[L];
        if(R < 0 || Br.left_of_right[L] != R)
            continue;
        if(!(retained[L] || retained[R]))
            continue;
        OpPart const* pa = part_from_left(Ar, L);
        OpPart const* pb = part_from_right(Br, L);
        TermSet<GTerm> terms(pole_res, coeff_tol);
        cplx zpw = 0;
        int nL = pa->n_to, nR = pa->n_from;
        for(int i1 = 0; i1 < nL; ++i1)
            for(int i2 = 0; i2 < nR; ++i2) {
                double a = pa->m[std::size_t(i1) * nR + i2], b = pb->m[std::size_t(i2) * nL + i1];
                if(a == 0.0 || b == 0.0)
                    continue;
                double P = evals[R][i2] - evals[L][i1];
                if(std::fabs(P) < pole_res)
                    zpw += cplx(a) * b * weights[L][i1];
                else {
                    cplx Res = cplx(a) * b * (weights[L][i1] - weights[R][i2]);
                    if(std::abs(Res) > coeff_tol)
                        terms.add(GTerm{Res, P});
                }
            }
        out.n_parts++;
        out.n_terms += long(terms.size());
        out.zero_pole_weight += zpw;
        for(std::size_t w = 0; w < z.size(); ++w) {
            cplx s = 0;
            for(auto const& t : terms)
                s += -t.R / (z[w] - t.P);
            if(std::abs(z[w]) <= std::numeric_limits<double>::epsilon())
                s += zpw * beta;
            out.values[w] += s;
        }
        for(std::size_t w = 0; w < tau.size(); ++w) {
            cplx s = 0;
            for(auto const& t : terms)
                s += chi_term_tau(t.R, t.P, tau[w], beta);
            out.tau_values[w] += s + zpw;
        }
    }
    if(subtract_disconnected) {
        cplx aA = ensemble_average(A), aB = ensemble_average(B);
        for(std::size_t w = 0; w < z.size(); ++w)
            if(std::abs(z[w]) <= std::numeric_limits<double>::epsilon())
                out.values[w] -= aA * aB * beta;
        for(std::size_t w = 0; w < tau.size(); ++w)
            out.tau_values[w] -= aA * aB;
    }
    return out;
}

// Misc.cpp:41-42
static int const perm3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
static int const perm3_sign[6] = {1, -1, -1, 1, 1, -1};

// TwoParticleGF.cpp:67-120 (parts), TwoParticleGFPart.cpp:189-322 (terms, addMultiterm,
// evaluation), TwoParticleGF.cpp:132-148 (frequency fill)
G2Result ED::g2(int i, int j, int k, int l, std::vector<cplx> const& zs, double pole_res, double coeff_tol,
                int max_parts, std::vector<int> const* part_subset) const {
    RotOp const* ops[3] = {&c_op(i), &c_op(j), &cdag_op(k)};
    RotOp const& CX4 = cdag_op(l);
    std::size_t nw = zs.size() / 3;
    G2Result out;
    out.values.assign(nw, 0.0);
    int part_index = -1;
    for(auto const& cx : CX4.parts) { // ascending S1 = right block of CX4
        int S1 = cx.from, S4 = cx.to;
        for(int p = 0; p < 6; ++p) {
            RotOp const& O1 = *ops[perm3[p][0]];
            RotOp const& O2 = *ops[perm3[p][1]];
            RotOp const& O3 = *ops[perm3[p][2]];
            int S3 = O3.left_of_right[S4];
            int S2 = O1.right_of_left[S1];
            if(S2 < 0 || S3 < 0 || O2.right_of_left[S2] != S3)
                continue;
            if(!(retained[S1] || retained[S2] || retained[S3] || retained[S4]))
                continue;
            ++part_index;
            if(max_parts >= 0 && part_index >= max_parts)
                continue;
            if(part_subset && !std::binary_search(part_subset->begin(), part_subset->end(), part_index))
                continue;
            OpPart const* o1 = part_from_left(O1, S1); // n1 x n2
            OpPart const* o2 = part_from_left(O2, S2); // n2 x n3
            OpPart const* o3 = part_from_left(O3, S3); // n3 x n4
            OpPart const* c4 = &cx;                     // n4 x n1
            int n1 = int(S.blocks[S1].size()), n2 = int(S.blocks[S2].size());
            int n3 = int(S.blocks[S3].size()), n4 = int(S.blocks[S4].size());
            auto const &E1 = evals[S1], &E2 = evals[S2], &E3 = evals[S3], &E4 = evals[S4];
            auto const &W1 = weights[S1], &W2 = weights[S2], &W3 = weights[S3], &W4 = weights[S4];
            auto const t_gen = std::chrono::steady_clock::now();
            TermSet<NRTerm> nr(pole_res, coeff_tol);
            TermSet<RTerm> rs(pole_res, coeff_tol);
            double sign = perm3_sign[p];
            std::vector<int> l4;
            for(int i1 = 0; i1 < n1; ++i1)
                for(int i3 = 0; i3 < n3; ++i3) {
                    l4.clear();
                    for(int i4 = 0; i4 < n4; ++i4)
                        if(c4->m[std::size_t(i4) * n1 + i1] != 0.0 && o3->m[std::size_t(i3) * n4 + i4] != 0.0)
                            l4.push_back(i4);
                    if(l4.empty())
                        continue;
                    for(int i2 = 0; i2 < n2; ++i2) {
                        double a = o1->m[std::size_t(i1) * n2 + i2], b = o2->m[std::size_t(i2) * n3 + i3];
                        if(a == 0.0 || b == 0.0)
                            continue;
                        for(int i4 : l4) {
                            if(!(W1[i1] + W2[i2] + W3[i3] + W4[i4] > coeff_tol))
                                continue;
                            ++out.n_quadruples;
                            cplx C = cplx(a) * b * o3->m[std::size_t(i3) * n4 + i4] * c4->m[std::size_t(i4) * n1 + i1];
                            C *= sign;
                            // addMultiterm (TwoParticleGFPart.cpp:267-298)
                            double P1 = E2[i2] - E1[i1], P2 = E3[i3] - E2[i2], P3 = E4[i4] - E3[i3];
                            double wi = W1[i1], wj = W2[i2], wk = W3[i3], wl = W4[i4];
                            cplx cz2 = -C * (wj + wk);
                            if(std::abs(cz2) > coeff_tol)
                                nr.add(NRTerm{cz2, {P1, P2, P3}, false, 1});
                            cplx cz4 = C * (wi + wl);
                            if(std::abs(cz4) > coeff_tol)
                                nr.add(NRTerm{cz4, {P1, P2, P3}, true, 1});
                            cplx r12 = C * beta * wi, n12 = C * (wk - wi);
                            if(std::abs(r12) > coeff_tol || std::abs(n12) > coeff_tol)
                                rs.add(RTerm{r12, n12, {P1, P2, P3}, true, 1});
                            cplx r23 = -C * beta * wj, n23 = C * (wj - wl);
                            if(std::abs(r23) > coeff_tol || std::abs(n23) > coeff_tol)
                                rs.add(RTerm{r23, n23, {P1, P2, P3}, false, 1});
                        }
                    }
                }
            out.n_parts++;
            out.n_nonres += long(nr.size());
            out.n_res += long(rs.size());
            out.n_erased += nr.erased() + rs.erased();
            out.part_nonres.push_back(long(nr.size()));
            out.part_res.push_back(long(rs.size()));
            for(int x : {S1, S2, S3, S4, p})
                out.part_blocks.push_back(x);
            std::vector<NRTerm> vnr(nr.begin(), nr.end());
            std::vector<RTerm> vr(rs.begin(), rs.end());
            auto const t_eval = std::chrono::steady_clock::now();
            out.gen_seconds += std::chrono::duration<double>(t_eval - t_gen).count();
#pragma omp parallel for schedule(static)
            for(long w = 0; w < long(nw); ++w) {
                cplx F[3] = {zs[3 * w], zs[3 * w + 1], -zs[3 * w + 2]};
                cplx z1 = F[perm3[p][0]], z2 = F[perm3[p][1]], z3 = F[perm3[p][2]];
                cplx s = 0;
                for(auto const& t : vnr)
                    s += t(z1, z2, z3);
                cplx s2 = 0;
                for(auto const& t : vr)
                    s2 += t(z1, z2, z3, pole_res);
                out.values[w] += s + s2;
            }
            out.eval_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_eval).count();
        }
    }
    return out;
}

// ---- 3-point susceptibility ------------------------------------------------------------------
// ThreePointSusceptibilityPart.hpp:66-168 and .cpp:34-43: fermion-fermion term
// C / ((z1 - P0)(z2 - P1)), merged on (bin P0, bin P1), count-weighted mean poles.
struct FFTerm {
    cplx C;
    std::array<double, 2> P;
    long W;
    struct Hash {
        double res;
        std::size_t operator()(FFTerm const& t) const {
            return mix(std::hash<int64_t>()(bin_of(t.P[0], res)), std::hash<int64_t>()(bin_of(t.P[1], res)));
        }
    };
    struct Eq {
        double res;
        bool operator()(FFTerm const& a, FFTerm const& b) const {
            return std::fabs(a.P[0] - b.P[0]) <= res && std::fabs(a.P[1] - b.P[1]) <= res;
        }
    };
    FFTerm& operator+=(FFTerm const& o) {
        long cw = W + o.W;
        for(int k = 0; k < 2; ++k)
            P[k] = (double(W) * P[k] + double(o.W) * o.P[k]) / double(cw);
        W = cw;
        C += o.C;
        return *this;
    }
    bool negligible(double tol, std::size_t div) const { return std::abs(C) <= tol / double(div); }
    // ThreePointSusceptibilityPart.hpp:527-529
    cplx operator()(cplx z1, cplx z2) const { return C / ((z1 - P[0]) * (z2 - P[1])); }
};

// ThreePointSusceptibilityPart.hpp:170-280 and .cpp:85-95: fermion-boson term
// C / ((z1 - P1)(z1 - xi z2 - P12)), merged on (xi, bin P1, bin P12).
struct FBTerm {
    cplx C;
    double P1, P12;
    int xi;
    long W;
    struct Hash {
        double res;
        std::size_t operator()(FBTerm const& t) const {
            std::size_t h = std::size_t(t.xi + 2);
            h = mix(h, std::hash<int64_t>()(bin_of(t.P1, res)));
            return mix(h, std::hash<int64_t>()(bin_of(t.P12, res)));
        }
    };
    struct Eq {
        double res;
        bool operator()(FBTerm const& a, FBTerm const& b) const {
            return a.xi == b.xi && std::fabs(a.P1 - b.P1) <= res && std::fabs(a.P12 - b.P12) <= res;
        }
    };
    FBTerm& operator+=(FBTerm const& o) {
        long cw = W + o.W;
        P1 = (double(W) * P1 + double(o.W) * o.P1) / double(cw);
        P12 = (double(W) * P12 + double(o.W) * o.P12) / double(cw);
        W = cw;
        C += o.C;
        return *this;
    }
    bool negligible(double tol, std::size_t div) const { return std::abs(C) <= tol / double(div); }
    // ThreePointSusceptibilityPart.hpp:531-533
    cplx operator()(cplx z1, cplx z2) const { return C / ((z1 - P1) * (z1 - double(xi) * z2 - P12)); }
};

// ThreePointSusceptibilityPart.hpp:282-370 and .cpp:141-149: resonant term, C / (z1 - P) where
// |z1 - xi z2| <= tolerance, merged on (xi, bin P).
struct R3Term {
    cplx C;
    double P;
    int xi;
    long W;
    struct Hash {
        double res;
        std::size_t operator()(R3Term const& t) const {
            return mix(std::size_t(t.xi + 2), std::hash<int64_t>()(bin_of(t.P, res)));
        }
    };
    struct Eq {
        double res;
        bool operator()(R3Term const& a, R3Term const& b) const { return a.xi == b.xi && std::fabs(a.P - b.P) <= res; }
    };
    R3Term& operator+=(R3Term const& o) {
        long cw = W + o.W;
        P = (double(W) * P + double(o.W) * o.P) / double(cw);
        W = cw;
        C += o.C;
        return *this;
    }
    bool negligible(double tol, std::size_t div) const { return std::abs(C) <= tol / double(div); }
    // ThreePointSusceptibilityPart.hpp:535-540
    cplx operator()(cplx z1, cplx z2, double dtol) const {
        return (std::abs(z1 - double(xi) * z2) <= dtol) ? (C / (z1 - P)) : cplx(0.0);
    }
};

// ThreePointSusceptibility.cpp:37-66 (operator roles per channel), :68-142 (parts),
// ThreePointSusceptibilityPart.cpp:243-348 (terms, addMultiterm), .cpp:350-359 and
// ThreePointSusceptibility.hpp:162-177 (evaluation).
Chi3Result ED::chi3(int channel, int i, int j, int k, int l, std::vector<cplx> const& zs, double pole_res,
                    double coeff_tol) const {
    RotOp const& CX1 = cdag_op(i);
    RotOp const& C2 = c_op(j);
    RotOp const& CX3 = cdag_op(k);
    RotOp const& C4 = c_op(l);
    bool const pp = channel == 0;
    RotOp const& F1 = CX1;
    RotOp const& F2 = channel == 0 ? CX3 : (channel == 1 ? C2 : C4);
    RotOp const& B1 = channel == 0 ? C2 : CX3;
    RotOp const& B2 = channel == 2 ? C2 : C4;
    int const xi = pp ? -1 : 1;
    std::size_t nw = zs.size() / 2;
    Chi3Result out;
    out.values.assign(nw, 0.0);

    // one part: <1|Fa|2> <2|Fb|3> <3|B1 B2|1>, blocks (S1, S2, S3)
    auto run_part = [&](OpPart const* fa, OpPart const* fb, OpPart const* b1, OpPart const* b2, int S1, int S2, int S3,
                        bool sw) {
        int n1 = int(S.blocks[S1].size()), n2 = int(S.blocks[S2].size()), n3 = int(S.blocks[S3].size());
        int nm = b1->n_from;
        std::vector<double> B(std::size_t(n3) * n1, 0.0);
        for(int r = 0; r < n3; ++r)
            for(int m = 0; m < nm; ++m) {
                double a = b1->m[std::size_t(r) * nm + m];
                if(a == 0.0)
                    continue;
                for(int c = 0; c < n1; ++c)
                    B[std::size_t(r) * n1 + c] += a * b2->m[std::size_t(m) * n1 + c];
            }
        auto const &E1 = evals[S1], &E2 = evals[S2], &E3 = evals[S3];
        auto const &W1 = weights[S1], &W2 = weights[S2], &W3 = weights[S3];
        double const prefactor = ((channel == 1) == sw) ? 1.0 : -1.0;
        TermSet<FFTerm> ff(pole_res, coeff_tol);
        TermSet<FBTerm> fbs(pole_res, coeff_tol);
        TermSet<R3Term> rs(pole_res, coeff_tol);
        for(int i1 = 0; i1 < n1; ++i1)
            for(int i2 = 0; i2 < n2; ++i2) {
                double x = fa->m[std::size_t(i1) * n2 + i2];
                if(x == 0.0)
                    continue;
                for(int i3 = 0; i3 < n3; ++i3) {
                    double y = fb->m[std::size_t(i2) * n3 + i3], b = B[std::size_t(i3) * n1 + i1];
                    if(y == 0.0 || b == 0.0)
                        continue;
                    ++out.n_triples;
                    cplx C = cplx(x * y * b);
                    C *= prefactor;
                    // addMultiterm (ThreePointSusceptibilityPart.cpp:298-348)
                    double Ei = E1[i1], Ej = E2[i2], Ek = E3[i3], Wi = W1[i1], Wj = W2[i2], Wk = W3[i3];
                    double Eij = Ei - Ej, Ejk = Ej - Ek, Eik = Ei - Ek;
                    cplx cff = C * (Wi + Wj);
                    if(std::abs(cff) > coeff_tol)
                        ff.add(FFTerm{pp ? -cff : cff, {sw ? Ejk : Eij, (sw ? Eij : Ejk) * (pp ? 1.0 : -1.0)}, 1});
                    if(std::fabs(Eik) <= pole_res) {
                        cplx cr = -C * beta * Wi;
                        if(std::abs(cr) > coeff_tol)
                            rs.add(R3Term{sw ? -cr : cr, sw ? Ejk : -Ejk, xi, 1});
                    } else {
                        cplx cfb = C * (Wk - Wi);
                        if(std::abs(cfb) > coeff_tol) {
                            fbs.add(FBTerm{sw ? -cfb : cfb, sw ? Ejk : Eij, Eik, xi, 1});
                            if(!sw)
                                ff.add(FFTerm{pp ? -cfb : cfb, {Eij, pp ? Ejk : -Ejk}, 1});
                        }
                    }
                }
            }
        out.n_parts++;
        out.n_ff += long(ff.size());
        out.n_fb += long(fbs.size());
        out.n_res += long(rs.size());
        out.part_ff.push_back(long(ff.size()));
        out.part_fb.push_back(long(fbs.size()));
        out.part_res.push_back(long(rs.size()));
        for(int v : {S1, S2, S3, int(sw)})
            out.part_blocks.push_back(v);
        std::vector<FFTerm> vff(ff.begin(), ff.end());
        std::vector<FBTerm> vfb(fbs.begin(), fbs.end());
        std::vector<R3Term> vr(rs.begin(), rs.end());
#pragma omp parallel for schedule(static)
        for(long w = 0; w < long(nw); ++w) {
            cplx z1 = zs[2 * w], z2 = zs[2 * w + 1];
            cplx s = 0;
            for(auto const& t : vff)
                s += t(z1, z2);
            for(auto const& t : vfb)
                s += t(z1, z2);
            for(auto const& t : vr)
                s += t(z1, z2, pole_res);
            out.values[w] += s;
        }
    };

    for(auto const& b2 : B2.parts) { // ascending right block of B2
        int B2right = b2.from, B2left = b2.to;
        int B1left = B1.left_of_right[B2left];
        if(B1left < 0)
            continue;
        OpPart const* b1 = part_from_left(B1, B1left);
        // F1, F2, B term: <B2right|F1|F1right><F2left|F2|B1left><B1left|B1|B2left><B2left|B2|B2right>
        int F1right = F1.right_of_left[B2right];
        int F2left = F2.left_of_right[B1left];
        if(F1right == F2left && F1right >= 0 && retained[B2right] && retained[F1right] && retained[B1left])
            run_part(part_from_left(F1, B2right), part_from_left(F2, F2left), b1, &b2, B2right, F1right, B1left, false);
        // F2, F1, B term
        int F2right = F2.right_of_left[B2right];
        int F1left = F1.left_of_right[B1left];
        if(F2right == F1left && F2right >= 0 && retained[B2right] && retained[F2right] && retained[B1left])
            run_part(part_from_left(F2, B2right), part_from_left(F1, F1left), b1, &b2, B2right, F2right, B1left, true);
    }
    return out;
}

} // namespace oracle
