// Fermionic second-quantised operator expressions (reference: include/pomerol/Operators.hpp:52-59,
// which re-exports libcommute's expression, c, c_dag, n and hc).
//
// An expression is a sum of monomials (ordered products of c / c^dag, the rightmost acting first)
// with real or complex coefficients, kept in the canonical normal order of libcommute's
// expressions: creation operators first in ascending index order, then annihilation operators in
// descending index order, reached with the anticommutation relations {c_i, c^dag_j} = delta_ij,
// c_i c_i = c^dag_i c^dag_i = 0.  Equal operators are therefore equal expressions (Sz * N ==
// N * Sz), and monomials that cancel disappear.
#pragma once

#include "Misc.hpp"

#include <map>
#include <ostream>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace Pomerol {
namespace Operators {

template <typename... IndexTypes> struct FermionOp {
    bool dag;
    std::tuple<IndexTypes...> index;
    bool operator<(FermionOp const& o) const { return std::tie(dag, index) < std::tie(o.dag, o.index); }
    bool operator==(FermionOp const& o) const { return dag == o.dag && index == o.index; }
};

struct hc_tag {};
/// Hermitian-conjugate placeholder: `expr + hc` adds the Hermitian conjugate of `expr`.
static constexpr hc_tag hc{};

template <typename ScalarType, typename... IndexTypes> class expression {
public:
    using scalar_type = ScalarType;
    using index_type = std::tuple<IndexTypes...>;
    using op_type = FermionOp<IndexTypes...>;
    using monomial_type = std::vector<op_type>;

private:
    std::map<monomial_type, ScalarType> terms_;

    // canonical position: creation operators before annihilation operators; ascending indices
    // among the former, descending among the latter
    static bool before(op_type const& a, op_type const& b) {
        if(a.dag != b.dag)
            return a.dag;
        return a.dag ? a.index < b.index : b.index < a.index;
    }

    // add c * m after bringing m to canonical order (an adjacent out-of-order pair x y becomes
    // -y x, plus delta for x = c_i, y = c^dag_i; a repeated operator gives zero)
    void add_canonical(monomial_type m, ScalarType c) {
        for(std::size_t k = 0; k + 1 < m.size(); ++k) {
            op_type const x = m[k], y = m[k + 1];
            if(x == y)
                return; // c_i c_i = c^dag_i c^dag_i = 0
            if(before(y, x)) {
                if(x.index == y.index) { // x = c_i, y = c^dag_i: c_i c^dag_i = 1 - c^dag_i c_i
                    monomial_type contracted;
                    contracted.insert(contracted.end(), m.begin(), m.begin() + std::ptrdiff_t(k));
                    contracted.insert(contracted.end(), m.begin() + std::ptrdiff_t(k + 2), m.end());
                    add_canonical(contracted, c);
                }
                std::swap(m[k], m[k + 1]);
                add_canonical(m, -c);
                return;
            }
        }
        add(m, c);
    }

    void add(monomial_type const& m, ScalarType c) {
        auto it = terms_.find(m);
        if(it == terms_.end()) {
            if(c != ScalarType(0))
                terms_.emplace(m, c);
        } else {
            it->second += c;
            if(it->second == ScalarType(0))
                terms_.erase(it);
        }
    }

public:
    expression() = default;
    expression(ScalarType c) { add({}, c); } // NOLINT
    expression(ScalarType c, monomial_type const& m) { add_canonical(m, c); }
    template <typename S2, typename = typename std::enable_if<!std::is_same<S2, ScalarType>::value>::type>
    expression(expression<S2, IndexTypes...> const& o) { // real -> complex promotion
        for(auto const& t : o.get_monomials())
            add(t.first, ScalarType(t.second));
    }

    std::map<monomial_type, ScalarType> const& get_monomials() const { return terms_; }
    std::size_t size() const { return terms_.size(); }

    expression& operator+=(expression const& o) {
        for(auto const& t : o.terms_)
            add(t.first, t.second);
        return *this;
    }
    expression& operator-=(expression const& o) {
        for(auto const& t : o.terms_)
            add(t.first, -t.second);
        return *this;
    }
    expression& operator*=(ScalarType c) {
        std::map<monomial_type, ScalarType> t2;
        for(auto const& t : terms_)
            if(t.second * c != ScalarType(0))
                t2.emplace(t.first, t.second * c);
        terms_.swap(t2);
        return *this;
    }
    expression operator-() const {
        expression r(*this);
        r *= ScalarType(-1);
        return r;
    }
    friend expression operator+(expression a, expression const& b) { return a += b; }
    friend expression operator-(expression a, expression const& b) { return a -= b; }
    friend expression operator*(expression const& a, expression const& b) {
        expression r;
        for(auto const& x : a.terms_)
            for(auto const& y : b.terms_) {
                monomial_type m = x.first;
                m.insert(m.end(), y.first.begin(), y.first.end());
                r.add_canonical(m, x.second * y.second);
            }
        return r;
    }
    friend expression operator*(ScalarType c, expression a) { return a *= c; }
    friend expression operator*(expression a, ScalarType c) { return a *= c; }
    friend expression operator+(expression const& a, hc_tag) { return a + a.hc(); }
    friend bool operator==(expression const& a, expression const& b) { return a.terms_ == b.terms_; }
    friend bool operator!=(expression const& a, expression const& b) { return !(a == b); }

    /// Hermitian conjugate.
    expression hc() const {
        expression r;
        for(auto const& t : terms_) {
            monomial_type m;
            for(auto it = t.first.rbegin(); it != t.first.rend(); ++it)
                m.push_back(op_type{!it->dag, it->index});
            r.add_canonical(m, conj_(t.second));
        }
        return r;
    }

    friend std::ostream& operator<<(std::ostream& os, expression const& e) {
        bool first = true;
        for(auto const& t : e.terms_) {
            if(!first)
                os << " + ";
            first = false;
            os << t.second;
            for(auto const& o : t.first) {
                os << "*C" << (o.dag ? "+" : "") << "(";
                print_tuple(os, o.index);
                os << ")";
            }
        }
        if(first)
            os << "0";
        return os;
    }

private:
    static RealType conj_(RealType x) { return x; }
    static ComplexType conj_(ComplexType x) { return std::conj(x); }
    template <std::size_t K = 0> static void print_tuple(std::ostream& os, index_type const& t) {
        if constexpr(K < sizeof...(IndexTypes)) {
            if(K)
                os << ",";
            os << std::get<K>(t);
            print_tuple<K + 1>(os, t);
        }
    }
};

template <typename T> struct norm_index { using type = T; };
template <> struct norm_index<char const*> { using type = std::string; };
template <> struct norm_index<char*> { using type = std::string; };
template <typename T> using norm_index_t = typename norm_index<typename std::decay<T>::type>::type;

template <typename ScalarType = double, typename... Args> expression<ScalarType, norm_index_t<Args>...> c_dag(Args&&... args) {
    using E = expression<ScalarType, norm_index_t<Args>...>;
    return E(ScalarType(1), {typename E::op_type{true, typename E::index_type(norm_index_t<Args>(args)...)}});
}
template <typename ScalarType = double, typename... Args> expression<ScalarType, norm_index_t<Args>...> c(Args&&... args) {
    using E = expression<ScalarType, norm_index_t<Args>...>;
    return E(ScalarType(1), {typename E::op_type{false, typename E::index_type(norm_index_t<Args>(args)...)}});
}
template <typename ScalarType = double, typename... Args> expression<ScalarType, norm_index_t<Args>...> n(Args&&... args) {
    return c_dag<ScalarType>(args...) * c<ScalarType>(args...);
}

/// Total particle number of a list of single-particle indices.
template <typename... IndexTypes> expression<double, IndexTypes...> N(std::vector<std::tuple<IndexTypes...>> const& v) {
    expression<double, IndexTypes...> r;
    for(auto const& t : v)
        r += std::apply([](auto const&... a) { return n<double>(a...); }, t);
    return r;
}
/// S_z = (N_up - N_down) / 2.
template <typename... IndexTypes>
expression<double, IndexTypes...> Sz(std::vector<std::tuple<IndexTypes...>> const& up, std::vector<std::tuple<IndexTypes...>> const& dn) {
    return 0.5 * (N(up) - N(dn));
}

} // namespace Operators
} // namespace Pomerol
