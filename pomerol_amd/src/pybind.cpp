// Thin Python binding of the pomerol-compatible C++ API (tests, benchmark, autograd wrapper).
//
// An ED session builds the model expression with the C++ operator DSL from a term list
// [(coeff, [(dagger, (label, orbital, spin)), ...]), ...] and then runs the API exactly as a
// pomerol program would: IndexClassification -> HilbertSpace -> StatesClassification ->
// Hamiltonian -> DensityMatrix -> FieldOperatorContainer -> GreensFunction / TwoParticleGF /
// Susceptibility.  Single-particle indices are ParticleIndex values (sorted index tuples).
#include "pomerol.hpp"

#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <map>
#include <memory>
#include <set>

namespace py = pybind11;
using namespace Pomerol;
using Expr = LatticePresets::RealExpr;
using IndexInfoT = IndexClassification<std::string, unsigned short, LatticePresets::spin>;
using HST = HilbertSpace<std::string, unsigned short, LatticePresets::spin>;

namespace {

Expr build_expr(py::list terms) {
    Expr H;
    for(auto t : terms) {
        auto tt = t.cast<py::tuple>();
        double coeff = tt[0].cast<double>();
        Expr::monomial_type m;
        for(auto o : tt[1].cast<py::list>()) {
            auto ot = o.cast<py::tuple>();
            auto idx = ot[1].cast<py::tuple>();
            m.push_back(Expr::op_type{ot[0].cast<bool>(),
                                      Expr::index_type(idx[0].cast<std::string>(), idx[1].cast<unsigned short>(),
                                                       LatticePresets::spin(idx[2].cast<int>()))});
        }
        H += Expr(coeff, m);
    }
    return H;
}

py::array_t<double> to_np(std::vector<double> const& v) { return py::array_t<double>(v.size(), v.data()); }
py::array_t<std::complex<double>> to_np(std::vector<ComplexType> const& v) {
    return py::array_t<std::complex<double>>(v.size(), v.data());
}
py::array_t<double> to_np(MatrixType<false> const& m) {
    py::array_t<double> a({m.rows(), m.cols()});
    std::copy(m.data(), m.data() + m.size(), a.mutable_data());
    return a;
}

struct ED {
    Expr HExpr;
    double beta;
    std::unique_ptr<IndexInfoT> info;
    std::unique_ptr<HST> HS;
    std::unique_ptr<StatesClassification> S;
    std::unique_ptr<Hamiltonian> H;
    std::unique_ptr<DensityMatrix> rho;
    std::unique_ptr<FieldOperatorContainer> ops;
    std::map<BlockNumber, MatrixType<false>> hmats; // Hamiltonian blocks kept before the eigensolve
    double t_eigh = 0, t_rot = 0, t_keep = 0;

    // keep: 0 no Hamiltonian block kept, 1 every block, 2 the largest block only (host copies taken
    // before the eigensolve overwrites them; their time is t_keep)
    ED(py::list terms, double beta, int keep) : HExpr(build_expr(terms)), beta(beta) {
        info.reset(new IndexInfoT(MakeIndexClassification(HExpr)));
        HS.reset(new HST(*info, HExpr));
        HS->compute();
        S.reset(new StatesClassification());
        S->compute(*HS);
        H.reset(new Hamiltonian(*S));
        H->prepare(HExpr, *HS);
        auto const tk = std::chrono::steady_clock::now();
        BlockNumber big = 0;
        for(BlockNumber b = 0; b < S->getNumberOfBlocks(); ++b)
            if(S->getBlockSize(b) > S->getBlockSize(big))
                big = b;
        for(BlockNumber b = 0; b < S->getNumberOfBlocks(); ++b)
            if(keep == 1 || (keep == 2 && b == big))
                hmats.emplace(b, H->getPart(b).getMatrix<false>());
        t_keep = std::chrono::duration<double>(std::chrono::steady_clock::now() - tk).count();
        auto t0 = std::chrono::steady_clock::now();
        H->compute();
        t_eigh = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        rho.reset(new DensityMatrix(*S, *H, beta));
        rho->prepare();
        rho->compute();
    }

    std::set<ParticleIndex> field_subset; // empty: every index (FieldOperatorContainer's default)

    FieldOperatorContainer& field_ops() {
        if(!ops) {
            auto t0 = std::chrono::steady_clock::now();
            ops.reset(new FieldOperatorContainer(*info, *HS, *S, *H, field_subset));
            ops->prepareAll(*HS);
            ops->computeAll();
            t_rot = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return *ops;
    }

    // Rotated monomial operators, computed once per session (susceptibilities share them).
    std::map<std::string, std::unique_ptr<MonomialOperator>> mono_cache;
    MonomialOperator& mono(py::list ops_) {
        std::string key = py::str(ops_);
        auto it = mono_cache.find(key);
        if(it == mono_cache.end()) {
            auto op = std::make_unique<MonomialOperator>(mono_expr(ops_), *HS, *S, *H);
            op->prepare(*HS);
            op->compute();
            it = mono_cache.emplace(key, std::move(op)).first;
        }
        return *it->second;
    }

    Expr mono_expr(py::list ops_) {
        Expr::monomial_type m;
        for(auto o : ops_) {
            auto ot = o.cast<py::tuple>();
            auto const& t = info->getInfo(ot[1].cast<ParticleIndex>());
            m.push_back(Expr::op_type{ot[0].cast<bool>(), t});
        }
        return Expr(1.0, m);
    }
};

// A TwoParticleGFContainer over an ED session (prepareAll on construction).
struct G2Box {
    ED& e;
    std::unique_ptr<TwoParticleGFContainer> C;
    std::vector<IndexCombination4> elems; // computed elements in the container's order

    G2Box(ED& e, std::vector<std::tuple<int, int, int, int>> const& idx, double pr, double ct) : e(e) {
        auto& ops = e.field_ops();
        C.reset(new TwoParticleGFContainer(*e.info, *e.S, *e.H, *e.rho, ops));
        C->PoleResolution = pr;
        C->CoefficientTolerance = ct;
        std::set<IndexCombination4> s;
        for(auto const& t : idx)
            s.emplace(ParticleIndex(std::get<0>(t)), ParticleIndex(std::get<1>(t)), ParticleIndex(std::get<2>(t)),
                      ParticleIndex(std::get<3>(t)));
        C->prepareAll(s);
        for(auto const& kv : C->getNonTrivialElements())
            elems.push_back(kv.first);
    }
    TwoParticleGF& element(std::size_t k) { return *C->getNonTrivialElements().at(elems.at(k)); }
};

FreqVec3 to_freqs(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> zs) {
    if(zs.ndim() != 2 || zs.shape(1) != 3)
        throw std::runtime_error("frequency triplets must have shape (N, 3)");
    FreqVec3 f;
    auto r = zs.unchecked<2>();
    for(py::ssize_t w = 0; w < zs.shape(0); ++w)
        f.emplace_back(r(w, 0), r(w, 1), r(w, 2));
    return f;
}

FreqVec2 to_pairs(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> zs) {
    if(zs.ndim() != 2 || zs.shape(1) != 2)
        throw std::runtime_error("frequency pairs must have shape (N, 2)");
    FreqVec2 f;
    auto r = zs.unchecked<2>();
    for(py::ssize_t w = 0; w < zs.shape(0); ++w)
        f.emplace_back(r(w, 0), r(w, 1));
    return f;
}

std::vector<ComplexType> to_points(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> z) {
    return std::vector<ComplexType>(z.data(), z.data() + z.size());
}

} // namespace

PYBIND11_MODULE(_pomerol, m) {
    m.doc() = "MI355X full-ED path behind pomerol's C++ API";
    m.def("eval_timing", [](int enable) {
        double ms = 0;
        int64_t n = 0;
        detail::check(pomx_eval_timing(detail::gpu(), enable, &ms, &n), "timing");
        return py::make_tuple(ms, n);
    }, py::arg("enable") = -1);
    m.def("kernel_stats", []() {
        // per-class statistics since eval_timing(1): launches, ms, flop, term_freqs, bytes
        pomx_kernel_stat st[POMX_STAT_NCLASS];
        detail::check(pomx_kernel_stats(detail::gpu(), st), "kernel statistics");
        static char const* names[POMX_STAT_NCLASS] = {"g2_sep", "g2_diag", "g2_grouped", "g2_general", "g2_reduce",
                                                      "pair", "g2_gen", "g2_sort"};
        py::dict d;
        for(int c = 0; c < POMX_STAT_NCLASS; ++c) {
            py::dict x;
            x["launches"] = st[c].launches;
            x["ms"] = st[c].ms;
            x["flop"] = st[c].flop;
            x["term_freqs"] = st[c].term_freqs;
            x["bytes"] = st[c].bytes;
            d[names[c]] = x;
        }
        return d;
    });
    m.def("device_ok", []() {
        try {
            detail::gpu();
            return true;
        } catch(std::exception const&) {
            return false;
        }
    });

    py::class_<BatchSolver>(m, "Batch")
        .def(py::init([](py::list models, std::vector<double> betas, py::tuple idx, int nfreq, double tol) {
                 std::vector<Expr> Hs;
                 for(auto t : models)
                     Hs.push_back(build_expr(t.cast<py::list>()));
                 BatchSolver::Index index(idx[0].cast<std::string>(), idx[1].cast<unsigned short>(),
                                          LatticePresets::spin(idx[2].cast<int>()));
                 return new BatchSolver(std::move(Hs), std::move(betas), index, nfreq, tol);
             }),
             py::arg("models"), py::arg("betas"), py::arg("index"), py::arg("nfreq"), py::arg("coeff_tol") = 1e-8)
        .def(py::init([](py::list param_models, std::vector<std::vector<double>> theta, std::vector<double> betas,
                         py::tuple idx, int nfreq, double tol) {
                 std::vector<std::vector<Expr>> ops;
                 for(auto inst : param_models) {
                     ops.emplace_back();
                     for(auto t : inst.cast<py::list>())
                         ops.back().push_back(build_expr(t.cast<py::list>()));
                 }
                 BatchSolver::Index index(idx[0].cast<std::string>(), idx[1].cast<unsigned short>(),
                                          LatticePresets::spin(idx[2].cast<int>()));
                 return new BatchSolver(std::move(ops), std::move(theta), std::move(betas), index, nfreq, tol);
             }),
             py::arg("param_models"), py::arg("theta"), py::arg("betas"), py::arg("index"), py::arg("nfreq"),
             py::arg("coeff_tol") = 1e-8)
        .def("set_params", &BatchSolver::setParameters)
        .def("gradient",
             [](BatchSolver& b, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> gbar) {
                 if(gbar.ndim() != 2)
                     throw std::runtime_error("gradient: Gbar must be (instances, nfreq)");
                 auto r = gbar.unchecked<2>();
                 std::vector<std::vector<ComplexType>> g(std::size_t(r.shape(0)));
                 for(std::size_t k = 0; k < g.size(); ++k)
                     for(py::ssize_t n = 0; n < r.shape(1); ++n)
                         g[k].push_back(r(py::ssize_t(k), n));
                 std::vector<std::vector<double>> d;
                 {
                     py::gil_scoped_release rel;
                     d = b.gradient(g);
                 }
                 return d;
             })
        .def("prepare", &BatchSolver::prepare)
        .def("solve", &BatchSolver::solve, py::arg("use_graph") = true)
        .def("autotune", &BatchSolver::autotune)
        .def("size", &BatchSolver::size)
        .def("n_structures", &BatchSolver::numStructures)
        .def("n_blocks", &BatchSolver::numBlocks)
        .def("E0", [](BatchSolver& b) { return to_np(b.getGroundEnergies()); })
        .def("G", [](BatchSolver& b) {
            auto g = b.getG();
            py::array_t<std::complex<double>> a({g.size(), g.empty() ? 0 : g[0].size()});
            auto r = a.mutable_unchecked<2>();
            for(std::size_t k = 0; k < g.size(); ++k)
                for(std::size_t n = 0; n < g[k].size(); ++n)
                    r(k, n) = g[k][n];
            return a;
        });

    py::class_<G2Box>(m, "G2Container")
        .def(py::init([](ED& e, std::vector<std::tuple<int, int, int, int>> idx, double pr, double ct) {
                 return new G2Box(e, idx, pr, ct);
             }),
             py::arg("ed"), py::arg("indices"), py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-16,
             py::keep_alive<1, 2>())
        .def("elements", [](G2Box& b) {
            std::vector<std::tuple<int, int, int, int>> out;
            for(auto const& I : b.elems)
                out.emplace_back(int(I.Index1), int(I.Index2), int(I.Index3), int(I.Index4));
            return out;
        })
        .def("parts", [](G2Box& b, std::size_t k) {
            // (S1, S2, S3, S4, permutation, n1, n2, n3, n4) per part of computed element k
            std::vector<std::vector<int64_t>> out;
            for(auto const& p : b.element(k).getParts()) {
                auto bl = p.getBlocks();
                int perm = 0;
                for(int pp = 0; pp < 6; ++pp)
                    if(permutations3[pp] == p.getPermutation())
                        perm = pp;
                out.push_back({bl[0], bl[1], bl[2], bl[3], perm, int64_t(b.e.S->getBlockSize(bl[0])),
                               int64_t(b.e.S->getBlockSize(bl[1])), int64_t(b.e.S->getBlockSize(bl[2])),
                               int64_t(b.e.S->getBlockSize(bl[3]))});
            }
            return out;
        })
        .def("compute_parts", [](G2Box& b, std::size_t k, std::vector<std::size_t> which,
                                 py::array_t<std::complex<double>> zs) {
            // TwoParticleGF::computeParts of computed element k (terms cleared after the fill)
            TwoParticleGF& G = b.element(k);
            FreqVec3 f = to_freqs(zs);
            auto t0 = std::chrono::steady_clock::now();
            std::vector<ComplexType> vals = G.computeParts(which, true, f);
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            py::dict d;
            d["values"] = to_np(vals);
            d["seconds"] = secs;
            std::vector<int64_t> pnr, pr_, praw;
            for(std::size_t q : which) {
                auto const& p = G.getParts().at(q);
                pnr.push_back(int64_t(p.getNumNonResonantTerms()));
                pr_.push_back(int64_t(p.getNumResonantTerms()));
                praw.push_back(p.getNumRawTerms());
            }
            d["part_nonres"] = pnr;
            d["part_res"] = pr_;
            d["part_raw"] = praw;
            return d;
        })
        .def("compute_all", [](G2Box& b, py::array_t<std::complex<double>> zs, bool clear, bool split) {
            auto all = b.C->computeAll(clear, to_freqs(zs), MPI_COMM_WORLD, split);
            py::dict d;
            for(auto const& kv : all) {
                auto const& I = kv.first;
                d[py::make_tuple(int(I.Index1), int(I.Index2), int(I.Index3), int(I.Index4))] = to_np(kv.second);
            }
            return d;
        }, py::arg("zs"), py::arg("clear") = true, py::arg("split") = true)
        .def("value", [](G2Box& b, std::tuple<int, int, int, int> idx, std::vector<std::tuple<long, long, long>> n) {
            // container element (computed element or permutation image) at Matsubara index triplets
            auto& el = (*b.C)(IndexCombination4(ParticleIndex(std::get<0>(idx)), ParticleIndex(std::get<1>(idx)),
                                                ParticleIndex(std::get<2>(idx)), ParticleIndex(std::get<3>(idx))));
            std::vector<ComplexType> v;
            for(auto const& t : n)
                v.push_back(el(std::get<0>(t), std::get<1>(t), std::get<2>(t)));
            return to_np(v);
        })
        .def("permutation", [](G2Box& b, std::tuple<int, int, int, int> idx) {
            // the frequency permutation (perm, sign) that maps idx to its computed element
            auto& el = (*b.C)(IndexCombination4(ParticleIndex(std::get<0>(idx)), ParticleIndex(std::get<1>(idx)),
                                                ParticleIndex(std::get<2>(idx)), ParticleIndex(std::get<3>(idx))));
            auto const& P = el.FrequenciesPermutation;
            return py::make_tuple(std::vector<int>(P.perm.begin(), P.perm.end()), P.sign);
        });
    py::class_<ED>(m, "ED")
        .def(py::init([](py::list terms, double beta, py::object keep_hmat) {
                 // False / True (every block), or "largest"
                 int keep = py::isinstance<py::str>(keep_hmat) ? (keep_hmat.cast<std::string>() == "largest" ? 2 : -1)
                                                               : (keep_hmat.cast<bool>() ? 1 : 0);
                 if(keep < 0)
                     throw std::invalid_argument("keep_hmat: False, True or 'largest'");
                 return new ED(terms, beta, keep);
             }),
             py::arg("terms"), py::arg("beta"), py::arg("keep_hmat") = false)
        .def_readonly("t_eigh", &ED::t_eigh)
        .def_readonly("t_keep", &ED::t_keep)
        .def_property_readonly("E0", [](ED& e) { return e.H->getGroundEnergy(); })
        .def_property_readonly("Z", [](ED& e) { return e.rho->getPartitionFunction(); })
        .def("indices", [](ED& e) {
            py::list l;
            for(ParticleIndex i = 0; i < e.info->getIndexSize(); ++i) {
                auto const& t = e.info->getInfo(i);
                l.append(py::make_tuple(std::get<0>(t), std::get<1>(t), int(std::get<2>(t))));
            }
            return l;
        })
        .def("n_blocks", [](ED& e) { return e.S->getNumberOfBlocks(); })
        .def("blocks", [](ED& e) {
            std::vector<std::vector<QuantumState>> b;
            for(BlockNumber k = 0; k < e.S->getNumberOfBlocks(); ++k)
                b.push_back(e.S->getFockStates(k));
            return b;
        })
        .def("hmat", [](ED& e, int b) { return to_np(e.hmats.at(BlockNumber(b))); })
        .def("evals", [](ED& e, int b) {
            auto const& v = e.H->getEigenValues(b);
            return to_np(std::vector<double>(v.begin(), v.end()));
        })
        .def("evecs", [](ED& e, int b) { return to_np(e.H->getPart(b).getMatrix<false>()); })
        .def("weights", [](ED& e, int b) {
            auto const& w = e.rho->getPart(b).getWeights();
            return to_np(std::vector<double>(w.begin(), w.end()));
        })
        .def("truncate_blocks", [](ED& e, double tol) { e.rho->truncateBlocks(tol, false); })
        .def("use_indices", [](ED& e, std::vector<int> idx) {
            if(e.ops)
                throw std::runtime_error("use_indices: field operators already computed");
            e.field_subset.clear();
            for(int i : idx)
                e.field_subset.insert(ParticleIndex(i));
        })
        .def_readonly("t_rot", &ED::t_rot)
        .def("retained", [](ED& e) {
            std::vector<bool> r;
            for(BlockNumber b = 0; b < e.S->getNumberOfBlocks(); ++b)
                r.push_back(e.rho->isRetained(b));
            return r;
        })
        .def("rotate", [](ED& e, py::list op, double tol) {
            MonomialOperator M(e.mono_expr(op), *e.HS, *e.S, *e.H);
            M.prepare(*e.HS);
            M.compute(tol);
            py::list out;
            for(auto const& p : M.getParts()) {
                py::dict d;
                d["from"] = p.getRightIndex();
                d["to"] = p.getLeftIndex();
                d["m"] = to_np(p.getRowMajorValue<false>());
                out.append(d);
            }
            return out;
        }, py::arg("op"), py::arg("tol") = 1e-8)
        .def("field_operator", [](ED& e, int i, bool dag) {
            auto& ops = e.field_ops();
            MonomialOperator const& M = dag ? static_cast<MonomialOperator const&>(ops.getCreationOperator(ParticleIndex(i)))
                                            : static_cast<MonomialOperator const&>(ops.getAnnihilationOperator(ParticleIndex(i)));
            py::list out;
            for(auto const& p : M.getParts()) {
                py::dict d;
                d["from"] = p.getRightIndex();
                d["to"] = p.getLeftIndex();
                d["m"] = to_np(p.getRowMajorValue<false>());
                out.append(d);
            }
            return out;
        })
        .def("gf", [](ED& e, int i, int j, py::array_t<std::complex<double>> z, double pr, double ct,
                      std::vector<double> tau) {
            auto& ops = e.field_ops();
            GreensFunction G(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(i)), ops.getCreationOperator(ParticleIndex(j)), *e.rho);
            G.PoleResolution = pr;
            G.CoefficientTolerance = ct;
            G.prepare();
            G.compute();
            py::dict d;
            d["values"] = to_np(G.evaluate(to_points(z)));
            d["n_parts"] = G.getNumParts();
            d["n_terms"] = G.getNumTerms();
            std::vector<double> poles;
            std::vector<ComplexType> res;
            for(auto const& p : G.getParts())
                for(auto const& t : p.getTerms()) {
                    poles.push_back(t.Pole);
                    res.push_back(t.Residue);
                }
            d["poles"] = to_np(poles);
            d["residues"] = to_np(res);
            if(!tau.empty()) {
                d["tau_values"] = to_np(G.evaluateTau(tau));
                ComplexType s0 = 0; // GreensFunction::of_tau as the sum of the parts' of_tau
                for(auto const& p : G.getParts())
                    s0 += p.of_tau(tau[0]);
                d["tau0_parts"] = s0;
                d["tau0"] = G.of_tau(tau[0]);
            }
            return d;
        }, py::arg("i"), py::arg("j"), py::arg("z"), py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-8,
           py::arg("tau") = std::vector<double>{})
        .def("gf_pointwise", [](ED& e, int i, int j, std::vector<long> n) {
            auto& ops = e.field_ops();
            GreensFunction G(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(i)), ops.getCreationOperator(ParticleIndex(j)), *e.rho);
            G.prepare();
            G.compute();
            std::vector<ComplexType> v;
            for(long k : n)
                v.push_back(G(k));
            return to_np(v);
        })
        .def("chi", [](ED& e, py::list A, py::list B, py::array_t<std::complex<double>> z, double pr, double ct, bool sub,
                       std::vector<double> tau) {
            MonomialOperator& MA = e.mono(A);
            MonomialOperator& MB = e.mono(B);
            Susceptibility X(*e.S, *e.H, MA, MB, *e.rho);
            X.PoleResolution = pr;
            X.CoefficientTolerance = ct;
            X.prepare();
            X.compute();
            if(sub)
                X.subtractDisconnected();
            py::dict d;
            d["values"] = to_np(X.evaluate(to_points(z)));
            d["n_parts"] = X.getNumParts();
            d["n_terms"] = X.getNumTerms();
            if(!tau.empty()) {
                d["tau_values"] = to_np(X.evaluateTau(tau));
                ComplexType s0 = 0; // Susceptibility::of_tau as the sum of the parts' of_tau
                for(auto const& p : X.getParts())
                    s0 += p.of_tau(tau[0]);
                d["tau0_parts"] = s0;
                d["tau0"] = X.of_tau(tau[0]);
            }
            return d;
        }, py::arg("A"), py::arg("B"), py::arg("z"), py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-8,
           py::arg("subtract_disconnected") = false, py::arg("tau") = std::vector<double>{})
        .def("ensemble_average", [](ED& e, py::list A) {
            MonomialOperator MA(e.mono_expr(A), *e.HS, *e.S, *e.H);
            MA.prepare(*e.HS);
            MA.compute();
            EnsembleAverage EA(MA, *e.rho);
            EA.compute();
            return EA();
        })
        .def("g2", [](ED& e, int i, int j, int k, int l, py::array_t<std::complex<double>> zs, double pr, double ct,
                      py::object parts, bool keep) {
            auto& ops = e.field_ops();
            TwoParticleGF G2(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(i)), ops.getAnnihilationOperator(ParticleIndex(j)),
                             ops.getCreationOperator(ParticleIndex(k)), ops.getCreationOperator(ParticleIndex(l)), *e.rho);
            G2.PoleResolution = pr;
            G2.CoefficientTolerance = ct;
            G2.prepare();
            FreqVec3 f = to_freqs(zs);
            std::vector<std::size_t> which;
            if(parts.is_none())
                for(std::size_t q = 0; q < G2.getNumParts(); ++q)
                    which.push_back(q);
            else
                which = parts.cast<std::vector<std::size_t>>();
            auto t0 = std::chrono::steady_clock::now();
            std::vector<ComplexType> vals = G2.computeParts(which, !keep, f);
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            py::dict d;
            d["values"] = to_np(vals);
            d["seconds"] = secs;
            d["n_parts_total"] = G2.getNumParts();
            int64_t nnr = 0, nr = 0, raw = 0;
            std::vector<int64_t> pnr, pr_, pb;
            for(std::size_t q : which) {
                auto const& p = G2.getParts()[q];
                nnr += int64_t(p.getNumNonResonantTerms());
                nr += int64_t(p.getNumResonantTerms());
                raw += p.getNumRawTerms();
                pnr.push_back(int64_t(p.getNumNonResonantTerms()));
                pr_.push_back(int64_t(p.getNumResonantTerms()));
                auto b = p.getBlocks();
                int perm = 0;
                for(int pp = 0; pp < 6; ++pp)
                    if(permutations3[pp] == p.getPermutation())
                        perm = pp;
                for(int x : {b[0], b[1], b[2], b[3], perm})
                    pb.push_back(x);
            }
            d["n_parts"] = which.size();
            d["n_nonres"] = nnr;
            d["n_res"] = nr;
            d["n_raw"] = raw;
            d["part_nonres"] = pnr;
            d["part_res"] = pr_;
            d["part_blocks"] = pb;
            if(keep) {
                // lazy path: TwoParticleGF::operator() on the kept terms, point by point
                std::vector<ComplexType> lazy;
                for(auto const& t : f)
                    lazy.push_back(G2(std::get<0>(t), std::get<1>(t), std::get<2>(t)));
                d["values_lazy"] = to_np(lazy);
            }
            return d;
        }, py::arg("i"), py::arg("j"), py::arg("k"), py::arg("l"), py::arg("zs"), py::arg("pole_res") = 1e-8,
           py::arg("coeff_tol") = 1e-16, py::arg("parts") = py::none(), py::arg("keep") = false)
        .def("spectrum_tangent", [](ED& e, py::list direction) {
            // per block: (dE, dw) along the Hamiltonian direction
            SpectrumTangent T(*e.S, *e.H, *e.rho);
            T.prepare(build_expr(direction), *e.HS);
            T.compute();
            py::list out;
            for(BlockNumber b = 0; b < e.S->getNumberOfBlocks(); ++b)
                out.append(py::make_tuple(to_np(T.getEigenValueTangents(b)), to_np(T.getWeightTangents(b))));
            return out;
        })
        .def("g2_tangent", [](ED& e, int i, int j, int k, int l, py::array_t<std::complex<double>> zs, py::list direction,
                              double pr, double ct) {
            auto& ops = e.field_ops();
            TwoParticleGF G2(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(i)), ops.getAnnihilationOperator(ParticleIndex(j)),
                             ops.getCreationOperator(ParticleIndex(k)), ops.getCreationOperator(ParticleIndex(l)), *e.rho);
            G2.PoleResolution = pr;
            G2.CoefficientTolerance = ct;
            G2.prepare();
            FreqVec3 f = to_freqs(zs);
            auto t0 = std::chrono::steady_clock::now();
            SpectrumTangent T(*e.S, *e.H, *e.rho);
            T.prepare(build_expr(direction), *e.HS);
            T.compute();
            std::vector<ComplexType> vals;
            std::vector<ComplexType> der = G2.computeTangent(T, f, &vals);
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            py::dict d;
            d["values"] = to_np(vals);
            d["tangent"] = to_np(der);
            d["seconds"] = secs;
            int64_t nnr = 0, nr = 0, raw = 0;
            for(auto const& p : G2.getParts()) {
                nnr += int64_t(p.getNumNonResonantTerms());
                nr += int64_t(p.getNumResonantTerms());
                raw += p.getNumRawTerms();
            }
            d["n_nonres"] = nnr;
            d["n_res"] = nr;
            d["n_raw"] = raw;
            return d;
        }, py::arg("i"), py::arg("j"), py::arg("k"), py::arg("l"), py::arg("zs"), py::arg("direction"),
           py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-16)
        .def("vertex4", [](ED& e, int i, int j, int k, int l, long n_iw, double pr, double ct,
                           std::vector<std::tuple<long, long, long>> query) {
            // Gamma4 of <c_i c_j c+_k c+_l> with G13 = G_ik, G24 = G_jl, G14 = G_il, G23 = G_jk
            auto& ops = e.field_ops();
            TwoParticleGF Chi(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(i)), ops.getAnnihilationOperator(ParticleIndex(j)),
                              ops.getCreationOperator(ParticleIndex(k)), ops.getCreationOperator(ParticleIndex(l)), *e.rho);
            Chi.PoleResolution = pr;
            Chi.CoefficientTolerance = ct;
            Chi.prepare();
            Chi.compute(false);
            auto gf = [&](int a, int b) {
                auto G = std::make_unique<GreensFunction>(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(a)),
                                                          ops.getCreationOperator(ParticleIndex(b)), *e.rho);
                G->prepare();
                G->compute();
                return G;
            };
            auto G13 = gf(i, k), G24 = gf(j, l), G14 = gf(i, l), G23 = gf(j, k);
            Vertex4 V(Chi, *G13, *G24, *G14, *G23);
            auto t0 = std::chrono::steady_clock::now();
            V.compute(n_iw);
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::vector<ComplexType> stored, direct, chi;
            for(auto const& q : query) {
                long n1 = std::get<0>(q), n2 = std::get<1>(q), n3 = std::get<2>(q);
                stored.push_back(V(n1, n2, n3));
                direct.push_back(V.value(n1, n2, n3));
                chi.push_back(Chi(n1, n2, n3));
            }
            py::dict d;
            d["gamma"] = to_np(stored);
            d["gamma_value"] = to_np(direct);
            d["chi"] = to_np(chi);
            d["seconds"] = secs;
            return d;
        })
        .def("chi3", [](ED& e, int channel, int i, int j, int k, int l, py::array_t<std::complex<double>> zs, double pr,
                        double ct, bool keep) {
            if(channel < 0 || channel > 2)
                throw std::runtime_error("channel must be 0 (PP), 1 (PH) or 2 (xPH)");
            auto& ops = e.field_ops();
            ThreePointSusceptibility X(Channel(channel), *e.S, *e.H, ops.getCreationOperator(ParticleIndex(i)),
                                       ops.getAnnihilationOperator(ParticleIndex(j)),
                                       ops.getCreationOperator(ParticleIndex(k)),
                                       ops.getAnnihilationOperator(ParticleIndex(l)), *e.rho);
            X.PoleResolution = pr;
            X.CoefficientTolerance = ct;
            X.prepare();
            FreqVec2 f = to_pairs(zs);
            auto t0 = std::chrono::steady_clock::now();
            std::vector<ComplexType> vals = X.compute(!keep, f);
            double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            py::dict d;
            d["values"] = to_np(vals);
            d["seconds"] = secs;
            d["vanishing"] = X.isVanishing();
            d["n_parts"] = X.getNumParts();
            int64_t nff = 0, nfb = 0, nr = 0, raw = 0;
            std::vector<int64_t> pff, pfb, pr_, pb;
            for(auto const& p : X.getParts()) {
                nff += int64_t(p.getNumNonResonantFFTerms());
                nfb += int64_t(p.getNumNonResonantFBTerms());
                nr += int64_t(p.getNumResonantTerms());
                raw += p.getNumRawTerms();
                pff.push_back(int64_t(p.getNumNonResonantFFTerms()));
                pfb.push_back(int64_t(p.getNumNonResonantFBTerms()));
                pr_.push_back(int64_t(p.getNumResonantTerms()));
                auto b = p.getBlocks();
                for(int x : {int(b[0]), int(b[1]), int(b[2]), int(p.getSwappedFermionOps())})
                    pb.push_back(x);
            }
            d["n_ff"] = nff;
            d["n_fb"] = nfb;
            d["n_res"] = nr;
            d["n_raw"] = raw;
            d["part_ff"] = pff;
            d["part_fb"] = pfb;
            d["part_res"] = pr_;
            d["part_blocks"] = pb;
            if(keep) {
                std::vector<ComplexType> lazy;
                for(auto const& t : f)
                    lazy.push_back(X(std::get<0>(t), std::get<1>(t)));
                d["values_lazy"] = to_np(lazy);
                d["values_bulk"] = to_np(X.evaluate(f));
                // merged terms of the first part, through the term getters
                if(X.getNumParts() > 0) {
                    auto const& p = X.getParts()[0];
                    ComplexType s = 0;
                    ComplexType z1 = std::get<0>(f.at(0)), z2 = std::get<1>(f.at(0));
                    for(auto const& t : p.getNonResonantFFTerms())
                        s += t(z1, z2);
                    for(auto const& t : p.getNonResonantFBTerms())
                        s += t(z1, z2);
                    for(auto const& t : p.getResonantTerms())
                        s += t(z1, z2, pr);
                    d["part0_from_terms"] = s;
                    d["part0"] = p(z1, z2);
                }
            }
            return d;
        }, py::arg("channel"), py::arg("i"), py::arg("j"), py::arg("k"), py::arg("l"), py::arg("zs"),
           py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-16, py::arg("keep") = false)
        .def("chi3_container", [](ED& e, int channel, std::vector<std::tuple<int, int, int, int>> idx,
                                  py::array_t<std::complex<double>> zs) {
            auto& ops = e.field_ops();
            ThreePointSusceptibilityContainer C(Channel(channel), *e.info, *e.S, *e.H, *e.rho, ops);
            std::set<IndexCombination4> want;
            for(auto const& t : idx)
                want.emplace(ParticleIndex(std::get<0>(t)), ParticleIndex(std::get<1>(t)), ParticleIndex(std::get<2>(t)),
                             ParticleIndex(std::get<3>(t)));
            C.prepareAll(want);
            auto all = C.computeAll(false, to_pairs(zs));
            py::dict d;
            for(auto const& kv : all) {
                auto const& I = kv.first;
                d[py::make_tuple(int(I.Index1), int(I.Index2), int(I.Index3), int(I.Index4))] = to_np(kv.second);
            }
            return d;
        })
        .def("g2_parts", [](ED& e, int i, int j, int k, int l) {
            // part list of TwoParticleGF::prepare: (S1, S2, S3, S4, permutation, n1, n2, n3, n4)
            auto& ops = e.field_ops();
            TwoParticleGF G2(*e.S, *e.H, ops.getAnnihilationOperator(ParticleIndex(i)), ops.getAnnihilationOperator(ParticleIndex(j)),
                             ops.getCreationOperator(ParticleIndex(k)), ops.getCreationOperator(ParticleIndex(l)), *e.rho);
            G2.prepare();
            std::vector<std::vector<int64_t>> out;
            for(auto const& p : G2.getParts()) {
                auto b = p.getBlocks();
                int perm = 0;
                for(int pp = 0; pp < 6; ++pp)
                    if(permutations3[pp] == p.getPermutation())
                        perm = pp;
                out.push_back({b[0], b[1], b[2], b[3], perm, int64_t(e.S->getBlockSize(b[0])), int64_t(e.S->getBlockSize(b[1])),
                               int64_t(e.S->getBlockSize(b[2])), int64_t(e.S->getBlockSize(b[3]))});
            }
            return out;
        })
        .def("g2_container", [](ED& e, std::vector<std::tuple<int, int, int, int>> idx, std::vector<std::tuple<long, long, long>> n,
                                double pr, double ct, std::vector<std::tuple<int, int, int, int>> query) {
            auto& ops = e.field_ops();
            TwoParticleGFContainer C(*e.info, *e.S, *e.H, *e.rho, ops);
            C.PoleResolution = pr;
            C.CoefficientTolerance = ct;
            std::set<IndexCombination4> s;
            for(auto const& t : idx)
                s.emplace(ParticleIndex(std::get<0>(t)), ParticleIndex(std::get<1>(t)), ParticleIndex(std::get<2>(t)),
                          ParticleIndex(std::get<3>(t)));
            C.prepareAll(s);
            C.computeAll(false);
            py::dict out;
            for(auto const& q : query) {
                auto& el = C(IndexCombination4(ParticleIndex(std::get<0>(q)), ParticleIndex(std::get<1>(q)),
                                               ParticleIndex(std::get<2>(q)), ParticleIndex(std::get<3>(q))));
                std::vector<ComplexType> v;
                for(auto const& t : n)
                    v.push_back(el(std::get<0>(t), std::get<1>(t), std::get<2>(t)));
                out[py::cast(q)] = to_np(v);
            }
            return out;
        });
}
