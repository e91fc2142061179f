// oracle/ref_cpu_py.cpp -- TEST INFRASTRUCTURE ONLY: Python binding of the CPU oracle.
#include "ref_cpu.hpp"

#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

namespace py = pybind11;
using namespace oracle;

static std::vector<HTerm> to_terms(py::list terms) {
    std::vector<HTerm> out;
    for(auto t : terms) {
        auto tt = t.cast<py::tuple>();
        HTerm h;
        h.coeff = tt[0].cast<double>();
        for(auto o : tt[1].cast<py::list>()) {
            auto ot = o.cast<py::tuple>();
            h.ops.push_back(FOp{ot[0].cast<bool>(), ot[1].cast<int>()});
        }
        out.push_back(h);
    }
    return out;
}
static Mono to_mono(py::list ops) {
    Mono m;
    for(auto o : ops) {
        auto ot = o.cast<py::tuple>();
        m.push_back(FOp{ot[0].cast<bool>(), ot[1].cast<int>()});
    }
    return m;
}
template <class T> static py::array_t<T> arr(std::vector<T> const& v) {
    return py::array_t<T>(v.size(), v.data());
}
static py::array_t<double> mat(std::vector<double> const& v, int r, int c) {
    py::array_t<double> a({r, c});
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}
static std::vector<cplx> zvec(py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> z) {
    return std::vector<cplx>(z.data(), z.data() + z.size());
}

PYBIND11_MODULE(pomerol_oracle, m) {
    m.doc() = "CPU restatement of pomerol's full-ED path (test oracle, not product)";
    py::class_<ED>(m, "ED")
        .def(py::init([](int n_modes, py::list terms, double beta, std::string const& eigensolver) {
                 if(eigensolver != "jacobi" && eigensolver != "qr")
                     throw std::runtime_error("eigensolver must be 'jacobi' or 'qr'");
                 return new ED(n_modes, to_terms(terms), beta, eigensolver == "qr" ? 1 : 0);
             }),
             py::arg("n_modes"), py::arg("terms"), py::arg("beta"), py::arg("eigensolver") = "jacobi")
        .def_static("eigh", [](py::array_t<double, py::array::c_style | py::array::forcecast> A, std::string const& how) {
            // one dense symmetric eigensolve (tests of the oracle's eigensolvers; CPU timings)
            if(A.ndim() != 2 || A.shape(0) != A.shape(1))
                throw std::runtime_error("eigh: square matrix expected");
            int n = int(A.shape(0));
            std::vector<double> a(A.data(), A.data() + A.size()), V, w;
            {
                py::gil_scoped_release rel;
                if(how == "qr")
                    tridiag_qr_eigh(n, a, V, w);
                else
                    jacobi_eigh(n, a, V, w);
            }
            return py::make_tuple(arr(w), mat(V, n, n));
        }, py::arg("A"), py::arg("how") = "qr")
        .def_readonly("E0", &ED::E0)
        .def_readonly("Z", &ED::Z)
        .def("n_blocks", [](ED const& e) { return e.S.blocks.size(); })
        .def("blocks", [](ED const& e) { return e.S.blocks; })
        .def("state_block", [](ED const& e) { return e.S.state_block; })
        .def("hmat", [](ED const& e, int b) { int n = e.S.blocks[b].size(); return mat(e.hmat[b], n, n); })
        .def("evals", [](ED const& e, int b) { return arr(e.evals[b]); })
        .def("evecs", [](ED const& e, int b) { int n = e.S.blocks[b].size(); return mat(e.evecs[b], n, n); })
        .def("weights", [](ED const& e, int b) { return arr(e.weights[b]); })
        .def("retained", [](ED const& e) { return std::vector<bool>(e.retained.begin(), e.retained.end()); })
        .def("truncate_blocks", &ED::truncate_blocks)
        .def("set_spectrum",
             [](ED& e, int b, py::array_t<double, py::array::c_style | py::array::forcecast> ev,
                py::array_t<double, py::array::c_style | py::array::forcecast> vec,
                py::array_t<double, py::array::c_style | py::array::forcecast> w) {
                 e.set_spectrum(b, std::vector<double>(ev.data(), ev.data() + ev.size()),
                                std::vector<double>(vec.data(), vec.data() + vec.size()),
                                std::vector<double>(w.data(), w.data() + w.size()));
             })
        .def("rotate",
             [](ED const& e, py::list op, double tol) {
                 RotOp R = e.rotate(to_mono(op), tol);
                 py::list out;
                 for(auto const& p : R.parts) {
                     py::dict d;
                     d["from"] = p.from;
                     d["to"] = p.to;
                     d["m"] = mat(p.m, p.n_to, p.n_from);
                     out.append(d);
                 }
                 return out;
             },
             py::arg("op"), py::arg("tol") = 1e-8)
        .def("gf",
             [](ED const& e, int i, int j, py::array_t<std::complex<double>> z, double pr, double ct,
                std::vector<double> tau) {
                 GFResult r = e.gf(i, j, zvec(z), pr, ct, tau);
                 py::dict d;
                 d["values"] = arr(r.values);
                 d["tau_values"] = arr(r.tau_values);
                 d["n_parts"] = r.n_parts;
                 d["n_terms"] = r.n_terms;
                 d["poles"] = arr(r.poles);
                 d["residues"] = arr(r.residues);
                 return d;
             },
             py::arg("i"), py::arg("j"), py::arg("z"), py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-8,
             py::arg("tau") = std::vector<double>{})
        .def("chi",
             [](ED const& e, py::list A, py::list B, py::array_t<std::complex<double>> z, double pr, double ct,
                bool sub, std::vector<double> tau) {
                 ChiResult r = e.chi(to_mono(A), to_mono(B), zvec(z), pr, ct, sub, tau);
                 py::dict d;
                 d["values"] = arr(r.values);
                 d["tau_values"] = arr(r.tau_values);
                 d["n_parts"] = r.n_parts;
                 d["n_terms"] = r.n_terms;
                 d["zero_pole_weight"] = r.zero_pole_weight;
                 return d;
             },
             py::arg("A"), py::arg("B"), py::arg("z"), py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-8,
             py::arg("subtract_disconnected") = false, py::arg("tau") = std::vector<double>{})
        .def("ensemble_average", [](ED const& e, py::list A) { return e.ensemble_average(to_mono(A)); })
        .def("chi3",
             [](ED const& e, int channel, int i, int j, int k, int l, py::array_t<std::complex<double>> zs, double pr,
                double ct) {
                 Chi3Result r;
                 std::vector<cplx> zv = zvec(zs);
                 {
                     py::gil_scoped_release rel;
                     r = e.chi3(channel, i, j, k, l, zv, pr, ct);
                 }
                 py::dict d;
                 d["values"] = arr(r.values);
                 d["n_parts"] = r.n_parts;
                 d["n_ff"] = r.n_ff;
                 d["n_fb"] = r.n_fb;
                 d["n_res"] = r.n_res;
                 d["n_triples"] = r.n_triples;
                 d["part_ff"] = arr(r.part_ff);
                 d["part_fb"] = arr(r.part_fb);
                 d["part_res"] = arr(r.part_res);
                 d["part_blocks"] = arr(r.part_blocks);
                 return d;
             },
             py::arg("channel"), py::arg("i"), py::arg("j"), py::arg("k"), py::arg("l"), py::arg("zs"),
             py::arg("pole_res") = 1e-8, py::arg("coeff_tol") = 1e-16)
        .def("g2",
             [](ED const& e, int i, int j, int k, int l, py::array_t<std::complex<double>> zs, double pr, double ct,
                int max_parts, py::object subset) {
                 std::vector<int> sub;
                 std::vector<int> const* psub = nullptr;
                 if(!subset.is_none()) {
                     sub = subset.cast<std::vector<int>>();
                     std::sort(sub.begin(), sub.end());
                     psub = &sub;
                 }
                 G2Result r;
                 std::vector<cplx> zv = zvec(zs);
                 {
                     py::gil_scoped_release rel;
                     r = e.g2(i, j, k, l, zv, pr, ct, max_parts, psub);
                 }
                 py::dict d;
                 d["values"] = arr(r.values);
                 d["n_parts"] = r.n_parts;
                 d["n_nonres"] = r.n_nonres;
                 d["n_res"] = r.n_res;
                 d["n_quadruples"] = r.n_quadruples;
                 d["n_erased"] = r.n_erased;
                 d["gen_seconds"] = r.gen_seconds;
                 d["eval_seconds"] = r.eval_seconds;
                 d["part_nonres"] = arr(r.part_nonres);
                 d["part_res"] = arr(r.part_res);
                 d["part_blocks"] = arr(r.part_blocks);
                 return d;
             },
             py::arg("i"), py::arg("j"), py::arg("k"), py::arg("l"), py::arg("zs"), py::arg("pole_res") = 1e-8,
             py::arg("coeff_tol") = 1e-16, py::arg("max_parts") = -1, py::arg("parts") = py::none());
}
