"""The command-line ED programs (pomerol_amd/bin/anderson.pomerol, hubbard2d.pomerol; reference
prog/): options and argument errors on the CPU, output files against the Python binding of the
same API on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "pomerol_amd", "bin")


def run(prog, args, cwd, check=True):
    print(f"[prog] {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BIN, prog)] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    if check:
        assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("prog,extra", [("anderson.pomerol", ["--U", "--ed", "--levels", "--hoppings"]),
                                        ("hubbard2d.pomerol", ["--U", "--mu", "--t", "--tp", "--x", "--y"])])
def test_help_lists_the_reference_options(prog, extra, tmp_path):
    """prog/quantum_model.cpp:28-42 and the models' own options."""
    out = run(prog, ["--help"], tmp_path).stdout
    for opt in ["--beta", "--calc_gf", "--calc_2gf", "--gf.eta", "--gf.step", "--gf.D", "--wf_min", "--wf_max",
                "--wb_min", "--wb_max", "--2pgf.indices", "--2pgf.reduce_tol", "--2pgf.coeff_tol"] + extra:
        assert opt in out, opt


def test_argument_errors(tmp_path):
    """Unknown flags exit with 1; levels and hoppings of different lengths with 2 (anderson_model.hpp:47-51)."""
    assert run("anderson.pomerol", ["--nonsense"], tmp_path, check=False).returncode == 1
    r = run("anderson.pomerol", ["--levels", "0.1,0.2", "--hoppings", "0.3"], tmp_path, check=False)
    assert r.returncode == 2 and "Number of levels != number of hoppings" in r.stderr


def load(path):
    return np.loadtxt(path, ndmin=2)


@pytest.mark.gpu
def test_anderson_program_outputs(pa, tmp_path):
    from pomerol_amd import workloads as W
    from pomerol_amd.presets import CoulombS, Hopping, Level

    U, ed, beta = 2.0, -1.0, 10.0
    levels, hops = [-0.5, 0.4], [0.3, 0.25]
    args = ["--beta", str(beta), "--U", str(U), "--ed", str(ed), "--levels", "-0.5,0.4", "--hoppings", "0.3,0.25",
            "--calc_2gf", "--wf_min", "-3", "--wf_max", "3", "--wb_min", "0", "--wb_max", "1",
            "--2pgf.indices", "0,1,0,1", "--gf.D", "2", "--gf.step", "0.1"]
    out = run("anderson.pomerol", args, tmp_path).stdout
    assert "Index 0 = (A, 0, dn)" in out and "2PGF : 98 freqs to evaluate" in out

    H = CoulombS("A", U, ed)
    for k, (e, v) in enumerate(zip(levels, hops)):
        H = H + Hopping("A", f"b{k}", v) + Level(f"b{k}", e)
    assert H.indices()[0] == ("A", 0, 0) and H.indices()[1] == ("A", 0, 1)
    ed_ = pa.ED(W.terms(H), beta)
    ev = np.sort(np.concatenate([ed_.evals(b) for b in range(ed_.n_blocks())]))
    spec = load(tmp_path / "spectrum.dat")
    assert np.array_equal(spec[:, 0], np.arange(len(ev))) and np.abs(spec[:, 1] - ev).max() < 1e-13

    gi = load(tmp_path / "gw_imfreq_00.dat")
    n = np.arange(-3, 13)
    assert np.allclose(gi[:, 0], np.pi * (2 * n + 1) / beta, rtol=0, atol=1e-14)
    g = ed_.gf(0, 0, 1j * gi[:, 0], 1e-8, 1e-8)["values"]
    assert np.abs(gi[:, 1] + 1j * gi[:, 2] - g).max() < 1e-14
    gr = load(tmp_path / "gw_refreq_00.dat")
    assert len(gr) == 41 and abs(gr[0, 0] + 2) < 1e-14 and abs(gr[-1, 0] - 2) < 1e-14
    g = ed_.gf(0, 0, gr[:, 0] + 0.05j, 1e-8, 1e-8)["values"]
    assert np.abs(gr[:, 1] + 1j * gr[:, 2] - g).max() < 1e-14

    chi = load(tmp_path / "chi0101.dat")
    assert chi.shape == (2 * 7 * 7, 5)
    zs = np.array([[1j * (W_ + w3), 1j * w2, 1j * w3] for W_, w3, w2 in chi[:, :3]])
    ref = ed_.g2(0, 1, 0, 1, zs, 1e-5, 1e-12)["values"]
    assert np.abs(chi[:, 3] + 1j * chi[:, 4] - ref).max() < 1e-12 * np.abs(ref).max()
    one = load(tmp_path / f"chi0101_W{2 * np.pi / beta:.6f}.dat")
    assert np.array_equal(one, chi[49:, 1:])


@pytest.mark.gpu
def test_hubbard2d_program_outputs(pa, tmp_path):
    from pomerol_amd import workloads as W
    from pomerol_amd.presets import CoulombS, Hopping

    U, beta = 3.0, 5.0
    out = run("hubbard2d.pomerol", ["--beta", str(beta), "--U", str(U), "--x", "2", "--y", "1", "--calc_gf",
                                    "--wf_min", "0", "--wf_max", "2", "--gf.D", "1", "--gf.step", "0.5"], tmp_path).stdout
    assert "Hamiltonian:" in out
    H = CoulombS("S0", U, -U / 2) + CoulombS("S1", U, -U / 2) + Hopping("S0", "S1", -1.0) + Hopping("S0", "S1", -1.0)
    ed_ = pa.ED(W.terms(H), beta)
    idx = H.indices()
    d0, d1 = idx.index(("S0", 0, 0)), idx.index(("S1", 0, 0))
    for j in (d0, d1):
        gi = load(tmp_path / f"gw_imfreq_{d0}{j}.dat")
        g = ed_.gf(d0, j, 1j * gi[:, 0], 1e-8, 1e-8)["values"]
        assert len(gi) == 9 and np.abs(gi[:, 1] + 1j * gi[:, 2] - g).max() < 1e-14
