"""The drop-in boundary against the reference's own callers: its Catch2 tests of the in-scope
path (test/*.cpp of the reference tree) compile unchanged against include/ (CPU; skipped where the
reference tree is absent), and, built by oracle/reftests.mk against this library into oracle/_ref,
pass on the device: their goldens and closed forms (G, anomalous G, 2PGF, vertex, susceptibility,
3-point susceptibility, Hamiltonian eigensystem and operator transposition) hold for the GPU path."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref")
MPI_INC = "/opt/conda/include"
TESTS = ["GF1siteTest", "GF2siteTest", "GF4siteTest", "GF1siteAnomalousTest", "AndersonTest", "Anderson2PGFTest",
         "Vertex4Test", "SusceptibilityTest", "HamiltonianTest", "3PSusc1siteTest", "3PSusc3siteTest"]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "test")), reason="reference tree absent")
@pytest.mark.parametrize("name", TESTS)
def test_reference_test_compiles_unchanged(name):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(REF, 'test')}",
           "-idirafter", MPI_INC, os.path.join(REF, "test", name + ".cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.gpu
@pytest.mark.parametrize("name", TESTS)
def test_reference_test_passes_on_device(name):
    exe = os.path.join(REF_BIN, name)
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref not built (needs the reference tree at build time)")
    print(f"[reference test] {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    print(tail.strip().splitlines()[-1] if tail.strip() else "", flush=True)
    assert r.returncode == 0, tail
    assert "All tests passed" in r.stdout, tail
