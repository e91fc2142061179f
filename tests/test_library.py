"""The native libraries load and export the C ABI (CPU only, no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pomx_c.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s+(pomx_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("pomx_ctx_create", "pomx_eigh_batched", "pomx_rotate_batched", "pomx_pair_terms", "pomx_pair_eval",
              "pomx_g2_part_compute", "pomx_g2_terms_eval", "pomx_boltzmann", "pomx_assemble"):
        assert n in names


def test_c_abi_library_exports_every_declared_symbol():
    import pomerol_amd

    lib = ctypes.CDLL(pomerol_amd.POMX_LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_cpp_api_library_loads():
    import pomerol_amd

    ctypes.CDLL(pomerol_amd.API_LIB)
    assert hasattr(pomerol_amd, "ED")


def test_no_cpu_fallback_without_device():
    """Without a gfx950 device the product path raises instead of computing on the CPU."""
    import pomerol_amd

    if pomerol_amd.device_ok():
        pytest.skip("a device is present")
    with pytest.raises(RuntimeError, match="gfx950"):
        pomerol_amd.ED([(1.0, [(True, ("A", 0, 1)), (False, ("A", 0, 1))])], 1.0)


def test_ctx_create_reports_missing_device():
    import pomerol_amd

    if pomerol_amd.device_ok():
        pytest.skip("a device is present")
    lib = ctypes.CDLL(pomerol_amd.POMX_LIB)
    ctx = ctypes.c_void_p()
    assert lib.pomx_ctx_create(0, ctypes.byref(ctx)) == 4  # POMX_ERR_NODEV


def test_product_does_not_reference_oracle():
    """pomerol_amd/ must not import, include or link anything under oracle/."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "pomerol_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "pomerol_oracle" not in text and "ref_cpu" not in text, f


def _build_cpp_caller(out):
    import subprocess

    lib = os.path.join(ROOT, "pomerol_amd", "lib")
    cmd = ["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}", os.path.join(ROOT, "tests", "cpp", "api_usage.cpp"),
           f"-L{lib}", "-lpomerol_amd", "-lpomx", f"-Wl,-rpath,{lib}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]


def test_cpp_caller_with_reference_header_names_builds(tmp_path):
    """A C++ caller that includes the reference's header names (DensityMatrix.hpp, GFContainer.hpp,
    Susceptibility.hpp, TwoParticleGFContainer.hpp, ...) compiles and links against the library."""
    out = tmp_path / "api_usage"
    _build_cpp_caller(out)
    assert out.exists()


@pytest.mark.gpu
def test_cpp_caller_runs_on_device(tmp_path):
    """The same C++ caller runs end to end on the device (no Python in between)."""
    import subprocess

    out = tmp_path / "api_usage"
    _build_cpp_caller(out)
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=300)
    print("[cpp] api_usage:", r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "nan" not in r.stdout and "inf" not in r.stdout
