"""Host-side checks that need no GPU: the C-ABI library loads and exports every
entry point include/sgm.h declares, its host logic (parameter defaults, the
seeded workload generators, the Gruneisen table builder) agrees with the
oracle, and the device entry points refuse to run without a GPU instead of
falling back to the CPU."""
import ctypes as C
import hashlib
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sgm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sgm_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(sgm):
    L = sgm.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python bindings cover them all
    assert set(names) <= set(sgm._SIGS), set(names) - set(sgm._SIGS)


def test_version_string(sgm):
    v = sgm.lib().sgm_version()
    assert v and b"gfx950" in v


def test_mix_params_defaults_match_oracle(sgm, oracle):
    """MixParams() defaults of singularity-eos/closure/mixed_cell_models.hpp:88-130
    and the 1e-10 set of example/pte_2mat.cpp:169-180, through both layers."""
    d = sgm.mix_params().as_oracle_array()
    t = sgm.mix_params(tight=True).as_oracle_array()
    np.testing.assert_array_equal(d, oracle.default_params())
    np.testing.assert_array_equal(t, oracle.tight_params())
    assert t[1] == t[2] == t[3] == 1e-10 and t[5] == t[6] == 0.0


def test_no_gpu_is_an_error_not_a_fallback(sgm):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert sgm.lib().sgm_init(4, C.byref(h)) == -5
    with pytest.raises(RuntimeError):
        sgm.Context(2)


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest()


def test_two_material_workload_fixture(workloads):
    """65,536-cell two-material inputs (mt19937_64, seed 0x5EED0000): SHA-256 of
    each array and the values of the first cells, committed."""
    fx = json.load(open(os.path.join(GOLDEN, "workload_pte2.json")))
    a = workloads.pte2(fx["ncell"])
    for k, h in fx["sha256"].items():
        assert sha(a[k]) == h, k
    for k, v in fx["head"].items():
        np.testing.assert_array_equal(a[k][: len(v)], np.array(v))


def test_two_material_workload_distribution(workloads):
    """Draws follow their stated laws: alpha_A ~ U(0.05,0.95) and alpha_B = 1-alpha_A,
    rho_A ~ LogNormal(ln 1e-3, 0.3), rho_B ~ U(0.9,1.1), T ~ U(300,3000)."""
    n = 1 << 16
    a = workloads.pte2(n)
    aA, aB = a["vfrac"][0::2], a["vfrac"][1::2]
    assert aA.min() >= 0.05 and aA.max() <= 0.95 and abs(aA.mean() - 0.5) < 0.01
    np.testing.assert_array_equal(aA + aB, np.ones(n))
    lr = np.log(a["rho"][0::2])
    assert abs(lr.mean() - np.log(1e-3)) < 0.01 and abs(lr.std() - 0.3) < 0.01
    rB = a["rho"][1::2]
    assert rB.min() >= 0.9 and rB.max() <= 1.1
    T = a["temp"]
    assert T.min() >= 300 and T.max() <= 3000 and abs(T.mean() - 1650) < 15


def test_two_material_workload_energies_match_oracle(workloads, oracle):
    """sie/press/sie_tot of the generator equal the oracle's EOS at (rho, T)."""
    n = 4096
    a = workloads.pte2(n)
    A, B = oracle.ideal(*workloads.IDEAL_A), oracle.stiff(*workloads.STIFF_B)
    for m, e in ((0, A), (1, B)):
        r, T = a["rho"][m::2], a["temp"][m::2]
        np.testing.assert_array_equal(a["sie"][m::2], e.vec(2, r, T))
        np.testing.assert_array_equal(a["press"][m::2], e.vec(3, r, T))
    mass = a["rho"] * a["vfrac"]
    st = (mass[0::2] * a["sie"][0::2] + mass[1::2] * a["sie"][1::2]) / (mass[0::2] + mass[1::2])
    np.testing.assert_allclose(a["sie_tot"], st, rtol=1e-15)


def test_gruneisen_table_builder_matches_oracle(workloads, oracle):
    """The library's host-side table sampler and the oracle's agree bit for bit
    on every node and on the derived cold curves."""
    t = workloads.GruneisenTable(nrho=64, nt=96)
    _, f, c = oracle.gruneisen_table(workloads.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                     t.lTMax, t.nT)
    for k in f:
        np.testing.assert_array_equal(t.f[k], f[k], err_msg=k)
    for k in c:
        np.testing.assert_array_equal(t.c[k], c[k], err_msg=k)


def test_cpp_adapter_on_host_only_registry():
    """include/sgm.hpp compiles into a host program (built by the package
    Makefile) that sets materials and gets SGM_E_NOGPU from a launch on a
    registry without a device."""
    import subprocess
    exe = os.path.join(ROOT, "singularity-eos_amd", "sgm_adapter_demo")
    out = subprocess.run([exe, "--host"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "adapter host OK" in out.stdout
