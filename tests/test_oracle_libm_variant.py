"""What replacing the reference's std::exp / std::cbrt (Vinet,
singularity-eos/eos/eos_vinet.hpp:235-239) by the portable routines that
host and device share changes in the closure: the oracle built both ways
(oracle/_build/liboracle.so and liboracle_libm.so) solves the first 16,384
cells of the uneven workload with PTESolverRhoT at the default MixParams and
at the 1e-10 set.

Measured (and asserted): cells without a Vinet material are bit-identical
(nothing else calls exp or cbrt). Nearly every Vinet cell differs in its
last bits (the two exp / cbrt agree within 1 ulp, and any last-bit change
moves the Newton path by rounding). At the default tolerance the
convergence flags and iteration counts are identical and the states agree
to 1e-10 relative in rho, alpha, e and T (4e-11 measured); pressures move
by the bulk modulus times the density change, up to 2e-8 of the cell's
largest pressure (condensed materials at low pressure, B / P ~ 1e3). At the
1e-10 set the convergence flag changes on under 1 % of the Vinet cells
(cells that stall near the roundoff floor converge or hit the cap the other
way) and iteration counts with it; where both converge, the states agree to
1e-10 (P to 1e-9 of the cell's largest pressure).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
N = 16384


def run(tmp_path, libm, tight):
    out = os.path.join(str(tmp_path), f"r_{int(libm)}_{int(tight)}.npz")
    env = dict(os.environ, SGO_ORACLE_LIBM="1" if libm else "0")
    subprocess.check_call([sys.executable, os.path.join(HERE, "_uneven_oracle_run.py"), out,
                           str(N), "1" if tight else "0"], env=env)
    return dict(np.load(out))


@pytest.mark.parametrize("tight", [False, True], ids=["default", "tight"])
def test_libm_exp_cbrt_changes_only_vinet_cells_within_tolerance(oracle, sgm, tmp_path, tight):
    a = run(tmp_path, False, tight)
    b = run(tmp_path, True, tight)
    off = a["offsets"]
    ncell = len(off) - 1
    has_vinet = np.zeros(ncell, dtype=bool)
    cell_of = np.repeat(np.arange(ncell), np.diff(off))
    has_vinet[np.unique(cell_of[a["mat_ids"] == 4])] = True
    same_entry = np.ones(len(cell_of), dtype=bool)
    for k in ("rho", "vfrac", "sie", "temp", "press"):
        same_entry &= (a[k] == b[k]) | (np.isnan(a[k]) & np.isnan(b[k]))
    same_cell = np.ones(ncell, dtype=bool)
    np.logical_and.at(same_cell, cell_of, same_entry)
    same_cell &= a["niter"] == b["niter"]
    assert same_cell[~has_vinet].all()
    frac = 1.0 - same_cell[has_vinet].mean()
    print(f"\n[libm exp/cbrt] tight={tight}: {has_vinet.sum()} Vinet cells, "
          f"{(~same_cell[has_vinet]).sum()} differ ({100 * frac:.2f} %)")
    conv_agree = (a["conv"] == b["conv"])[has_vinet].mean()
    print(f"[libm exp/cbrt] convergence flags agree on {100 * conv_agree:.2f} % of Vinet cells")
    both = (a["conv"] == 1) & (b["conv"] == 1)
    ok = both[cell_of]
    if tight:
        assert conv_agree > 0.98
    else:
        assert np.array_equal(a["conv"], b["conv"]) and np.array_equal(a["niter"], b["niter"])
    for k in ("rho", "vfrac", "sie", "temp"):
        rel = np.abs(a[k] - b[k]) / np.maximum(np.abs(a[k]), 1e-300)
        assert rel[ok].max() <= 1e-10, (k, rel[ok].max())
    cmax = np.zeros(ncell)
    np.maximum.at(cmax, cell_of, np.abs(a["press"]))
    relp = np.abs(a["press"] - b["press"]) / np.maximum(cmax[cell_of], 1.0)
    assert relp[ok].max() <= (1e-9 if tight else 1e-7), relp[ok].max()
    assert frac > 0.5  # the substitution is visible in the last bits
