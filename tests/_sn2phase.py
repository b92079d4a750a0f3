"""The two-phase tin (Vinet beta/gamma) PTE fixture of the reference
(test/pte_test_2phaseVinetSn.hpp, test/material_2phaseVinetSn.hpp), loaded
from tests/golden/pte_2phase_vinet_sn.json, and its trial cells built as
test/test_pte_2phaseVinetSn.cpp builds them."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "pte_2phase_vinet_sn.json")))
PHASES = ("beta", "gamma")


def vinet_args(name):
    m = FIX["materials"][name]
    return (m["rho0"], m["T0"], m["B0"], m["BP0"], m["A0"], m["Cv0"], m["E0"], m["S0"])


def cells(which):
    """CSR arrays of the `which` ("long" / "short") trials: set_trial_state
    (pte_test_2phaseVinetSn.hpp:84-101), then sie_tot and rho_tot as the
    test's kernel sums them (test_pte_2phaseVinetSn.cpp:173-181)."""
    s = FIX[which]
    n = len(s["in_rho_tot"])
    lam, tv = FIX["in_lambda"], FIX["trial_vfrac"]
    rho = np.zeros(2 * n)
    vf = np.zeros(2 * n)
    sie = np.zeros(2 * n)
    sie_tot = np.zeros(n)
    u_tot = np.zeros(n)
    for t in range(n):
        vsum = 0.0
        for i in range(2):
            vf[2 * t + i] = tv[i]
            rho[2 * t + i] = lam[i] * s["in_rho_tot"][t] / vf[2 * t + i]
            sie[2 * t + i] = s["in_sie_tot"][t]
            vsum += vf[2 * t + i]
        for i in range(2):
            vf[2 * t + i] *= 1.0 / vsum
        st = rt = 0.0
        for i in range(2):
            rt += rho[2 * t + i] * vf[2 * t + i]
            st += rho[2 * t + i] * vf[2 * t + i] * sie[2 * t + i]
        st /= rt
        sie_tot[t] = st
        u_tot[t] = rt * st
    return dict(ncell=n, nnz=2 * n, offsets=np.arange(0, 2 * n + 1, 2, dtype=np.int32),
                mat_ids=np.tile(np.array([0, 1], dtype=np.uint8), n), rho=rho, vfrac=vf,
                sie=sie, temp=np.zeros(2 * n), press=np.zeros(2 * n), vfrac_tot=np.ones(n),
                sie_tot=sie_tot, u_tot=u_tot)


def expected(which):
    s = FIX[which]
    return {k: np.array(s[k], dtype=np.float64) for k in
            ("out_press", "out_temp", "out_rho0", "out_rho1", "out_vfrac0", "out_vfrac1")}
