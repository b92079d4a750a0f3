"""GPU parity of the get_sg_eos closure driver (sgm_get_sg_eos) against the
oracle's restatement, through the C ABI, in all four input modes.

Cells come from the uneven workload (IdealGas x2, StiffGas, Gruneisen, Vinet
and the rho-T table), every 7th reduced to one material (the single-material
paths) and every 5th given a trace material below the mass-fraction cutoff.
The loop visits a shuffled subset of the cells through 1-based offsets; the
cells it skips must come back untouched. Bar: every output array bit for bit,
and the PTE convergence flags.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
CELL_FIELDS = ("press", "pmax", "vol", "spvol", "sie", "temp", "bmod", "dpde", "cv")
FRAC_FIELDS = ("frac_mass", "frac_vol", "frac_ie", "frac_bmod", "frac_dpde", "frac_cv")


@pytest.fixture(scope="module")
def setup(sgm, workloads, oracle, torch_gpu):
    wl = workloads
    ctx, t = wl.pool_context()
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    pool = [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
            oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*wl.VINET_CU), otab.eos()]
    return ctx, pool, otab


def inputs(workloads, ctx, n, seed):
    d = workloads.driver_cells(ctx, n, seed=seed)
    nmat = d["nmat"]
    rng = np.random.default_rng(seed)
    arr = {k: np.zeros(n) for k in CELL_FIELDS}
    arr["spvol"], arr["sie"], arr["temp"], arr["press"] = (d["spvol"], d["sie"], d["temp"],
                                                          d["press"])
    arr["vol"] = rng.uniform(0.5, 2.0, n)
    arr["pmax"] = rng.uniform(0, 1e9, n)
    arr["frac_mass"] = d["frac_mass"]
    for k in FRAC_FIELDS[1:]:
        arr[k] = np.full(n * nmat, -1.0)
    offsets = (rng.permutation(n)[: n - n // 16] + 1).astype(np.int32)
    return arr, d["eos_offsets"], offsets


@pytest.mark.parametrize("mode", [1, 0, -3, -2, -1])
def test_driver_bit_exact_vs_oracle(sgm, workloads, oracle, setup, torch_gpu, mode):
    torch = torch_gpu
    ctx, pool, _ = setup
    n = 3000
    arr, eos_offsets, offsets = inputs(workloads, ctx, n, 0x5EED0000 + 70 + mode)
    cutoff = 1e-6
    params = sgm.mix_params(tight=mode in (0, 1, -3))
    ref = {k: v.copy() for k, v in arr.items()}
    nfail, err, conv_o = oracle.get_sg_eos(pool, mode, eos_offsets, offsets, ref, cutoff,
                                           params.as_oracle_array())
    dev = {k: torch.tensor(v, device="cuda") for k, v in arr.items()}
    conv = torch.full((len(offsets),), -1, dtype=torch.int32, device="cuda")
    rc = sgm.get_sg_eos(ctx, mode, eos_offsets.tolist(),
                        torch.tensor(offsets, device="cuda"), dev, cutoff, params=params,
                        converged=conv)
    assert rc == 0
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in dev.items()}
    for k in CELL_FIELDS + FRAC_FIELDS:
        a, b = got[k], ref[k]
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), (mode, k, int((~same).sum()))
    assert np.array_equal(conv.cpu().numpy(), conv_o)
    # the cells the loop skips are untouched
    skipped = np.setdiff1d(np.arange(n), offsets - 1)
    for k in CELL_FIELDS:
        assert np.array_equal(got[k][skipped], arr[k][skipped]), k
    if mode in (0, 1, -3):
        assert conv_o.mean() > 0.85


def test_driver_rejects_bad_arguments(sgm, setup, torch_gpu):
    torch = torch_gpu
    ctx, _, _ = setup
    z = torch.zeros(8, dtype=torch.float64, device="cuda")
    off = torch.ones(1, dtype=torch.int32, device="cuda")
    arrays = {k: z for k in CELL_FIELDS + FRAC_FIELDS}
    arrays["frac_cv"] = None  # frac_dpde without frac_cv
    assert sgm.get_sg_eos(ctx, 1, [1, 2], off, arrays, 0.0) == -1
    arrays["frac_cv"] = z
    assert sgm.get_sg_eos(ctx, 5, [1, 2], off, arrays, 0.0) == -1  # unknown mode
    assert sgm.get_sg_eos(ctx, 1, [1, 99], off, arrays, 0.0) == -2  # unknown material


def test_driver_reuses_plan_only_for_the_same_topology(sgm, workloads, oracle, setup,
                                                        torch_gpu):
    """The context keeps the plan of the last call's topology: a repeated call
    (same cells and materials, new states) and a call with another topology
    after it both match the oracle bit for bit."""
    torch = torch_gpu
    ctx, pool, _ = setup
    n = 2500
    base, eos_offsets, offsets = inputs(workloads, ctx, n, 0x5EED0000 + 80)
    params = sgm.mix_params(tight=True)
    rng = np.random.default_rng(81)
    for call in range(3):
        arr = {k: v.copy() for k, v in base.items()}
        offs = offsets
        if call == 1:  # same topology, other energies and temperatures
            arr["sie"] = arr["sie"] * rng.uniform(0.9, 1.1, n)
            arr["temp"] = arr["temp"] * rng.uniform(0.9, 1.1, n)
        if call == 2:  # another topology: a different subset of the cells
            offs = offsets[: len(offsets) // 2]
        ref = {k: v.copy() for k, v in arr.items()}
        nfail, err, conv_o = oracle.get_sg_eos(pool, 1, eos_offsets, offs, ref, 1e-6,
                                               params.as_oracle_array())
        dev = {k: torch.tensor(v, device="cuda") for k, v in arr.items()}
        conv = torch.full((len(offs),), -1, dtype=torch.int32, device="cuda")
        assert sgm.get_sg_eos(ctx, 1, eos_offsets.tolist(), torch.tensor(offs, device="cuda"),
                              dev, 1e-6, params=params, converged=conv) == 0
        torch.cuda.synchronize()
        for k in CELL_FIELDS + FRAC_FIELDS:
            a, b = dev[k].cpu().numpy(), ref[k]
            same = (a == b) | (np.isnan(a) & np.isnan(b))
            assert same.all(), (call, k, int((~same).sum()))
        assert np.array_equal(conv.cpu().numpy(), conv_o), call
