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


def cache_info(sgm, ctx):
    import ctypes as C
    out = (C.c_long * 4)()
    assert sgm.lib().sgm_driver_cache_info(ctx.h, out) == 0
    return dict(builds=out[0], reuses=out[1], bytes=out[2], held=out[3])


def test_driver_reuses_plan_only_for_the_same_topology(sgm, workloads, oracle, setup,
                                                        torch_gpu):
    """The context keeps the plan of the last call's topology (the CSR
    offsets and material ids of the participating entries, and the
    formulation). Every call below matches the oracle bit for bit, and the
    context's counters show which calls reused the plan:
      0 first call (build); 1 same topology, other states (reuse);
      2 the same cells visited in reverse order: same ncell and nnz, other
        offsets (build); 3 one cell trades a participating material for one
        that did not take part: same offsets, other material ids (build);
      4 rho-T mode on that topology: another formulation (build);
      5 rho-e again (build); 6 the same call again (reuse);
    then sgm_release_driver_cache frees the plan and buffers and the next
    call builds them again."""
    torch = torch_gpu
    wl = workloads
    ctx, _ = wl.pool_context()
    _, pool, _ = setup
    n = 2500
    base, eos_offsets, offsets = inputs(workloads, ctx, n, 0x5EED0000 + 80)
    nmat = len(eos_offsets)
    params = sgm.mix_params(tight=True)
    rng = np.random.default_rng(81)
    fm = base["frac_mass"].reshape(nmat, n)  # column-major [cell x mat]
    # a visited cell with a participating material p and a missing one q
    visited = offsets - 1
    swap = None
    for c in visited:
        part = np.flatnonzero(fm[:, c] > 1e-6)
        miss = np.flatnonzero(fm[:, c] == 0)
        if len(part) >= 2 and len(miss):
            swap = (c, part[0], miss[0])
            break
    assert swap is not None
    expect = [(1, 0), (1, 1), (2, 1), (3, 1), (4, 1), (5, 1), (5, 2)]
    state = {"arr": base, "offs": offsets}
    for call, (nb, nr) in enumerate(expect):
        arr = {k: v.copy() for k, v in state["arr"].items()}
        offs = state["offs"]
        mode = 1
        if call == 1:
            arr["sie"] = arr["sie"] * rng.uniform(0.9, 1.1, n)
            arr["temp"] = arr["temp"] * rng.uniform(0.9, 1.1, n)
        if call == 2:
            offs = offsets[::-1].copy()
            state["offs"] = offs
        if call == 3:
            c, p_, q_ = swap
            f = arr["frac_mass"].reshape(nmat, n)
            f[q_, c] = f[p_, c]
            f[p_, c] = 0.0
            state["arr"] = {k: v.copy() for k, v in arr.items()}
        if call == 4:
            mode = -3
        params_c = params
        ref = {k: v.copy() for k, v in arr.items()}
        nfail, err, conv_o = oracle.get_sg_eos(pool, mode, eos_offsets, offs, ref, 1e-6,
                                               params_c.as_oracle_array())
        dev = {k: torch.tensor(v, device="cuda") for k, v in arr.items()}
        conv = torch.full((len(offs),), -1, dtype=torch.int32, device="cuda")
        assert sgm.get_sg_eos(ctx, mode, eos_offsets.tolist(), torch.tensor(offs, device="cuda"),
                              dev, 1e-6, params=params_c, converged=conv) == 0
        torch.cuda.synchronize()
        for k in CELL_FIELDS + FRAC_FIELDS:
            a, b = dev[k].cpu().numpy(), ref[k]
            same = (a == b) | (np.isnan(a) & np.isnan(b))
            assert same.all(), (call, k, int((~same).sum()))
        assert np.array_equal(conv.cpu().numpy(), conv_o), call
        info = cache_info(sgm, ctx)
        assert (info["builds"], info["reuses"]) == (nb, nr), (call, info)
        assert info["held"] == 1 and info["bytes"] > 0
    assert sgm.lib().sgm_release_driver_cache(ctx.h) == 0
    info = cache_info(sgm, ctx)
    assert info["held"] == 0 and info["bytes"] == 0
    dev = {k: torch.tensor(v, device="cuda") for k, v in state["arr"].items()}
    assert sgm.get_sg_eos(ctx, 1, eos_offsets.tolist(), torch.tensor(state["offs"], device="cuda"),
                          dev, 1e-6, params=params) == 0
    torch.cuda.synchronize()
    assert cache_info(sgm, ctx)["builds"] == 6
    ctx.close()
