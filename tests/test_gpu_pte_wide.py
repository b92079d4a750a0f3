"""GPU parity of the one-wave-per-cell closure -- cells of 9..64 materials,
9..32 for RhoU -- against the oracle, through the C ABI: the plain per-call
entry points in all five formulations, batched plans whose cells mix the
register-resident counts with wide ones, and get_sg_eos with 14 material
slots in every input mode.

Bar: bit for bit -- states, convergence flags, iteration counts and error
bits -- as for the narrower cells: the one-wave solvers form every value in
the serial order of the oracle's loops, compiled without FMA contraction.

Materials: the six pool materials plus eight variants of the analytic ones
(ideal and stiffened gases, Gruneisen, Vinet with other parameters), so a
cell holds up to 14 distinct materials; cells of 16..64 materials repeat
pool entries (the solver treats each entry as a material of its own). The
64-material cells put the FixedP system at 65 equations, past one column per
lane.
"""
import numpy as np
import pytest

from test_gpu_pte import FIELDS, assert_identical, run_batched, run_oracle, run_plain  # noqa: F401

pytestmark = pytest.mark.gpu

EXTRA = [("ideal", (0.4, 1.0e7)), ("ideal", (0.25, 1.5e7)), ("stiff", (3.0, 3.0e7, 5e9, 0.0)),
         ("gruneisen", (5.0e5, 1.4, 0.0, 0.0, 1.9, 0.0, 2.6, 298.0, 0.0, 8.5e6)),
         ("vinet", (8.5, 298.0, 1.2e12, 4.8, 5.0e-5, 3.9e6, 0.0, 5.0e8)),
         ("stiff", (4.0, 4.2e7, 7e9, 0.0)), ("ideal", (0.5, 5.0e6)), ("ideal", (1.0, 8.0e6))]
EV2K = 1.160451930280894026e4


def ref_density(kind, prm):
    """A condensed material's reference density; None for a gas."""
    return {"ideal": lambda: None, "stiff": lambda: 1.0, "gruneisen": lambda: prm[6],
            "vinet": lambda: prm[0]}[kind]()


def oracle_pool(oracle, wl):
    """The 14 materials as oracle EOS objects, their reference densities (None
    for gases) and the oracle's table; the library's host table comes along
    for the device registry."""
    table = wl.GruneisenTable()
    t = table
    otab, _, _ = oracle.gruneisen_table(wl.GRUN_C1, t.lRhoMin, t.lRhoMax, t.nRho, t.lTMin,
                                        t.lTMax, t.nT)
    o = [oracle.ideal(*wl.IDEAL_A), oracle.ideal(*wl.IDEAL_M1), oracle.stiff(*wl.STIFF_B),
         oracle.gruneisen(*wl.GRUN_C1), oracle.vinet(*wl.VINET_CU), otab.eos()]
    ref = [None, None, 1.0, wl.GRUN_C1[6], wl.VINET_CU[0], 2.7]
    for kind, prm in EXTRA:
        o.append(getattr(oracle, kind)(*prm))
        ref.append(ref_density(kind, prm))
    return o, ref, otab, table


@pytest.fixture(scope="module")
def wide_pool(sgm, workloads, oracle, torch_gpu):
    wl = workloads
    o, ref, otab, table = oracle_pool(oracle, wl)
    ctx = sgm.Context(6 + len(EXTRA))
    ctx.set_ideal(0, *wl.IDEAL_A)
    ctx.set_ideal(1, *wl.IDEAL_M1)
    ctx.set_stiff(2, *wl.STIFF_B)
    ctx.set_gruneisen(3, *wl.GRUN_C1)
    ctx.set_vinet(4, *wl.VINET_CU)
    ctx.set_table(5, table.desc())
    ctx._keep.append(table)
    for i, (kind, prm) in enumerate(EXTRA, start=6):
        getattr(ctx, "set_" + kind)(i, *prm)
    return ctx, o, ref, otab


def wide_cells(o, ref, counts, seed, method=0):
    """CSR cells with the given material counts: per material a density near
    its reference (gases log-normal around 1e-3), an independent temperature
    in [300, 3000] K, volume fractions from exponential draws; the cell input
    is the mass-weighted energy (RhoT, PT), the first temperature (FixedT) or
    the volume-weighted pressure magnitude + 1e5 (FixedP)."""
    rng = np.random.default_rng(seed)
    npool = len(o)
    mats, rho, vf, T, P, off, cin = [], [], [], [], [], [0], []
    for nm in counts:
        ids = rng.choice(npool, size=nm, replace=nm > npool)
        al = rng.exponential(size=nm) + 1e-3
        al /= al.sum()
        rr = np.array([rng.lognormal(np.log(1e-3), 0.3) if ref[i] is None
                       else ref[i] * rng.uniform(0.9, 1.1) for i in ids])
        tt = rng.uniform(300.0, 3000.0, nm)
        ee = np.array([o[i].vec(2, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        pp = np.array([o[i].vec(3, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        rb = rr * al
        if method == 2:
            cin.append(float(tt[0]))
        elif method == 3:
            cin.append(float(abs((al * pp).sum()) + 1e5))
        else:
            cin.append(float((rb * ee).sum() / rb.sum()))
        mats += list(ids)
        rho += list(rr)
        vf += list(al)
        T += list(tt)
        P += list(pp)
        off.append(off[-1] + nm)
    return dict(rho=np.array(rho), vfrac=np.array(vf), sie=np.zeros(len(rho)), temp=np.array(T),
                press=np.array(P), vfrac_tot=np.ones(len(counts)), sie_tot=np.array(cin),
                offsets=np.array(off, dtype=np.int32), mat_ids=np.array(mats, dtype=np.uint8),
                ncell=len(counts), nnz=off[-1])


EDGE_COUNTS = [9, 10, 12, 14, 16, 20, 24, 31, 32, 33, 47, 63, 64]


def edge_counts(method):
    """EDGE_COUNTS within the formulation's one-wave size (RhoU: 32)."""
    return [c for c in EDGE_COUNTS if method != 4 or c <= 32]


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
def test_wide_plain_bit_exact(sgm, oracle, wide_pool, torch_gpu, method):
    ctx, o, ref, _ = wide_pool
    rng = np.random.default_rng(0x5EED0000 + 90 + method)
    counts = edge_counts(method) + list(rng.integers(9, 15, 96))
    a = wide_cells(o, ref, counts, 0x5EED0000 + 91 + method, method)
    for tight in (False, True):
        params = sgm.mix_params(tight=tight)
        print(f"[wide] method {method} tight {tight} cells {a['ncell']}", flush=True)
        g = run_plain(torch_gpu, ctx, a, method, params)
        r = run_oracle(oracle, o, a, method, params.as_oracle_array())
        assert_identical(g, r, f"wide method={method} tight={tight}")
        assert np.array_equal(g["err"], r["err"])
        assert (g["err"] & 32).sum() == 0  # no cell refused as malformed
        if method == 2:
            assert np.array_equal(g["temp"], a["temp"])
        print(f"[wide] method {method} tight {tight} converged {g['conv'].mean():.3f}", flush=True)


def test_wide_rhou_past_32_flagged(sgm, oracle, wide_pool, torch_gpu):
    """RhoU's 2 nm equations fit the one-wave workspace up to 32 materials:
    its wider cells come back flagged as malformed (status bit 32, not
    converged) in the plain path and in a plan, the others solved."""
    ctx, o, ref, _ = wide_pool
    a = wide_cells(o, ref, [3, 33, 2, 40, 32, 64, 9], 0x5EED0000 + 97)
    wide = [1, 3, 5]
    for g in (run_plain(torch_gpu, ctx, a, 4, sgm.mix_params()),
              run_batched(torch_gpu, sgm, ctx, a, 4, sgm.mix_params())[0]):
        assert list(g["err"][wide] & 32) == [32, 32, 32]
        assert list(g["conv"][wide]) == [0, 0, 0]
        assert (g["err"][[0, 2, 4, 6]] & 32).sum() == 0


@pytest.mark.parametrize("method", [0, 2, 4])
def test_wide_cells_in_plans(sgm, oracle, wide_pool, torch_gpu, method):
    """A plan over cells of 1..64 materials: the wide ones form a list bin of
    their own (one wave per cell), the rest the usual bins; results equal the
    plain path and the oracle."""
    ctx, o, ref, _ = wide_pool
    rng = np.random.default_rng(0x5EED0000 + 93 + method)
    counts = list(rng.integers(1, 9, 400)) + edge_counts(method) + list(rng.integers(9, 17, 40))
    rng.shuffle(counts)
    a = wide_cells(o, ref, counts, 0x5EED0000 + 94 + method, method)
    params = sgm.mix_params(tight=True)
    g = run_plain(torch_gpu, ctx, a, method, params)
    r = run_oracle(oracle, o, a, method, params.as_oracle_array())
    assert_identical(g, r, f"plan cells method={method} plain vs oracle")
    b, info = run_batched(torch_gpu, sgm, ctx, a, method, params, replays=2)
    for k, out in enumerate(b):
        assert_identical(out, g, f"plan cells method={method} replay {k}")


def driver_inputs(o, ref, n, seed):
    """get_sg_eos inputs over all 14 materials as slots: per cell 2..14
    distinct materials (every 9th cell one), mass fractions from the cell's
    state, every 5th cell with a trace of a material it lacks."""
    rng = np.random.default_rng(seed)
    nmat = len(o)
    counts = [1 if c % 9 == 0 else int(rng.integers(2, nmat + 1)) for c in range(n)]
    fm = np.zeros((n, nmat))
    spvol, sie, temp, press = (np.zeros(n) for _ in range(4))
    for c, nm in enumerate(counts):
        ids = rng.choice(nmat, size=nm, replace=False)
        al = rng.exponential(size=nm) + 1e-3
        al /= al.sum()
        rr = np.array([rng.lognormal(np.log(1e-3), 0.3) if ref[i] is None
                       else ref[i] * rng.uniform(0.9, 1.1) for i in ids])
        tt = rng.uniform(300.0, 3000.0, nm)
        ee = np.array([o[i].vec(2, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        pp = np.array([o[i].vec(3, [r_], [t_])[0] for i, r_, t_ in zip(ids, rr, tt)])
        rb = rr * al
        x = rb / rb.sum()
        fm[c, ids] = x
        spvol[c] = 1.0 / (rr * al).sum()
        sie[c] = (x * ee).sum()
        temp[c] = (x * tt).sum() / EV2K
        press[c] = abs((al * pp).sum()) + 1e5
        if c % 5 == 0 and nm < nmat:
            fm[c, np.setdiff1d(np.arange(nmat), ids)[0]] = 1e-9
    arr = {k: np.zeros(n) for k in ("press", "pmax", "vol", "spvol", "sie", "temp", "bmod", "dpde",
                                    "cv")}
    arr["spvol"], arr["sie"], arr["temp"], arr["press"] = spvol, sie, temp, press
    arr["vol"] = rng.uniform(0.5, 2.0, n)
    arr["pmax"] = rng.uniform(0, 1e9, n)
    arr["frac_mass"] = np.ascontiguousarray(fm.T).ravel()
    for k in ("frac_vol", "frac_ie", "frac_bmod", "frac_dpde", "frac_cv"):
        arr[k] = np.full(n * nmat, -1.0)
    offsets = (rng.permutation(n)[: n - n // 16] + 1).astype(np.int32)
    return arr, np.arange(1, nmat + 1, dtype=np.int32), offsets, np.array(counts)


@pytest.mark.parametrize("mode", [1, 0, -3, -2, -1])
def test_driver_wide_bit_exact(sgm, oracle, wide_pool, torch_gpu, mode):
    torch = torch_gpu
    ctx, o, ref, _ = wide_pool
    n = 600
    arr, eos_offsets, offsets, counts = driver_inputs(o, ref, n, 0x5EED0000 + 95 + mode)
    assert (counts > 8).sum() > 100
    cutoff = 1e-6
    params = sgm.mix_params(tight=mode in (0, 1, -3))
    refd = {k: v.copy() for k, v in arr.items()}
    print(f"[wide driver] mode {mode} oracle", flush=True)
    nfail, err, conv_o = oracle.get_sg_eos(o, mode, eos_offsets, offsets, refd, cutoff,
                                           params.as_oracle_array())
    dev = {k: torch.tensor(v, device="cuda") for k, v in arr.items()}
    conv = torch.full((len(offsets),), -1, dtype=torch.int32, device="cuda")
    print(f"[wide driver] mode {mode} device", flush=True)
    rc = sgm.get_sg_eos(ctx, mode, eos_offsets.tolist(), torch.tensor(offsets, device="cuda"),
                        dev, cutoff, params=params, converged=conv)
    assert rc == 0
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in dev.items()}
    for k in arr:
        a_, b_ = got[k], refd[k]
        same = (a_ == b_) | (np.isnan(a_) & np.isnan(b_))
        assert same.all(), (mode, k, int((~same).sum()))
    assert np.array_equal(conv.cpu().numpy(), conv_o)
    print(f"[wide driver] mode {mode} converged {conv_o.mean():.3f}", flush=True)


def test_wide_approx_temperature(sgm, oracle, wide_pool, torch_gpu):
    """ApproxTemperatureFromRhoMatU (sgm_approx_temperature) on cells of 9..64
    materials against the oracle's, bit for bit, without and with a guess,
    and with lambda caches (which change no value)."""
    import ctypes as C
    torch = torch_gpu
    ctx, o, ref, _ = wide_pool
    counts = EDGE_COUNTS + [3, 8, 11]
    a = wide_cells(o, ref, counts, 0x5EED0000 + 99)
    n, nnz, off = a["ncell"], a["nnz"], a["offsets"]
    rb = a["rho"] * a["vfrac"]
    cell = np.repeat(np.arange(n), np.diff(off))
    u = np.bincount(cell, weights=rb, minlength=n) * a["sie_tot"]
    d = {k: torch.tensor(a[k], device="cuda", dtype=torch.float64) for k in ("rho", "vfrac")}
    offs = torch.tensor(off, device="cuda", dtype=torch.int32)
    mats = torch.tensor(a["mat_ids"], device="cuda", dtype=torch.uint8)
    ud = torch.tensor(u, device="cuda", dtype=torch.float64)

    def oracle_T(guess):
        out = np.zeros(n)
        for c in range(n):
            s = slice(off[c], off[c + 1])
            ids = a["mat_ids"][s]
            hs = (C.c_void_p * len(ids))(*[o[i].h for i in ids])
            out[c] = oracle.lib().sgo_approx_temperature(
                len(ids), hs, float(u[c]), oracle.p(np.ascontiguousarray(a["rho"][s])),
                oracle.p(np.ascontiguousarray(a["vfrac"][s])), float(guess[c]))
        return out

    r0 = oracle_T(np.zeros(n))
    for guess, lam in ((None, None), (r0 * 1.3, None), (None, True)):
        T = torch.zeros(n, dtype=torch.float64, device="cuda")
        gd = torch.tensor(guess, device="cuda") if guess is not None else None
        ld = torch.zeros(3 * nnz, dtype=torch.float64, device="cuda") if lam else None
        assert ctx.approx_temperature(n, nnz, offs, mats, d["rho"], d["vfrac"], ud, T,
                                      Tguess=gd, lam=ld) == 0
        torch.cuda.synchronize()
        want = r0 if guess is None else oracle_T(guess)
        assert np.array_equal(T.cpu().numpy(), want)
