"""The device-side closure interface (include/sgm_device.hpp): a caller's
own HIP kernel (singularity-eos_amd/device_demo.hip, built into
libsgm_device_demo.so) constructs PTESolverRhoT / PTESolverPT per cell with
pointer indexers and calls PTESolver, as hydro codes call the reference's
device-callable objects. Its results equal the library's batch entry points
(sgm_pte_rhoT / sgm_pte_PT) bit for bit — states, convergence flags and
iteration counts — on the uneven workload's 2..6-material cells, and with
the temperature guess from the device ApproxTemperatureFromRhoMatU on the
reference's two-phase tin cells."""
import ctypes as C
import os

import numpy as np
import pytest

import _sn2phase as SN
from test_gpu_pte import FIELDS, assert_identical, dev, run_plain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def demo(sgm):
    path = os.path.join(os.path.dirname(sgm.LIB_PATH), "libsgm_device_demo.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built; run __graft_entry__.build()")
    sgm.lib()
    L = C.CDLL(path)
    f = L.sgm_demo_pte_per_cell
    vp = C.c_void_p
    f.restype = C.c_int
    f.argtypes = [vp, C.c_int, C.c_long] + [vp] * 10 + [C.c_int, C.POINTER(sgm.MixParams)] + \
        [vp] * 5 + [vp]
    return f


def run_demo(torch, sgm, f, ctx, a, method, params, approx=False, Tnorm=None):
    d = dev(torch, a)
    n = a["ncell"]
    slow = torch.zeros(n, dtype=torch.int32, device="cuda")
    mline = torch.zeros(n, dtype=torch.int32, device="cuda")
    p = sgm.ptr
    rc = f(ctx.h, method, n, p(d["offsets"]), p(d["mat_ids"]), p(d["rho"]), p(d["vfrac"]),
           p(d["sie"]), p(d["temp"]), p(d["press"]), p(d["vfrac_tot"]), p(d["sie_tot"]),
           p(Tnorm), 1 if approx else 0, C.byref(params), p(d["conv"]), p(d["niter"]),
           p(d["resid"]), p(slow), p(mline), None)
    assert rc == 0
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in d.items()}
    out["slow"] = slow.cpu().numpy()
    out["max_line"] = mline.cpu().numpy()
    return out


@pytest.mark.parametrize("method,tight", [(0, False), (0, True), (1, True)])
def test_device_interface_equals_batch_entry_points(sgm, workloads, demo, torch_gpu, method,
                                                    tight):
    ctx, _ = workloads.pool_context()
    a = workloads.uneven(ctx, 8192, seed=0x5EED0000 + 41)
    params = sgm.mix_params(tight=tight)
    g = run_plain(torch_gpu, ctx, a, method, params)
    u = run_demo(torch_gpu, sgm, demo, ctx, a, method, params)
    assert_identical(u, g, f"device api m{method}")
    assert np.array_equal(u["resid"], g["resid"])
    # SolverStatus fields the batch calls do not report: slow convergence is
    # only ever flagged on converged cells; line-search counts stay below the cap
    assert np.all(u["slow"] <= u["conv"])
    assert u["max_line"].max() < params.line_search_max_iter
    ctx.close()


def test_device_interface_sn_fixture_with_approx_temperature(sgm, demo, torch_gpu):
    torch = torch_gpu
    ctx = sgm.Context(2)
    ctx.set_vinet(0, *SN.vinet_args("beta"))
    ctx.set_vinet(1, *SN.vinet_args("gamma"))
    a = SN.cells("long")
    n = a["ncell"]
    params = sgm.mix_params()
    u = run_demo(torch, sgm, demo, ctx, a, 0, params, approx=True)
    # the batch path with Tnorm from sgm_approx_temperature
    d = dev(torch, a)
    ut = torch.tensor(a["u_tot"], device="cuda")
    Tg = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert ctx.approx_temperature(n, a["nnz"], d["offsets"], d["mat_ids"], d["rho"], d["vfrac"],
                                  ut, Tg) == 0
    rc = ctx.pte(0, n, a["nnz"], d["offsets"], d["mat_ids"], d["rho"], d["vfrac"], d["sie"],
                 d["temp"], d["press"], d["vfrac_tot"], d["sie_tot"], Tnorm=Tg, params=params,
                 conv=d["conv"], niter=d["niter"], resid=d["resid"], err=d["err"])
    assert rc == 0
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in d.items()}
    assert_identical(u, g, "device api sn")
    assert u["conv"].all()
    ctx.close()
