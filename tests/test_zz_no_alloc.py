"""sgm_get_sg_eos_noalloc, the get_sg_eos entry point that allocates
nothing (its work arrays live in a static device buffer), against
sgm_get_sg_eos on the same inputs: every output array and the convergence
flags bit for bit, on two calls in a row (the second smaller, so it reuses
the buffer the first left behind), in the rho-e, rho-T and P-T modes; a call
larger than the buffer is refused with SGM_E_BUDGET and touches nothing."""
import numpy as np
import pytest

from test_gpu_driver import CELL_FIELDS, FRAC_FIELDS, inputs

pytestmark = pytest.mark.gpu


def run(sgm, torch, ctx, mode, arr, eos_offsets, offsets, params, noalloc):
    dev = {k: torch.tensor(v, device="cuda") for k, v in arr.items()}
    conv = torch.full((len(offsets),), -1, dtype=torch.int32, device="cuda")
    rc = sgm.get_sg_eos(ctx, mode, eos_offsets.tolist(), torch.tensor(offsets, device="cuda"),
                        dev, 1e-6, params=params, converged=conv, noalloc=noalloc)
    torch.cuda.synchronize()
    return rc, {k: v.cpu().numpy() for k, v in dev.items()}, conv.cpu().numpy()


@pytest.mark.parametrize("mode", [1, -3, -1])
def test_noalloc_equals_get_sg_eos(sgm, workloads, torch_gpu, mode):
    torch = torch_gpu
    ctx, _ = workloads.pool_context()
    params = sgm.mix_params(tight=mode != -1)
    for n, seed in ((3000, 90), (1000, 91)):
        arr, eos_offsets, offsets = inputs(workloads, ctx, n, 0x5EED0000 + seed + 10 * mode)
        rc0, ref, conv0 = run(sgm, torch, ctx, mode, arr, eos_offsets, offsets, params, False)
        rc1, got, conv1 = run(sgm, torch, ctx, mode, arr, eos_offsets, offsets, params, True)
        assert rc0 == 0 and rc1 == 0
        for k in CELL_FIELDS + FRAC_FIELDS:
            assert np.array_equal(got[k], ref[k], equal_nan=True), (mode, n, k)
        assert np.array_equal(conv1, conv0), (mode, n)
    ctx.close()


def test_noalloc_refuses_calls_beyond_its_buffer(sgm, workloads, torch_gpu):
    torch = torch_gpu
    ctx, _ = workloads.pool_context()
    n = 200000  # 1.2e6 entries of 6 materials: more than the 64 MiB buffer holds
    arrays = {k: torch.full((n,), 7.0, dtype=torch.float64, device="cuda") for k in CELL_FIELDS}
    for k in FRAC_FIELDS:
        arrays[k] = torch.full((n * 6,), 7.0, dtype=torch.float64, device="cuda")
    off = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
    rc = sgm.get_sg_eos(ctx, 1, [1, 2, 3, 4, 5, 6], off, arrays, 1e-6, noalloc=True)
    assert rc == -6
    torch.cuda.synchronize()
    for k, v in arrays.items():
        assert bool((v == 7.0).all()), k
    ctx.close()
