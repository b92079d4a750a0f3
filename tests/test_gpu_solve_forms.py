"""The closure's dense linear solve in its two device forms: solve_small
(every lane on its own copy: the register-resident bin kernels) and
solve_small_wave (a whole wave, lane j holding column j, pivoting through a
position map: the whole-wave straggler tails). The reference solves these
systems in solve_Ax_b_wscr (singularity-eos/closure/mixed_cell_models.hpp:176-278:
Jacobi row scaling, then Eigen's colPivHouseholderQr for n >= 3). Bar: both
forms bit-identical to each other and to the oracle's restatement, including
the failure flag, on every lane of the wave, for 3..8 equations, on systems
that reach every branch: well conditioned, column scales over twelve decades,
near-singular, exactly rank-deficient (the zero-pivot cut), a zero row
(failure), and tiny-tail columns (the reflection skipped)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def forms(sgm):
    path = os.path.join(os.path.dirname(sgm.LIB_PATH), "libsgm_device_demo.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built; run __graft_entry__.build()")
    sgm.lib()
    f = C.CDLL(path).sgm_demo_solve_forms
    vp = C.c_void_p
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_long] + [vp] * 6 + [vp]
    return f


def systems(n, rng, per_kind=60):
    out = []
    for _ in range(per_kind):
        out.append(rng.normal(size=(n, n)) + n * np.eye(n))
        out.append(rng.normal(size=(n, n)) * 10.0 ** rng.uniform(-6, 6, size=n))
        A = rng.normal(size=(n, n))
        A[:, -1] = A[:, 0] + 1e-13 * rng.normal(size=n)  # near-singular
        out.append(A)
        A = rng.normal(size=(n, n))
        A[:, 1] = 2.0 * A[:, 0]  # rank-deficient
        A[:, -1] = A[:, 0] - A[:, 1]
        out.append(A)
        A = rng.normal(size=(n, n))
        A[rng.integers(n), :] = 0.0  # zero row: both forms fail
        out.append(A)
        A = np.triu(rng.normal(size=(n, n)))  # tails below the diagonal vanish
        A[np.arange(n), np.arange(n)] += 1.0
        out.append(A)
        # rows of a closure Jacobian: pressure rows near 1e10, the energy row near 1
        A = rng.normal(size=(n, n)) * np.r_[np.full(n - 1, 1e10), 1.0][:, None]
        out.append(A)
    return np.array(out)


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7, 8])
def test_solve_forms_bit_identical(sgm, oracle, forms, torch_gpu, n):
    torch = torch_gpu
    rng = np.random.default_rng(900 + n)
    A = systems(n, rng)
    b = rng.normal(size=(len(A), n))
    b[::5] *= 1e8
    count = len(A)
    dA = torch.tensor(A.reshape(count, n * n), dtype=torch.float64, device="cuda")
    db = torch.tensor(b, dtype=torch.float64, device="cuda")
    xs = torch.full((count, n), -7.0, dtype=torch.float64, device="cuda")
    xw = torch.full((count, n), -7.0, dtype=torch.float64, device="cuda")
    ok = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    lanes = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    print(f"[solve forms] n={n} systems={count}", flush=True)
    p = sgm.ptr
    assert forms(n, count, p(dA), p(db), p(xs), p(xw), p(ok), p(lanes), None) == 0
    torch.cuda.synchronize()
    xs, xw, ok, lanes = (t.cpu().numpy() for t in (xs, xw, ok, lanes))
    assert np.all(lanes == 0)
    serial_ok, wave_ok = (ok & 1) != 0, (ok & 2) != 0
    assert np.array_equal(serial_ok, wave_ok)
    assert np.array_equal(xs[serial_ok].view(np.uint64), xw[serial_ok].view(np.uint64))
    # the oracle's solve_Ax_b_wscr on the same systems, bit for bit
    lib = oracle.lib()
    for k in range(count):
        a = np.ascontiguousarray(A[k]).ravel().copy()
        x = b[k].copy()
        ok_o = bool(lib.sgo_solve(n, a.ctypes.data, x.ctypes.data))
        assert ok_o == serial_ok[k], k
        if ok_o:
            assert np.array_equal(x.view(np.uint64), xs[k].view(np.uint64)), k
    # every kind of system was there: failures (zero rows) and successes
    assert (~serial_ok).sum() >= 60 and serial_ok.sum() >= 6 * 60 - 5


def test_solve_forms_rejects_bad_sizes(forms):
    assert forms(2, 1, None, None, None, None, None, None, None) != 0
    assert forms(9, 1, None, None, None, None, None, None, None) != 0
    assert forms(5, 0, None, None, None, None, None, None, None) == 0
