"""Pins the FP64 restatement of the learned-guess network (oracle/mlp_ref.py)
that the device MLP kernels are checked against: its hand-derived backward
pass against central finite differences of its own forward loss, with the
bf16 rounding switched off (rounding makes the loss piecewise constant in the
weights), on parameters of every layer, weights and biases, to 1e-6
relative. A planted mistake (the GELU itself in place of its derivative, on
the CPU) fails the same check."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref():
    return importlib.import_module("mlp_ref")


def batch(ref, B, seed):
    rng = np.random.default_rng(seed)
    P = rng.normal(0, 1, ref.NPARAM)
    fan = [ref.KIN, ref.H, ref.H, ref.H]
    for (ow, ob), f in zip(((ref.OW1, ref.OB1), (ref.OW2, ref.OB2), (ref.OW3, ref.OB3),
                            (ref.OW4, ref.OB4)), fan):
        P[ow:ob] *= 1.0 / np.sqrt(f)
    P[ref.OB1:ref.OW2] *= 0.1
    X = rng.normal(0, 1, (B, ref.KIN))
    T = np.zeros((B, 8))
    nm = rng.integers(2, 7, B)
    for r in range(B):
        a = rng.exponential(size=nm[r])
        T[r, :nm[r]] = a / a.sum()
        T[r, 6] = rng.normal()
        T[r, 7] = nm[r]
    return P, X, T


def mean_loss(ref, P, X, T, wT):
    Y, _ = ref.forward(P, X, quant=False)
    L, _ = ref.loss_grad(Y, T, wT)
    return L.sum() / len(L)


def check(ref, wT, seed):
    P, X, T = batch(ref, 48, seed)
    g, _ = ref.gradients(P, X, T, wT, quant=False)
    rng = np.random.default_rng(seed + 100)
    layers = [(ref.OW1, ref.OB1), (ref.OB1, ref.OW2), (ref.OW2, ref.OB2), (ref.OB2, ref.OW3),
              (ref.OW3, ref.OB3), (ref.OB3, ref.OW4), (ref.OW4, ref.OB4), (ref.OB4, ref.NPARAM)]
    idx = []
    for lo, hi in layers:
        idx += list(rng.integers(lo, hi, 8))
    # output rows 7..15 of the last layer do not enter the loss: zero gradient
    idx += [ref.OW4 + 10 * ref.H + 3, ref.OB4 + 12]
    worst = 0.0
    for i in idx:
        h = 1e-6 * max(1.0, abs(P[i]))
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mean_loss(ref, Pp, X, T, wT) - mean_loss(ref, Pm, X, T, wT)) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(abs(fd), 1e-3 * np.abs(g).max()))
    return worst, g


@pytest.mark.parametrize("wT", [1.0, 0.25])
def test_mlp_reference_gradients_match_finite_differences(ref, wT):
    worst, g = check(ref, wT, 5)
    assert worst < 1e-6, worst
    assert g[ref.OW4 + 10 * ref.H + 3] == 0.0 and g[ref.OB4 + 12] == 0.0


def test_mlp_gradient_check_catches_a_planted_mistake(ref, monkeypatch):
    monkeypatch.setattr(ref, "gelu_d", ref.gelu)
    worst, _ = check(ref, 1.0, 5)
    assert worst > 1e-3
