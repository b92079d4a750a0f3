"""ORACLE — test infrastructure only. FP64 numpy restatement of the learned
initial guess network (csrc/sgm_mlp.hip): the same parameter layout, the
same masked-softmax + ln T loss, forward and backward by hand, used to check
the device kernels' outputs and gradients. BF16 rounding of weights and
activations is applied where the device applies it, so the only remaining
differences are FP32 accumulation order and BF16 rounding of the gradients
fed to the matrix cores."""
import numpy as np

KIN, H, KOUT = 64, 128, 16
OW1 = 0
OB1 = OW1 + H * KIN
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * H
OW4 = OB3 + H
OB4 = OW4 + KOUT * H
NPARAM = OB4 + KOUT


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits_to_f64(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(
        np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    W = [P[OW1:OB1].reshape(H, KIN), P[OW2:OB2].reshape(H, H), P[OW3:OB3].reshape(H, H),
         P[OW4:OB4].reshape(KOUT, H)]
    b = [P[OB1:OW2], P[OB2:OW3], P[OB3:OW4], P[OB4:NPARAM]]
    return W, b


def gelu(z):
    k, c = 0.7978845608028654, 0.044715
    return 0.5 * z * (1 + np.tanh(k * (z + c * z ** 3)))


def gelu_d(z):
    k, c = 0.7978845608028654, 0.044715
    t = np.tanh(k * (z + c * z ** 3))
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * k * (1 + 3 * c * z * z)


def forward(P, X, quant=True):
    """X: [B, 64] (already bf16 values). -> Y [B, 16], cache. quant=False
    leaves out the bf16 rounding of weights and activations (the plain FP64
    network, whose gradient finite differences can check)."""
    W, b = unpack(P)
    q = bf16 if quant else (lambda v: np.asarray(v, dtype=np.float64))
    Wq = [q(w) for w in W]
    A = [np.asarray(X, dtype=np.float64)]
    Z = [None]
    for l in range(3):
        z = A[-1] @ Wq[l].T + b[l]
        Z.append(z)
        A.append(q(gelu(z)))
    Y = A[-1] @ Wq[3].T + b[3]
    return Y, (A, Z, Wq)


def loss_grad(Y, T, wT=1.0):
    """T: [B, 8] targets (alphas, standardized ln T, nmat). Squared error of
    the log volume fractions (log-softmax over the cell's materials, targets
    floored at 1e-8) + wT (ln T error)^2. -> loss per row, dY (mean over B)"""
    B = len(Y)
    dY = np.zeros_like(Y)
    L = np.zeros(B)
    for r in range(B):
        nm = int(T[r, 7])
        if nm < 2 or nm > 6:
            continue
        y = Y[r, :nm]
        lse = y.max() + np.log(np.exp(y - y.max()).sum())
        p = np.exp(y - lse)
        e = (y - lse) - np.log(np.maximum(T[r, :nm], 1e-8))
        dY[r, :nm] = 2 * (e - p * e.sum()) / B
        et = Y[r, 6] - T[r, 6]
        dY[r, 6] = 2 * wT * et / B
        L[r] = (e * e).sum() + wT * et * et
    return L, dY


def gradients(P, X, T, wT=1.0, quant=True):
    """Flat gradient of the mean loss in the device parameter layout."""
    Y, (A, Z, Wq) = forward(P, X, quant)
    L, dZ = loss_grad(Y, T, wT)
    g = np.zeros(NPARAM)
    offs = [(OW1, OB1), (OW2, OB2), (OW3, OB3), (OW4, OB4)]
    for l in range(3, -1, -1):
        ow, ob = offs[l]
        g[ow:ow + Wq[l].size] = (dZ.T @ A[l]).ravel()
        g[ob:ob + Wq[l].shape[0]] = dZ.sum(0)
        if l > 0:
            dZ = (dZ @ Wq[l]) * gelu_d(Z[l])
    return g, L
