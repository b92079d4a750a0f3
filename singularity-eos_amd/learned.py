"""Learned initial guess for the closure (the sgm_mlp_* kernels).

Training data are closures the solver itself solved: features come from a
cell's inputs, targets from its converged state. Everything runs on the
device through libsgm; torch only holds the device buffers.
"""
import ctypes as C

import numpy as np

from . import Plan, check, lib, ptr

FIELDS = ("rho", "vfrac", "sie", "temp", "press")
KIN, KOUT = 64, 16


class MLP:
    def __init__(self, seed=1):
        self.h = C.c_void_p()
        check(lib().sgm_mlp_create(seed, C.byref(self.h)), "mlp_create")
        self.nparam = lib().sgm_mlp_nparam()
        self.mu = np.zeros(KIN, dtype=np.float32)
        self.sigma = np.ones(KIN, dtype=np.float32)
        self.t_mu, self.t_sigma = 0.0, 1.0

    def close(self):
        if self.h:
            lib().sgm_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        check(lib().sgm_mlp_params(self.h, out.ctypes.data, None), "mlp_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().sgm_mlp_params(self.h, None, p.ctypes.data), "mlp_params")

    def set_norm(self, mu, sigma, t_mu, t_sigma):
        self.mu = np.ascontiguousarray(mu, dtype=np.float32)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float32)
        self.t_mu, self.t_sigma = float(t_mu), float(t_sigma)
        check(lib().sgm_mlp_set_norm(self.h, self.mu.ctypes.data, self.sigma.ctypes.data,
                                     self.t_mu, self.t_sigma), "mlp_set_norm")

    def features(self, torch, d, ncell, nnz, raw=False, stream=0):
        """d: device arrays of a closure's inputs. -> (X bf16 as int16 [n,64] or
        raw float32 [n,64], nmat int32 [n])"""
        nm = torch.zeros(ncell, dtype=torch.int32, device="cuda")
        if raw:
            X = torch.zeros((ncell, KIN), dtype=torch.float32, device="cuda")
            args = (ptr(X), None)
        else:
            X = torch.zeros((ncell, KIN), dtype=torch.int16, device="cuda")
            args = (None, ptr(X))
        check(lib().sgm_mlp_features(self.h, ncell, nnz, ptr(d["offsets"]), ptr(d["mat_ids"]),
                                     ptr(d["rho"]), ptr(d["vfrac"]), ptr(d["temp"]),
                                     ptr(d["sie_tot"]), *args, ptr(nm), stream), "features")
        return X, nm

    def targets(self, torch, d, out, conv, nmat, ncell, nnz, stream=0):
        t = torch.zeros((ncell, 8), dtype=torch.float32, device="cuda")
        check(lib().sgm_mlp_targets(self.h, ncell, nnz, ptr(d["offsets"]), ptr(out["vfrac"]),
                                    ptr(out["temp"]), ptr(d["vfrac_tot"]), ptr(nmat), ptr(conv),
                                    ptr(t), stream), "targets")
        return t

    def gradients(self, torch, X, tgt, idx, wT=1.0, stream=0):
        g = torch.zeros(self.nparam, dtype=torch.float32, device="cuda")
        loss = torch.zeros(len(idx), dtype=torch.float32, device="cuda")
        check(lib().sgm_mlp_gradients(self.h, ptr(X), ptr(tgt), ptr(idx), len(idx), wT, ptr(g),
                                      ptr(loss), stream), "gradients")
        return g, loss

    def train_step(self, X, tgt, idx, lr, wT=1.0, loss=None, stream=0):
        check(lib().sgm_mlp_train_step(self.h, ptr(X), ptr(tgt), ptr(idx), len(idx), wT, lr,
                                       ptr(loss), stream), "train_step")

    def forward(self, torch, X, idx, stream=0):
        Y = torch.zeros((len(idx), KOUT), dtype=torch.float32, device="cuda")
        check(lib().sgm_mlp_forward(self.h, ptr(X), ptr(idx), len(idx), ptr(Y), stream),
              "forward")
        return Y

    def guess(self, torch, ctx, d, ncell, nnz, stream=0):
        gv = torch.zeros(nnz, dtype=torch.float64, device="cuda")
        gT = torch.zeros(ncell, dtype=torch.float64, device="cuda")
        check(lib().sgm_mlp_guess(self.h, ctx.h, ncell, nnz, ptr(d["offsets"]), ptr(d["mat_ids"]),
                                  ptr(d["rho"]), ptr(d["vfrac"]), ptr(d["temp"]),
                                  ptr(d["vfrac_tot"]), ptr(d["sie_tot"]), ptr(gv), ptr(gT),
                                  stream), "guess")
        return gv, gT


def to_device(torch, a):
    d = {k: torch.tensor(a[k], device="cuda", dtype=torch.float64)
         for k in FIELDS + ("vfrac_tot", "sie_tot")}
    d["offsets"] = torch.tensor(a["offsets"], device="cuda", dtype=torch.int32)
    d["mat_ids"] = torch.tensor(a["mat_ids"], device="cuda", dtype=torch.uint8)
    return d


def solve(torch, ctx, d, ncell, nnz, params, plan=None, guess=None):
    """Batched closure on copies of d's inputs -> (outputs, conv, niter, used)."""
    out = {k: d[k].clone() for k in FIELDS}
    conv = torch.zeros(ncell, dtype=torch.int32, device="cuda")
    niter = torch.zeros(ncell, dtype=torch.int32, device="cuda")
    used = torch.zeros(ncell, dtype=torch.int32, device="cuda")
    own = plan is None
    if own:
        plan = Plan(ctx, ncell, nnz, d["offsets"], d["mat_ids"], 0)
    if guess is None:
        rc = plan.execute(out["rho"], out["vfrac"], out["sie"], out["temp"], out["press"],
                          d["vfrac_tot"], d["sie_tot"], params=params, conv=conv, niter=niter)
    else:
        rc = plan.execute_guided(out["rho"], out["vfrac"], out["sie"], out["temp"],
                                 out["press"], d["vfrac_tot"], d["sie_tot"], guess[0], guess[1],
                                 params=params, conv=conv, niter=niter, used=used)
    check(rc, "plan execute")
    torch.cuda.synchronize()
    if own:
        plan.close()
    return out, conv, niter, used


def build_training_set(torch, mlp, ctx, a, params, log=None):
    """Features and targets of a solved workload; sets the model's
    standardization from the data. -> (X, targets, valid row indices)"""
    d = to_device(torch, a)
    n, nnz = a["ncell"], a["nnz"]
    raw, nm = mlp.features(torch, d, n, nnz, raw=True)
    out, conv, niter, _ = solve(torch, ctx, d, n, nnz, params)
    ok = ((nm > 0) & (conv > 0)).cpu().numpy()
    R = raw.cpu().numpy()[ok]
    mu = np.zeros(KIN, dtype=np.float32)
    sg = np.ones(KIN, dtype=np.float32)
    cont = [9 * s + j for s in range(6) for j in (6, 7, 8)] + [54, 55]
    nmv = nm.cpu().numpy()[ok]
    for i in cont:
        s_slot = i // 9 if i < 54 else None
        sel = nmv > s_slot if s_slot is not None else np.ones(len(R), dtype=bool)
        v = R[sel, i]
        mu[i] = float(v.mean())
        sg[i] = float(max(v.std(), 1e-3))
    lnT = np.log(out["temp"].cpu().numpy()[a["offsets"][:-1]][ok])
    mlp.set_norm(mu, sg, float(lnT.mean()), float(max(lnT.std(), 1e-3)))
    X, nm2 = mlp.features(torch, d, n, nnz)
    tgt = mlp.targets(torch, d, out, conv, nm2, n, nnz)
    rows = torch.tensor(np.flatnonzero(ok).astype(np.int32), device="cuda")
    if log:
        log(f"[learned] training set: {int(ok.sum())} of {n} cells "
            f"(mean Newton iterations {float(niter.float().mean()):.2f})")
    return X, tgt, rows


def train(torch, mlp, X, tgt, rows, steps=2000, batch=4096, lr=2e-3, seed=0, log=None):
    """Adam on random batches of the valid rows; a cosine learning-rate decay."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    loss = torch.zeros(batch, dtype=torch.float32, device="cuda")
    hist = []
    nrow = rows.numel()
    for it in range(steps):
        pick = torch.randint(0, nrow, (batch,), device="cuda", generator=g)
        idx = rows[pick].contiguous()
        lr_t = lr * 0.5 * (1 + np.cos(np.pi * it / steps))
        mlp.train_step(X, tgt, idx, lr_t, loss=loss)
        if it % 200 == 0 or it == steps - 1:
            hist.append((it, float(loss.mean())))
            if log:
                log(f"[learned] step {it}: loss {hist[-1][1]:.5f}")
    return hist
