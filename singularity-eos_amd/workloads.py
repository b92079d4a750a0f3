"""Synthetic workloads for the EOS / closure hot path (host numpy arrays).

All random draws come from libsgm's mt19937_64 generators (include/sgm.h,
sgm_workload_*), so the same seed gives the same inputs everywhere.

  pte2     65,536 two-material cells: IdealGas(0.4, 7.176e6) + StiffGas(3.4, 3.6e7, 6e9, 0)
  table    256 x 512 rho-T table sampled from Gruneisen(rho0=2.7, C0=5.35e5, s1=1.34,
           G0=2.0, Cv=9e6; unlisted s2=s3=b=P0=0, T0=298) on NQT-log10 grids
           rho in [1e-3, 1e2], T in [1e2, 1e6]; 2^24 log-uniform queries
  uneven   2^20 cells with 2..6 materials from a pool of six
"""
import ctypes as C

import numpy as np

from . import Context, TableDesc, lib, check

SEED_BASE = 0x5EED0000

IDEAL_A = (0.4, 7.176e6)
STIFF_B = (3.4, 3.6e7, 6e9, 0.0)
IDEAL_M1 = (2.0 / 3.0, 3.1e7)
# Gruneisen used for the table and for the pool's condensed slot:
# (C0, s1, s2, s3, G0, b, rho0, T0, P0, Cv). Parameters the workload does not
# list are zero (T0 = 298 K). b = 0 matters: with the Al test's b = 0.48
# (test_eos_gruneisen.cpp:282-291) the Hugoniot term P_H (1 - Gamma rho eta /
# (2 rho0)) turns negative above rho ~ 8.2 g/cc, inside the table's range, the
# table's density of minimum pressure becomes ~10.9 g/cc at every T, and
# the closure bounds every table material above it: no table cell can reach
# equilibrium.
GRUN_C1 = (5.35e5, 1.34, 0.0, 0.0, 2.0, 0.0, 2.7, 298.0, 0.0, 9.0e6)
GRUN_AL_TEST_B = 0.48
# Vinet Cu-like: rho0, T0, B0, BP0, A0, Cv0, E0, S0
VINET_CU = (8.93, 298.0, 1.37e12, 5.0, 5.19245e-5, 3.83e6, 0.0, 5.05e8)
TABLE_NRHO, TABLE_NT = 256, 512


class GruneisenTable:
    """Host arrays of the Gruneisen-sampled rho-T table plus its descriptor."""

    def __init__(self, nrho=TABLE_NRHO, nt=TABLE_NT, grun=GRUN_C1, lrho=None, lt=None):
        from .nqt import log10 as nlog10
        self.lRhoMin, self.lRhoMax = lrho if lrho else (nlog10(1e-3), nlog10(1e2))
        self.lTMin, self.lTMax = lt if lt else (nlog10(1e2), nlog10(1e6))
        self.nRho, self.nT = nrho, nt
        n2 = nrho * nt
        self.f = {k: np.zeros(n2) for k in ("P", "sie", "bMod", "dPdRho", "dPdE", "dEdRho", "dEdT")}
        self.c = {k: np.zeros(nrho) for k in ("PCold", "sieCold", "bModCold")}
        args = list(grun) + [self.lRhoMin, self.lRhoMax, nrho, self.lTMin, self.lTMax, nt]
        ptrs = [self.f[k].ctypes.data for k in ("P", "sie", "bMod", "dPdRho", "dPdE", "dEdRho",
                                                "dEdT")]
        ptrs += [self.c[k].ctypes.data for k in ("PCold", "sieCold", "bModCold")]
        check(lib().sgm_build_gruneisen_table(*args, *ptrs), "build_gruneisen_table")

    def desc(self):
        d = TableDesc()
        d.lRhoMin, d.lRhoMax, d.nRho = self.lRhoMin, self.lRhoMax, self.nRho
        d.lTMin, d.lTMax, d.nT = self.lTMin, self.lTMax, self.nT
        d.rho_offset = d.T_offset = 0.0
        for k, v in self.f.items():
            setattr(d, k, v.ctypes.data)
        for k, v in self.c.items():
            setattr(d, k, v.ctypes.data)
        d.reproducible = 0
        d.pmin_vapor_dome = 0
        return d


def pool_context(table=None, host_only=False):
    """Context with the six pool materials in slots 0..5:
    0 IdealGas(0.4), 1 IdealGas(2/3), 2 StiffGas water-like, 3 Gruneisen Al-like,
    4 Vinet Cu-like, 5 the Gruneisen-sampled table."""
    ctx = Context(6, host_only=host_only)
    ctx.set_ideal(0, *IDEAL_A)
    ctx.set_ideal(1, *IDEAL_M1)
    ctx.set_stiff(2, *STIFF_B)
    ctx.set_gruneisen(3, *GRUN_C1)
    ctx.set_vinet(4, *VINET_CU)
    table = table or GruneisenTable()
    ctx.set_table(5, table.desc())
    ctx._keep.append(table)
    return ctx, table


def pte2(n, seed=SEED_BASE + 0):
    """Two-material inputs: dict of host arrays (CSR with 2 materials per cell)."""
    a = {k: np.zeros(2 * n) for k in ("rho", "vfrac", "sie", "temp", "press")}
    a["sie_tot"] = np.zeros(n)
    check(lib().sgm_workload_pte2(seed, n, a["rho"].ctypes.data, a["vfrac"].ctypes.data,
                                  a["sie"].ctypes.data, a["temp"].ctypes.data,
                                  a["press"].ctypes.data, a["sie_tot"].ctypes.data), "workload_pte2")
    a["vfrac_tot"] = np.ones(n)
    a["offsets"] = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    a["mat_ids"] = np.tile(np.array([0, 1], dtype=np.uint8), n)
    a["ncell"], a["nnz"] = n, 2 * n
    return a


def uneven(ctx, n, seed=SEED_BASE + 2, pool=(0, 1, 2, 3, 4, 5)):
    cap = 6 * n
    a = {k: np.zeros(cap) for k in ("rho", "vfrac", "sie", "temp", "press")}
    a["offsets"] = np.zeros(n + 1, dtype=np.int32)
    a["mat_ids"] = np.zeros(cap, dtype=np.uint8)
    a["sie_tot"] = np.zeros(n)
    pl = (C.c_int * 6)(*pool)
    nnz = lib().sgm_workload_uneven(ctx.h, seed, n, C.cast(pl, C.c_void_p),
                                    a["offsets"].ctypes.data, a["mat_ids"].ctypes.data,
                                    a["rho"].ctypes.data, a["vfrac"].ctypes.data,
                                    a["sie"].ctypes.data, a["temp"].ctypes.data,
                                    a["press"].ctypes.data, a["sie_tot"].ctypes.data)
    if nnz < 0:
        raise RuntimeError(f"workload_uneven {nnz}")
    for k in ("rho", "vfrac", "sie", "temp", "press"):
        a[k] = a[k][:nnz].copy()
    a["mat_ids"] = a["mat_ids"][:nnz].copy()
    a["vfrac_tot"] = np.ones(n)
    a["ncell"], a["nnz"] = n, int(nnz)
    return a


EV2K = 1.160451930280894026e4  # kelvin per eV (singularity-eos/eos/get_sg_eos.hpp:68)


def driver_cells(ctx, n, seed=SEED_BASE + 3, pure_every=7, cutoff_trace=1e-9):
    """Inputs of the get_sg_eos closure driver from the uneven workload's cells:
    per cell the specific volume 1 / sum(rho_m vfrac_m), the mass-weighted
    specific energy, a mass-weighted temperature guess (eV) and the
    volume-weighted pressure (absolute value + 1e5); per material (six slots =
    the six pool materials, column-major [n x 6]) the mass fractions, zero for
    materials not in the cell. Every `pure_every`-th cell keeps only its first
    material; every 5th cell gets a trace material (mass fraction
    `cutoff_trace`) of a pool material it lacks, which a cutoff above that
    excludes from the closure. Returns a dict of numpy arrays plus
    `eos_offsets` (1-based pool indices)."""
    a = uneven(ctx, n, seed=seed)
    off = a["offsets"].astype(np.int64)
    nmat = 6
    cell = np.repeat(np.arange(n), np.diff(off))
    keep = np.ones(len(cell), dtype=bool)
    if pure_every:
        pure = np.zeros(n, dtype=bool)
        pure[::pure_every] = True
        first = np.zeros(len(cell), dtype=bool)
        first[off[:-1]] = True
        keep = ~pure[cell] | first
    c, mats = cell[keep], a["mat_ids"][keep].astype(np.int64)
    rho, vf = a["rho"][keep], a["vfrac"][keep]
    vsum = np.bincount(c, weights=vf, minlength=n)
    al = vf / vsum[c]
    rb = rho * vf
    rbs = np.bincount(c, weights=rb, minlength=n)
    x = rb / rbs[c]
    fm = np.zeros((n, nmat))
    fm[c, mats] = x
    spvol = 1.0 / np.bincount(c, weights=rho * al, minlength=n)
    sie = np.bincount(c, weights=x * a["sie"][keep], minlength=n)
    temp = np.bincount(c, weights=x * a["temp"][keep], minlength=n) / EV2K
    press = np.abs(np.bincount(c, weights=al * a["press"][keep], minlength=n)) + 1e5
    tc = np.arange(0, n, 5)
    for k in range(nmat):  # the trace: the first pool material the cell lacks
        lack = tc[fm[tc, k] == 0]
        has_trace = (fm[lack] == cutoff_trace).any(axis=1)
        fm[lack[~has_trace], k] = cutoff_trace
    return dict(frac_mass=np.ascontiguousarray(fm.T).ravel(), spvol=spvol, sie=sie, temp=temp,
                press=press, eos_offsets=np.arange(1, nmat + 1, dtype=np.int32), ncell=n,
                nmat=nmat)


def table_queries(n, table, seed=SEED_BASE + 1):
    """log10 rho ~ U(-3, 2), log10 T ~ U(2, 6) (true log-uniform) — drawn with
    numpy's PCG64 from the seed (query points only; any seeded draw works)."""
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-3.0, 2.0, n)
    T = 10.0 ** rng.uniform(2.0, 6.0, n)
    return rho, T
