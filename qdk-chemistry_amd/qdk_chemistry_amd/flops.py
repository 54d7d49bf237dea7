"""Algorithmic operation counts of the ERI class kernels (the roofline's F).

    flops(class) = N_prim_quartets * f_prim(class) + N_quartets * f_contr(class)

The counts follow csrc/eri_core.hpp (contract_chunk, rys_roots, rys_1d_x,
ket_to_sph) and csrc/eri_class.hpp (k_eri_jk_seg's digestion) operation by
operation: FMA = 2, add / mul / hardware rcp / rsq = 1.  The chunk geometry
-- the bra functions [Q0, Q1) and ket cartesians [K0, K1) that each kernel
of a class evaluates -- is read from the compiled library (qfe_class_chunks,
i.e. JKChunks of eri_class.hpp itself, including its LDS-row rule), so the
model follows the kernels that run.  Every chunk kernel repeats the primitive
loop, so f_prim sums over the chunks.

f_prim, per primitive quartet of one chunk:
  setup     PA, PQ (6), p + q (1), rsqrt with two Newton steps (10), rpq (1),
            rho (2), T (6), prefactor (3), q/(p+q) and p/(p+q) (2) = 31.
            1/(2p) and 1/(2q) are table loads (no flops).
  roots     the cheapest branch of rys_roots (large-T Boys moments for one or
            two roots, the large-T Hermite limit for three or more), so the
            count is a lower bound of what executes.  (ps|ss) has no root
            and no per-root terms: see _ps_ss.
  per root  B00, B10, B01, the two root scalings (10), C00 and C00' in x, y,
            z (12), the weight on Iz (1) = 23; the 2-D recursion and the two
            1-D transfers in each direction (rys_1d_x); the chunk's assembly,
            per (bra function, ket cartesian): 2 mul per product of the three
            1-D integrals, 1 add (unit coefficient) or 1 FMA per product, and
            1 add into the accumulator.
f_contr, per shell quartet and chunk: the ket solid-harmonic transform
(1 FMA per nonzero coefficient and ket function) and the digestion (the deg
scaling and 6 FMAs per bra function x ket function), plus 20 for the task
set-up.  The ket-primitive work of the outer loop (amortised over the bra
primitives) is not counted.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from .basis import solid_harmonic_matrix


def ncart(l):
    return (l + 1) * (l + 2) // 2


def nsph(l, pure=True):
    return 2 * l + 1 if (pure and l >= 2) else ncart(l)


def _sph(l, pure=True):
    if pure and l >= 2:
        return solid_harmonic_matrix(l)
    return np.eye(ncart(l))


def chunks(la, lb, lc, ld):
    """[(Q0, Q1, K0, K1)] of the compiled J/K kernels of the class."""
    import ctypes
    from . import _lib
    cap = 4096
    out = np.zeros(4 * cap, dtype=np.int32)
    n = _lib.lib().qfe_class_chunks(la, lb, lc, ld, cap, _lib.ptr(out, ctypes.c_int))
    if n < 0:
        raise ValueError(f"class ({la}{lb}|{lc}{ld}) is not compiled into the library")
    if n > cap:
        raise ValueError("more chunks than the query holds")
    return [tuple(int(x) for x in out[4 * i:4 * i + 4]) for i in range(n)]


def _onedim(la, lb, lc, ld):
    """flops of rys_1d_x for one direction (EA = la + lb, FC = lc + ld)."""
    EA, FC = la + lb, lc + ld
    f = 0
    if EA >= 1:
        f += 1                      # g[1][0] = C00 * s
    f += max(EA - 1, 0) * 4         # g[i+1][0] = C00 g + i B10 g   (mul, mul, fma)
    if FC >= 1:
        f += 1                      # g[0][1]
        f += EA * 4                 # g[i][1]
        f += max(FC - 1, 0) * (4 + EA * 7)  # g[0][k+1]; g[i][k+1] (3 mul + 2 fma)
    # ket transfer (per i): t[k][d] = t[k+1][d-1] + CD t[k][d-1]  (fma)
    f += (EA + 1) * sum((FC + 1 - d) * 2 for d in range(1, ld + 1))
    # bra transfer (per c, d)
    f += (lc + 1) * (ld + 1) * sum((EA + 1 - b) * 2 for b in range(1, lb + 1))
    return f


def _roots(nr, L=1):
    if nr == 1:
        if L == 0:
            return 9                # F0 only (rsqrt with Newton steps, mul): no root is needed
        return 10 + 2 + 10          # large-T F0, F1 (rcp + rsqrt with Newton steps, mul) + u = F1 / F0
    if nr == 2:
        # F0..F3 (large T: rsqrt with Newton steps, 1/T as its square, F0
        # and the upward steps) + the 2 x 2 moment solve
        return 10 + 1 + 6 + 30 + 10 + 14
    return 20 + 2 * nr               # large-T Hermite limit: rcp, rsqrt, 2n mul


def _ps_ss(nq):
    """f_prim of one (ps|ss) chunk of nq bra functions.  The kernel skips the
    root: with one root the integral is linear in it, and w = F_0, w u = F_1
    give (p_i s|ss) = pref (F_0 PA_i - q/(p+q) F_1 PQ_i).  Set-up (30), the
    large-T F_0, F_1 (rsqrt with Newton steps 10, 1/T as its square 1,
    F_0 1, F_1 2), pref F_0 and pref F_1 q/(p+q) (3), and per bra function
    1 mul, 1 FMA and the accumulator add (4)."""
    return _setup(1, 0) + 14 + 3 + 4 * nq


SETUP = 3 + 3 + 1 + 10 + 1 + 2 + 6 + 3 + 2   # 31 with every term used
PER_ROOT = 10 + 12 + 1                      # 23 with every term used


def _setup(EA, FC):
    """The primitive set-up terms the class uses (the compiler drops the rest):
    PA (3) with a bra shell above s, PQ (3), p + q (1), rsqrt with two Newton
    steps (10), rpq (1), rho (2), T (6), prefactor (3), q/(p+q) and p/(p+q)
    (1 each) where a root scaling uses them."""
    return (3 if EA else 0) + 3 + 1 + 10 + 1 + 2 + 6 + 3 + (1 if EA else 0) + (1 if FC else 0)


def _per_root(EA, FC):
    """Per root: B00 = rpq u / 2 (2) for bra and ket above s; B10 (3) from EA = 2,
    B01 (3) from FC = 2; the root scalings q/(p+q) u (1) and p/(p+q) u (1);
    C00 (3 FMA) and C00' (3 FMA) where needed; the weight on Iz (1)."""
    return ((2 if EA and FC else 0) + (3 if EA >= 2 else 0) + (3 if FC >= 2 else 0) +
            (1 + 6 if EA else 0) + (1 + 6 if FC else 0) + 1)


@lru_cache(maxsize=None)
def class_flops(la, lb, lc, ld, pure=True):
    """(f_prim, f_contr) for class (la lb | lc ld) as compiled."""
    L = la + lb + lc + ld
    nr = L // 2 + 1
    Sa, Sb, Sc, Sd = _sph(la, pure), _sph(lb, pure), _sph(lc, pure), _sph(ld, pure)
    nsb, nsc, nsd = nsph(lb, pure), nsph(lc, pure), nsph(ld, pure)
    ncd = ncart(ld)
    onedim = 3 * _onedim(la, lb, lc, ld)
    f_prim = f_contr = 0
    for q0, q1, k0, k1 in chunks(la, lb, lc, ld):
        nk = k1 - k0
        asm = 0
        for q in range(q0, q1):
            ma, mb = divmod(q, nsb)
            nnz = int(np.count_nonzero(Sa[ma])) * int(np.count_nonzero(Sb[mb]))
            ones = sum(1 for a in np.nonzero(Sa[ma])[0] for b in np.nonzero(Sb[mb])[0]
                       if abs(Sa[ma, a] * Sb[mb, b] - 1.0) < 1e-15)
            asm += nk * (nnz * 2 + ones * 1 + (nnz - ones) * 2 + 1)
        EA, FC = la + lb, lc + ld
        if L == 1 and la == 1:
            f_prim += _ps_ss(q1 - q0)
        else:
            f_prim += _setup(EA, FC) + _roots(nr, L) + nr * (_per_root(EA, FC) + onedim + asm)
        # ket transform of this chunk's cartesians into every ket function
        nnz_ket = sum(1 for mc in range(nsc) for md in range(nsd) for k in range(k0, k1)
                      if Sc[mc, k // ncd] * Sd[md, k % ncd] != 0.0)
        f_contr += (q1 - q0) * (2 * nnz_ket + nsc * nsd * (1 + 12)) + 20
    return f_prim, f_contr


def build_flops(class_rows):
    """class_rows: iterable of ((la, lb, lc, ld), quartets, prim_quartets)."""
    tot = 0.0
    for cls, nq, npq in class_rows:
        fp, fc = class_flops(*cls)
        tot += npq * fp + nq * fc
    return tot


def table(maxl=2, pure=True):
    rows = []
    pcs = [(a, b) for a in range(maxl + 1) for b in range(a + 1)]
    for i, (la, lb) in enumerate(pcs):
        for lc, ld in pcs[: i + 1]:
            fp, fc = class_flops(la, lb, lc, ld, pure)
            rows.append(((la, lb, lc, ld), fp, fc))
    return rows


def write_table(path, maxl=3):
    """The committed per-class table (markdown) that the bench and the roofline
    use: class, chunks, f_prim, f_contr."""
    lines = ["| class | chunks | f_prim | f_contr |", "|---|---|---|---|"]
    for (la, lb, lc, ld), fp, fc in table(maxl):
        lines.append(f"| ({la}{lb}\\|{lc}{ld}) | {len(chunks(la, lb, lc, ld))} | {fp} | {fc} |")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
