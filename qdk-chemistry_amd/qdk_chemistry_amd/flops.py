"""Algorithmic FP64 operation counts of the ERI class kernels.

The count follows the source of csrc/eri_core.hpp / csrc/eri_class.hpp
operation by operation (FMA = 2, add/mul/rcp/rsq = 1) and is the figure the
roofline uses:

    flops(class) = N_prim_quartets * f_prim(class) + N_quartets * f_contr(class)

f_prim: one pass of the inner (bra primitive) loop body: Gaussian-product
setup, roots and weights on their cheapest path (large-T limit for 3+ roots,
large-T Boys moments for 1-2 roots), per root the 2-D recursion and the 1-D
transfers in x, y, z, and the assembly of the chunk's integrals -- counted
once per bra-function chunk, since every chunk kernel repeats the loop.
Operations the compiler may remove (terms a chunk never uses) are counted only
where the source skips them (zero solid-harmonic coefficients).
f_contr: ket solid-harmonic transform and J/K digestion per shell quartet.
The ket-primitive work of the outer loop is not counted (it is amortised over
the bra primitives), so the figure is a lower bound of what executes.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from .basis import cart_list, solid_harmonic_matrix

ACC_MAX = 48  # must match kAccMax in csrc/eri_class.hpp


def ncart(l):
    return (l + 1) * (l + 2) // 2


def nsph(l, pure=True):
    return 2 * l + 1 if (pure and l >= 2) else ncart(l)


def _sph(l, pure=True):
    if pure and l >= 2:
        return solid_harmonic_matrix(l)
    return np.eye(ncart(l))


def chunks(la, lb, lc, ld, pure=True):
    nbra = nsph(la, pure) * nsph(lb, pure)
    nketc = ncart(lc) * ncart(ld)
    ch = max(1, min(ACC_MAX // nketc, nbra))
    return [(q0, min(q0 + ch, nbra)) for q0 in range(0, nbra, ch)]


def _onedim(la, lb, lc, ld):
    """flops of rys_1d for one direction."""
    EA, FC = la + lb, lc + ld
    f = 0
    if EA >= 1:
        f += 1                      # g[1][0] = C00 * s
    f += max(EA - 1, 0) * 4         # g[i+1][0] = C00 g + i*B10*g  (mul, mul, fma)
    if FC >= 1:
        f += 1                      # g[0][1]
        f += EA * 4                 # g[i][1]
        f += max(FC - 1, 0) * (4 + EA * 7)  # g[0][k+1]; g[i][k+1] (3 mul + 2 fma)
    # ket transfer (per i): t[k][d] = t[k+1][d-1] + CD t[k][d-1]  (fma)
    f += (EA + 1) * sum((FC + 1 - d) * 2 for d in range(1, ld + 1))
    # bra transfer (per c, d)
    f += (lc + 1) * (ld + 1) * sum((EA + 1 - b) * 2 for b in range(1, lb + 1))
    return f


def _roots(nr):
    if nr == 1:
        return 10 + 2 + 10          # large-T Boys F0,F1 (frcp, frsq, mul) + u = F1*rcp(F0)
    if nr == 2:
        return 10 + 10 + 6 + 30 + 10 + 14  # Boys F0..F3 (large T) + 2x2 moment solve
    return 20 + 2 * nr               # large-T Hermite limit: frcp, frsq, 2n mul


@lru_cache(maxsize=None)
def class_flops(la, lb, lc, ld, pure=True):
    """(f_prim, f_contr) for class (la lb | lc ld) as evaluated."""
    L = la + lb + lc + ld
    nr = L // 2 + 1
    Sa, Sb, Sc, Sd = _sph(la, pure), _sph(lb, pure), _sph(lc, pure), _sph(ld, pure)
    nketc = ncart(lc) * ncart(ld)
    nkets = nsph(lc, pure) * nsph(ld, pure)
    nsb = nsph(lb, pure)
    setup = 3 + 3 + 1 + 10 + 1 + 2 + 5 + 1 + 3 + 2 + 20   # PA, PQ, p+q, rsq, rpq, rho, T, pref, qr/pr, 1/2p, 1/2q
    per_root_common = 2 + 3 + 3 + 2 + 12          # B00, B10, B01, xa/xc, C00/C00' (x, y, z)
    onedim = 3 * _onedim(la, lb, lc, ld)
    f_prim = 0
    for q0, q1 in chunks(la, lb, lc, ld, pure):
        asm = 0
        for q in range(q0, q1):
            ma, mb = divmod(q, nsb)
            nnz = int(np.count_nonzero(Sa[ma])) * int(np.count_nonzero(Sb[mb]))
            ones = sum(1 for a in np.nonzero(Sa[ma])[0] for b in np.nonzero(Sb[mb])[0]
                       if abs(Sa[ma, a] * Sb[mb, b] - 1.0) < 1e-15)
            # per ket cartesian: v = Ix*Iy*Iz (2 mul), s += v (1) or fma (2), acc += s (1)
            asm += nketc * (nnz * 2 + ones * 1 + (nnz - ones) * 2 + 1)
        f_prim += setup + _roots(nr) + nr * (per_root_common + onedim + asm)
    # contracted: ket transform + digestion (6 fma per element) + deg scaling
    nnz_ket = sum(int(np.count_nonzero(Sc[mc])) * int(np.count_nonzero(Sd[md]))
                  for mc in range(Sc.shape[0]) for md in range(Sd.shape[0]))
    nbra = nsph(la, pure) * nsph(lb, pure)
    f_contr = nbra * (2 * nnz_ket + nkets * (1 + 12)) + 20
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
