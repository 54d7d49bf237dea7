"""Generate the Rys quadrature roots/weights table used by the HIP ERI kernels.

The Rys rule with n points for parameter T has nodes u_i = t_i^2 in (0,1) and
weights w_i with  sum_i w_i u_i^k = F_k(T) = int_0^1 t^{2k} exp(-T t^2) dt
for k = 0 .. 2n-1 (Boys function moments).  For each n we fit u_i(T) and
w_i(T) piecewise on intervals of width H in [0, TMAX_n) with polynomials of
degree DEG in the local variable x = T - T_mid (Horner form on the device).
Beyond TMAX_n the rule is replaced by its large-T limit: u_i = x_i^2 / T,
w_i = h_i / sqrt(T), with (x_i, h_i) the positive nodes/weights of the 2n-point
Gauss-Hermite rule (the neglected tail is below 1e-16 relative there).

Reference values: 40-digit mpmath.  Moments F_k(T) = gamma(k+1/2, T) /
(2 T^(k+1/2)); recurrence coefficients by Gautschi's Chebyshev algorithm;
nodes and weights as the eigen-decomposition of the Jacobi matrix
(Golub-Welsch).  The output is a flat float64 binary read by the host library:

  header: int32 NMAX, int32 DEG, float64 H, then for n = 1..NMAX:
          int32 nint_n (intervals), float64 TMAX_n,
          float64 coef[nint_n][2n][DEG+1]   (u_0..u_{n-1}, w_0..w_{n-1}),
          float64 hermite_x2[n], float64 hermite_w[n]
"""
from __future__ import annotations

import os
import struct
import sys

import mpmath as mp
import numpy as np

NMAX = 7
DEG = 11
H = 1.0
mp.mp.dps = 40


def boys_moments(T, kmax):
    T = mp.mpf(T)
    if T == 0:
        return [mp.mpf(1) / (2 * k + 1) for k in range(kmax + 1)]
    return [mp.gammainc(k + mp.mpf(0.5), 0, T) / (2 * T ** (k + mp.mpf(0.5))) for k in range(kmax + 1)]


def rys_rule(T, n, mom=None):
    """(nodes u_i ascending, weights) of the n-point Rys rule at T (mpmath)."""
    mom = mom if mom is not None else boys_moments(T, 2 * n - 1)
    # Chebyshev algorithm (Gautschi 2004, Alg. 2.1) on ordinary moments
    alpha = [mp.mpf(0)] * n
    beta = [mp.mpf(0)] * n
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = [mom[l] for l in range(2 * n)]
    alpha[0] = mom[1] / mom[0]
    beta[0] = mom[0]
    for k in range(1, n):
        new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = new[k + 1] / new[k] - sig[k] / sig[k - 1]
        beta[k] = new[k] / sig[k - 1]
        sig_prev, sig = sig, new
    Jm = mp.matrix(n, n)
    for i in range(n):
        Jm[i, i] = alpha[i]
        if i + 1 < n:
            Jm[i, i + 1] = Jm[i + 1, i] = mp.sqrt(beta[i + 1])
    E, Q = mp.eigsy(Jm)
    order = sorted(range(n), key=lambda i: E[i])
    u = [E[i] for i in order]
    w = [beta[0] * Q[0, i] ** 2 for i in order]
    return u, w


def hermite_half(n):
    """Positive nodes squared and weights of the 2n-point Gauss-Hermite rule."""
    x, w = np.polynomial.hermite.hermgauss(2 * n)
    # refine in mpmath via Newton on H_{2n}
    xs, ws = [], []
    for xi in x[x > 0]:
        r = mp.mpf(xi)
        for _ in range(8):
            h2n = mp.hermite(2 * n, r)
            dh = 4 * n * mp.hermite(2 * n - 1, r)
            r = r - h2n / dh
        # w = 2^(N-1) N! sqrt(pi) / (N^2 H_{N-1}(x)^2), N = 2n
        N = 2 * n
        wi = 2 ** (N - 1) * mp.factorial(N) * mp.sqrt(mp.pi) / (N ** 2 * mp.hermite(N - 1, r) ** 2)
        xs.append(r * r)
        ws.append(wi)
    order = sorted(range(n), key=lambda i: xs[i])
    return [xs[i] for i in order], [ws[i] for i in order]


def asymptotic(T, n, hx2, hw):
    T = mp.mpf(T)
    return [x / T for x in hx2], [w / mp.sqrt(T) for w in hw]


def find_tmax(n, hx2, hw, tol=2e-16):
    """Smallest integer T beyond which the large-T rule is within tol (relative)."""
    T = 10
    while True:
        ok = True
        for Tt in (T, T + 0.37, T + 1.5, T + 7.3):
            u, w = rys_rule(Tt, n)
            ua, wa = asymptotic(Tt, n, hx2, hw)
            for a, b in zip(u + w, ua + wa):
                if abs(a - b) > tol * abs(a):
                    ok = False
        if ok:
            return T
        T += 2


def cheb_nodes(a, b, m):
    k = np.arange(m)
    return 0.5 * (a + b) + 0.5 * (b - a) * np.cos((2 * k + 1) * np.pi / (2 * m))


def fit_interval(n, a, b, npts):
    Ts = cheb_nodes(a, b, npts)
    vals = []
    for T in Ts:
        u, w = rys_rule(T, n)
        vals.append([float(x) for x in u] + [float(x) for x in w])
    vals = np.array(vals)  # npts x 2n
    mid = 0.5 * (a + b)
    x = Ts - mid
    V = np.vander(x, DEG + 1, increasing=True)
    coefs, *_ = np.linalg.lstsq(V, vals, rcond=None)  # (DEG+1) x 2n
    return coefs.T  # 2n x (DEG+1)


def generate(path, nmax=NMAX, verbose=True):
    blobs = [struct.pack("<iid", nmax, DEG, H)]
    for n in range(1, nmax + 1):
        hx2, hw = hermite_half(n)
        tmax = find_tmax(n, hx2, hw)
        nint = int(np.ceil(tmax / H))
        tab = np.zeros((nint, 2 * n, DEG + 1))
        for k in range(nint):
            tab[k] = fit_interval(n, k * H, (k + 1) * H, DEG + 9)
        if verbose:
            print(f"n={n}: TMAX={tmax} intervals={nint}", flush=True)
        blobs.append(struct.pack("<id", nint, float(nint * H)))
        blobs.append(tab.astype("<f8").tobytes())
        blobs.append(np.array([float(x) for x in hx2], dtype="<f8").tobytes())
        blobs.append(np.array([float(x) for x in hw], dtype="<f8").tobytes())
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        for b in blobs:
            f.write(b)
    os.replace(tmp, path)


def load(path):
    """Parse the binary table (for host-side checks)."""
    raw = open(path, "rb").read()
    nmax, deg, h = struct.unpack_from("<iid", raw, 0)
    off = 16
    out = {}
    for n in range(1, nmax + 1):
        nint, tmax = struct.unpack_from("<id", raw, off)
        off += 12
        cnt = nint * 2 * n * (deg + 1)
        tab = np.frombuffer(raw, dtype="<f8", count=cnt, offset=off).reshape(nint, 2 * n, deg + 1)
        off += cnt * 8
        hx2 = np.frombuffer(raw, dtype="<f8", count=n, offset=off)
        off += n * 8
        hw = np.frombuffer(raw, dtype="<f8", count=n, offset=off)
        off += n * 8
        out[n] = dict(nint=nint, tmax=tmax, tab=tab, hx2=hx2, hw=hw)
    return dict(nmax=nmax, deg=deg, h=h, rules=out)


def evaluate(tables, n, T):
    r = tables["rules"][n]
    if T >= r["tmax"]:
        return r["hx2"] / T, r["hw"] / np.sqrt(T)
    k = int(T / tables["h"])
    x = T - (k + 0.5) * tables["h"]
    c = r["tab"][k]
    v = np.zeros(2 * n)
    for d in range(tables["deg"], -1, -1):
        v = v * x + c[:, d]
    return v[:n], v[n:]


if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "..", "qdk_chemistry_amd", "lib", "rys_table.bin")
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    generate(out)
