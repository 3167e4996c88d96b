"""Writes kh_j0_tables.hpp: the coefficients of the device J0 (kh_j0.hpp).

Run from any directory:  python gen_j0_tables.py   (needs mpmath; 50-digit
arithmetic, a few seconds).  The output is committed, so building the
library does not need mpmath.

Two regimes, split at x = J0_XA:

* x >= J0_XA: Hankel's large-argument form
      J0(x) = sqrt(2 / (pi x)) * (P(x) cos(x - pi/4) - Q(x) sin(x - pi/4)),
  with P and Q obtained from mpmath's J0 and Y0 through
      P = sqrt(pi x / 2) * ( J0 cos chi + Y0 sin chi),
      Q = sqrt(pi x / 2) * (-J0 sin chi + Y0 cos chi),   chi = x - pi/4,
  and fitted by least squares on Chebyshev nodes as polynomials in
  z = 1/x^2:  P ~ sum_k JP[k] z^k,  Q ~ (1/x) sum_k JQ[k] z^k.
  (Asymptotically P -> 1, Q -> -1/(8x).)

* 0 <= x < J0_XA: one polynomial per unit interval [n, n+1) in
  t = x - (n + 1/2), fitted the same way to mpmath's J0.

The script checks every fit against mpmath on a fine grid, in the units the
kernel needs (absolute error in J0 / sqrt(2/(pi x)) for the large regime,
absolute error in J0 for the small one), and refuses to write a table whose
error exceeds 2e-17 (large) or 1e-16 (small: the coefficients themselves are
rounded to doubles of magnitude up to 1, half an ulp of which is 5.6e-17).
"""

from __future__ import annotations

import os

import mpmath as mp

mp.mp.dps = 50

XA = 12          # regime split
DEG_PQ = 9       # degree in z of the P and Q polynomials
DEG_SMALL = 14   # degree in t of the per-interval polynomials
TOL_LARGE = 2e-17
TOL_SMALL = 1e-16


def _cheb_nodes(a, b, n):
    return [a + (b - a) * (1 + mp.cos(mp.pi * (i + mp.mpf(1) / 2) / n)) / 2 for i in range(n)]


def _lsq(nodes, vals, deg, scale, shift=0):
    """Monomial coefficients c_k of sum c_k (s - shift)^k, fitted in the
    scaled variable (s - shift) / scale for conditioning."""
    A = mp.matrix([[((s - shift) / scale) ** k for k in range(deg + 1)] for s in nodes])
    c = mp.qr_solve(A, mp.matrix(vals))[0]
    return [c[k] / scale ** k for k in range(deg + 1)]


def _pq(x):
    chi = x - mp.pi / 4
    J, Y = mp.besselj(0, x), mp.bessely(0, x)
    s = mp.sqrt(mp.pi * x / 2)
    return s * (J * mp.cos(chi) + Y * mp.sin(chi)), s * (-J * mp.sin(chi) + Y * mp.cos(chi))


def _horner(c, v):
    acc = mp.mpf(0)
    for ck in reversed(c):
        acc = acc * v + ck
    return acc


def fit_large():
    zmax = mp.mpf(1) / XA ** 2
    nodes = _cheb_nodes(mp.mpf(0) + zmax * mp.mpf("1e-6"), zmax, DEG_PQ + 30)
    P, Q = [], []
    for z in nodes:
        x = 1 / mp.sqrt(z)
        p, q = _pq(x)
        P.append(p)
        Q.append(q * x)
    cP = _lsq(nodes, P, DEG_PQ, zmax)
    cQ = _lsq(nodes, Q, DEG_PQ, zmax)
    cPd = [float(c) for c in cP]
    cQd = [float(c) for c in cQ]
    err = 0.0
    for i in range(400):
        x = mp.mpf(XA) * (1 + mp.mpf(i) ** 2 / 400)
        z = 1 / x ** 2
        p, q = _pq(x)
        ep = abs(_horner(cPd, z) - p)
        eq = abs(_horner(cQd, z) / x - q)
        err = max(err, float(ep + eq))
    return cPd, cQd, err


def fit_small():
    rows, err = [], 0.0
    for n in range(XA):
        c = mp.mpf(n) + mp.mpf(1) / 2
        nodes = _cheb_nodes(mp.mpf(n), mp.mpf(n + 1), DEG_SMALL + 30)
        vals = [mp.besselj(0, s) for s in nodes]
        co = [float(v) for v in _lsq(nodes, vals, DEG_SMALL, mp.mpf(1) / 2, shift=c)]
        for i in range(201):
            s = mp.mpf(n) + mp.mpf(i) / 200
            err = max(err, float(abs(_horner(co, s - c) - mp.besselj(0, s))))
        rows.append(co)
    return rows, err


def main():
    cP, cQ, e_large = fit_large()
    rows, e_small = fit_small()
    print(f"large-argument fit: max error {e_large:.3g}; small-argument fit: {e_small:.3g}")
    if e_large > TOL_LARGE or e_small > TOL_SMALL:
        raise SystemExit("fit error above tolerance; table not written")
    here = os.path.dirname(os.path.abspath(__file__))
    out = [
        "// Generated by gen_j0_tables.py (mpmath, 50 digits); do not edit.",
        f"// Fit errors against mpmath: large regime {e_large:.2e} (in units of",
        f"// sqrt(2/(pi x))), small regime {e_small:.2e} (absolute).",
        "#pragma once",
        "",
        "namespace kh {",
        "",
        f"constexpr int J0_XA = {XA};",
        f"constexpr int J0_DEG_PQ = {DEG_PQ};",
        f"constexpr int J0_DEG_SMALL = {DEG_SMALL};",
        "",
        "// P(z) and x Q(z), z = 1/x^2, x >= J0_XA; index = power of z",
        "constexpr double J0_P[J0_DEG_PQ + 1] = {",
        *[f"    {c!r}," for c in cP],
        "};",
        "constexpr double J0_Q[J0_DEG_PQ + 1] = {",
        *[f"    {c!r}," for c in cQ],
        "};",
        "",
        "// J0 on [n, n+1) as a polynomial in t = x - (n + 1/2); row n, index = power of t",
        "constexpr double J0_SMALL[J0_XA][J0_DEG_SMALL + 1] = {",
    ]
    for co in rows:
        out.append("    {" + ", ".join(repr(c) for c in co) + "},")
    out += ["};", "", "}  // namespace kh", ""]
    with open(os.path.join(here, "kh_j0_tables.hpp"), "w", encoding="utf-8") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
