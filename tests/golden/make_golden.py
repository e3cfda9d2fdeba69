"""Writes the small golden vectors under tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

The reference (JAX 0.4.33) cannot be imported in this environment, so the
expected outputs come from the float64 oracle, and every case is accepted only
if an independent brute-force form (direct O(N^4) Rayleigh-Sommerfeld sum,
dense chirp-z DFT, scipy.signal.czt) agrees to 1e-11 relative L2.  Inputs are
seeded (numpy default_rng) and stored next to the outputs.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import brute as B  # noqa: E402
from oracle import xlumina_ref as R  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cfield(rng, n, m=None):
    m = n if m is None else m
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


def main():
    cases = {}
    rng = np.random.default_rng(20240601)
    # scalar RS: even / odd N, z > 0 and z < 0
    for n, half, z in ((16, 60.0, 300.0), (24, 80.0, -250.0), (17, 50.0, 500.0)):
        x, y = R.space(half, n)
        u = cfield(rng, n)
        out, qf = R.rs_propagation(u, x, y, 0.65, z)
        d = B.rs_direct(u, x, y, 0.65, z)
        assert rel(out, d) < 1e-11, rel(out, d)
        cases[f"rs_n{n}"] = dict(x=x, y=y, wavelength=0.65, z=z, field=u, out=out, qf=qf)
    # vectorial RS with per-pixel LCD before it
    n = 20
    x, y = R.space(90.0, n)
    Ex, Ey = cfield(rng, n), cfield(rng, n)
    eta = rng.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    th = rng.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    mx, my, _ = R.LCD(Ex, Ey, None, eta.astype(np.float64), th.astype(np.float64))
    (px, py, pz), qf = R.vrs_propagation(mx, my, x, y, 0.635, 700.0)
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2 + 700.0 ** 2)
    for comp, src in ((px, mx), (py, my), (pz, mx * X / r + my * Y / r)):
        assert rel(comp, B.rs_direct(src, x, y, 0.635, 700.0)) < 1e-11
    cases["vrs_lcd_n20"] = dict(x=x, y=y, wavelength=0.635, z=700.0, Ex=Ex, Ey=Ey, eta=eta, theta=th,
                                out=np.stack([px, py, pz]), qf=qf)
    # scalar CZT: M < m, M == m (missing-tap quirk), M > m, rectangular output
    for n, mxo, myo in ((24, 10, 12), (20, 20, 20), (16, 21, 18)):
        x, y = R.space(100.0, n)
        xo = np.linspace(-20.0, 20.0, mxo)
        yo = np.linspace(-15.0, 12.0, myo)
        u = cfield(rng, n)
        out = R.czt(u, x, y, 0.65, 1000.0, xo, yo)
        d = B.czt2d_direct(u, x, y, 0.65, 1000.0, xo, yo)
        assert rel(out, d) < 1e-11, rel(out, d)
        cases[f"czt_n{n}_{mxo}x{myo}"] = dict(x=x, y=y, wavelength=0.65, z=1000.0, xout=xo, yout=yo, field=u,
                                              out=out)
    # VCZT
    n = 24
    x, y = R.space(100.0, n)
    xo = np.linspace(-20.0, 20.0, 14)
    Ex, Ey = cfield(rng, n), cfield(rng, n)
    out = np.stack(R.vczt(Ex, Ey, x, y, 0.65, 1200.0, xo, xo))
    X, Y = np.meshgrid(x, y)
    r = np.sqrt(X ** 2 + Y ** 2 + 1200.0 ** 2)
    ez = (Ex * X / r + Ey * Y / r) * 1200.0 / r
    for i, src in enumerate((Ex, Ey, ez)):
        assert rel(out[i], B.czt2d_direct(src, x, y, 0.65, 1200.0, xo, xo)) < 1e-11
    cases["vczt_n24"] = dict(x=x, y=y, wavelength=0.65, z=1200.0, xout=xo, yout=xo, Ex=Ex, Ey=Ey, out=out)
    # high-NA objective lens (even N: no rho = 0 sample)
    n = 32
    x, y = R.space(1800.0, n)
    rad, f = 1800.0, 2000.0
    xo = np.linspace(-2.0, 2.0, 16)
    Ex, Ey = cfield(rng, n), cfield(rng, n)
    out = R.vczt_objective_lens(Ex, Ey, x, y, 0.65, rad, f, xo, xo)
    lens, stm = R.high_NA_objective_lens(Ex, Ey, x, y, rad, f)
    Dm = f * 0.65 * (n - 1) / (2 * rad)
    K = B.czt_matrix(n, xo[0] + Dm / 2, xo[-1] + Dm / 2, Dm, len(xo))
    cte = -(1j * stm ** 2 / (f * 0.65))
    for i in range(3):
        assert rel(out[i], cte * (K @ lens[i] @ K.T)) < 1e-11
    cases["hna_n32"] = dict(x=x, y=y, wavelength=0.65, radius=rad, f=f, xout=xo, yout=xo, Ex=Ex, Ey=Ey, out=out,
                            lens=lens)
    path = os.path.join(HERE, "golden_small.npz")
    flat = {}
    for name, d in cases.items():
        for k, v in d.items():
            flat[f"{name}__{k}"] = np.asarray(v)
    np.savez_compressed(path, **flat)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
