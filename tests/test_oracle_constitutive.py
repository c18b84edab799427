"""Pins of the oracle's Mualem-van Genuchten evaluation (src/global.cu:32-105).

The reference value is the closed form evaluated with 50-digit decimal arithmetic
(Python's decimal module: ln/exp/sqrt are correctly rounded there), written out from the
textbook definitions (van Genuchten 1980, eqs. 3, 8, 21), not from the oracle's code:
    Se = [1 + (alpha |psi|)^n]^(-m),  m = 1 - 1/n
    theta = theta_r + (theta_s - theta_r) Se
    K = Ks Se^(1/2) [1 - (1 - Se^(1/m))^m]^2
    C = dtheta/dpsi  (checked against a finite difference of theta)
"""
from decimal import Decimal, getcontext

import numpy as np
import pytest

from dhara_amd import _abi
from oracle import oracle

getcontext().prec = 50

LOAM = dict(ts=0.43, tr=0.078, alpha=3.6, n=1.56, Ks=2.89e-6)


def D(x):
    return Decimal(repr(float(x)))


def vg_decimal(psi, ts, tr, alpha, n, Ks):
    psi, ts, tr, alpha, n, Ks = map(D, (psi, ts, tr, alpha, n, Ks))
    if psi >= 0:
        return float(ts), float(Ks)
    m = 1 - 1 / n
    x = alpha * (-psi)
    Se = (1 + (x.ln() * n).exp()).ln() * (-m)
    Se = Se.exp()
    theta = tr + (ts - tr) * Se
    inner = 1 - (Se.ln() / m).exp()           # 1 - Se^(1/m)
    w = 1 - (inner.ln() * m).exp()            # 1 - inner^m
    K = Ks * Se.sqrt() * w * w
    return float(theta), float(K)


def model(n=1, **kw):
    g = _abi.Grid(nx=n, ny=1, nz=1, nroot_groups=1, dx=1.0, dy=1.0, dz=0.05)
    p = _abi.default_params(theta_s=LOAM["ts"], theta_r=LOAM["tr"], alpha=LOAM["alpha"],
                            n_vg=LOAM["n"], porosity=LOAM["ts"], **kw)
    m = oracle.OracleModel(g, p)
    m.upload("KSAT", np.full(n, LOAM["Ks"]))
    return m


PSIS = np.array([-1e-7, -1e-4, -0.01, -0.1, -0.5, -1.0, -3.0, -10.0, -100.0])


def test_loam_table_against_decimal_closed_form():
    m = model(len(PSIS))
    th, K, C = m.eval_constitutive(PSIS)
    for i, psi in enumerate(PSIS):
        tref, kref = vg_decimal(psi, **{k: LOAM[k] for k in ("ts", "tr", "alpha", "n", "Ks")})
        assert abs(th[i] - tref) <= 1e-15, psi
        assert abs(K[i] - kref) <= 2e-13 * kref, psi


def test_reference_pow_form_agrees_up_to_its_cancellation():
    m = model(len(PSIS))
    th, K, C = m.eval_constitutive(PSIS)
    th2, K2, C2 = m.eval_constitutive_refform(PSIS)
    assert np.max(np.abs(th - th2)) < 1e-15
    assert np.max(np.abs(C - C2) / np.abs(C)) < 1e-12
    # the reference's 1 - pow(Se, 1/m) cancels near saturation: relative error ~ eps / (1-Se^(1/m))
    far = PSIS <= -0.1
    assert np.max(np.abs(K[far] - K2[far]) / K[far]) < 1e-10
    assert np.max(np.abs(K - K2) / K) < 1e-3


def test_capacity_is_derivative_of_theta():
    psi = -np.geomspace(1e-3, 50.0, 25)
    m = model(len(psi))
    _, _, C = m.eval_constitutive(psi)
    h = 1e-6 * np.abs(psi)
    tp, _, _ = m.eval_constitutive(psi + h)
    tm, _, _ = m.eval_constitutive(psi - h)
    fd = (tp - tm) / (2 * h)
    assert np.max(np.abs(C - fd) / np.abs(C)) < 1e-6


def test_saturated_limits_and_monotonicity():
    psi = np.array([0.0, 1e-3, 2.0, -0.0])
    m = model(4)
    th, K, C = m.eval_constitutive(psi)
    assert np.all(th == LOAM["ts"]) and np.all(K == LOAM["Ks"]) and np.all(C == 0.0)
    psi = -np.geomspace(1e-6, 1e3, 200)[::-1]     # increasing psi
    m = model(len(psi))
    th, K, C = m.eval_constitutive(psi)
    assert np.all(np.diff(th) > 0) and np.all(np.diff(K) > 0) and np.all(C > 0)
    Se = (th - LOAM["tr"]) / (LOAM["ts"] - LOAM["tr"])
    assert np.all((Se > 0) & (Se <= 1))


def test_inverse_round_trip():
    theta = np.linspace(0.0785, 0.4299, 50)
    m = model(50)
    psi = m.eval_inverse(theta)
    th2, _, _ = m.eval_constitutive(psi)
    assert np.max(np.abs(th2 - theta)) < 1e-14
    assert m.eval_inverse(np.array([0.43] * 50))[0] == 0.0


def test_per_cell_alpha_n_fields():
    rng = np.random.default_rng(3)
    n = 30
    a = rng.uniform(2, 5, n)
    nv = rng.uniform(1.3, 1.8, n)
    psi = -rng.uniform(0.01, 5.0, n)
    m = model(n)
    m.upload("ALPHA", a)
    m.upload("NVG", nv)
    th, K, C = m.eval_constitutive(psi)
    for i in range(n):
        tref, kref = vg_decimal(psi[i], LOAM["ts"], LOAM["tr"], a[i], nv[i], LOAM["Ks"])
        assert abs(th[i] - tref) <= 1e-15
        assert abs(K[i] - kref) <= 2e-13 * kref
