"""Pins of the oracle's LW pieces: Gauss quadrature, Planck source, the
no-scattering solver (closed forms in limits), optimal angles, heating rate."""
import copy

import numpy as np
import pytest


def test_gauss_nodes_match_scipy_and_closed_form():
    from oracle import oracle as O
    from scipy.special import roots_jacobi
    D, w = O.gauss(2)
    mu = np.array([(6 + np.sqrt(6)) / 10, (6 - np.sqrt(6)) / 10])
    assert np.allclose(D, 1 / mu, rtol=1e-13)
    assert np.allclose(w, [(9 + np.sqrt(6)) / 36, (9 - np.sqrt(6)) / 36], rtol=1e-13)
    for n in (2, 3, 4):
        x, wx = roots_jacobi(n, 0, 1)
        m = (x + 1) / 2
        order = np.argsort(-m)
        D, w = O.gauss(n)
        assert np.allclose(D, 1 / m[order], rtol=1e-12)
        assert np.allclose(w, wx[order] / 4, rtol=1e-12)
        assert w.sum() == pytest.approx(0.5, rel=1e-14)
    D, w = O.gauss(1)
    assert D[0] == 1.66 and w[0] == 0.5


def _isothermal(kd, T, nlay, ncol=2):
    from ccpprad import synthetic as syn
    pr = syn.make_profiles(ncol, nlay, seed=3, clouds=False, kd=kd)
    pr.t_lay = np.full_like(pr.t_lay, T)
    pr.t_lev = np.full_like(pr.t_lev, T)
    pr.tsfc = np.full_like(pr.tsfc, T)
    return pr


def test_planck_source_band_sums_and_jacobian(kd_lw):
    """sum over a band of lay_src = B_b(T) (Planck fractions sum to 1), and
    sfc_jac equals the table slope over 1 K."""
    from oracle import oracle as O
    from ccpprad import synthetic as syn
    pr = syn.make_profiles(3, 20, seed=4, clouds=False, kd=kd_lw)
    _, pf = O.gas_optics(kd_lw, pr)
    src = O.lw_sources(kd_lw, pr, pf)
    nb = kd_lw.nband
    T = kd_lw.temp_ref
    tpl = np.linspace(T[0], T[-1], kd_lw.totplnk.shape[1])
    for c in range(3):
        for l in (0, 7, 19):
            Bb = np.array([np.interp(pr.t_lay[l, c], tpl, kd_lw.totplnk[b]) for b in range(nb)])
            s = src["lay_src"][c, l].reshape(nb, 8).sum(axis=1)
            assert np.allclose(s, Bb, rtol=1e-12)
        B0 = np.array([np.interp(pr.tsfc[c], tpl, kd_lw.totplnk[b]) for b in range(nb)])
        B1 = np.array([np.interp(pr.tsfc[c] + 1.0, tpl, kd_lw.totplnk[b]) for b in range(nb)])
        assert np.allclose(src["sfc_jac"][c].reshape(nb, 8).sum(axis=1), B1 - B0, rtol=1e-9)


def test_isothermal_blackbody_column(kd_lw):
    """Isothermal atmosphere over a black surface at the same temperature,
    Planck fractions equal in every layer: F_up = pi * sum_b B_b(T) at every
    level, and with one angle (D = 1.66, w = 1/2) the surface downward flux
    is pi * sum_g S_g (1 - exp(-1.66 tau_col,g))."""
    from oracle import oracle as O
    T, nlay = 260.0, 30
    pr = _isothermal(kd_lw, T, nlay)
    tau, pf = O.gas_optics(kd_lw, pr)
    pf = np.ascontiguousarray(np.repeat(pf[:, :1, :], nlay, axis=1))
    src = O.lw_sources(kd_lw, pr, pf)
    emis = np.ones((pr.ncol, kd_lw.nband))
    up, dn, _ = O.lw_solve(kd_lw, pr.ncol, nlay, tau, src, emis, nang=1)
    tpl = np.linspace(kd_lw.temp_ref[0], kd_lw.temp_ref[-1], kd_lw.totplnk.shape[1])
    Btot = sum(np.interp(T, tpl, kd_lw.totplnk[b]) for b in range(kd_lw.nband))
    assert np.allclose(up, np.pi * Btot, rtol=1e-10)
    S = src["lay_src"][:, 0, :]
    expect = np.pi * (S * (1.0 - np.exp(-1.66 * tau.sum(axis=1)))).sum(axis=1)
    assert np.allclose(dn[0], expect, rtol=1e-10)


def test_transparent_and_opaque_limits(kd_lw):
    """tau -> 0: F_up = surface emission everywhere, F_dn = 0.  tau large: the
    TOA upward flux is the top layer's blackbody emission."""
    from oracle import oracle as O
    pr = _isothermal(kd_lw, 250.0, 10)
    _, pf = O.gas_optics(kd_lw, pr)
    src = O.lw_sources(kd_lw, pr, pf)
    emis = np.full((pr.ncol, kd_lw.nband), 0.95)
    z = np.zeros((pr.ncol, 10, kd_lw.ngpt))
    up, dn, _ = O.lw_solve(kd_lw, pr.ncol, 10, z, src, emis, nang=3)
    assert np.abs(dn).max() == 0.0
    sfc = np.pi * (0.95 * src["sfc_src"]).sum(axis=1)
    assert np.allclose(up, sfc[None, :], rtol=1e-12)
    big = np.full_like(z, 1e3)
    up, dn, _ = O.lw_solve(kd_lw, pr.ncol, 10, big, src, emis, nang=3)
    top_emis = np.pi * src["lev_inc"][:, 9, :].sum(axis=1)   # level source at the top level
    assert np.allclose(up[-1], top_emis, rtol=1e-9)


def test_small_tau_series_continuous():
    """The linear-in-tau weight switches to its series at sqrt(eps); both
    sides agree with the exact expression to the series' truncation error."""
    import math
    x0 = math.sqrt(np.finfo(float).eps)
    exact = lambda x: (-math.expm1(-x)) / x - math.exp(-x)
    series = lambda x: x * (0.5 - x / 3.0)
    for x in (x0 * 0.999, x0 * 1.001):
        assert abs(series(x) - exact(x)) < 1e-12 * max(x, 1e-300) / x0 + 1e-16


def test_optimal_angle_limits(kd_lw):
    """Ds = fit1 * exp(-tau_col) + fit2: tau = 0 gives fit1 + fit2, large tau fit2."""
    from oracle import oracle as O
    z = np.zeros((2, 5, kd_lw.ngpt))
    Ds = O.optimal_angles(kd_lw, 2, 5, z)
    f = kd_lw.optimal_angle_fit
    assert np.allclose(Ds, np.repeat(f[:, 0] + f[:, 1], 8)[None, :])
    Ds = O.optimal_angles(kd_lw, 2, 5, z + 100.0)
    assert np.allclose(Ds, np.repeat(f[:, 1], 8)[None, :])


def test_heating_rate_sign_and_units():
    """Net upward flux increasing by 1 W/m2 across a layer 100 Pa thick
    (surface first, so p decreases upward) cools at g/(cp*100) K/s."""
    from oracle import oracle as O
    up = np.array([[300.0], [301.0]])
    dn = np.array([[200.0], [200.0]])
    p = np.array([[1000.0], [900.0]])
    h = O.heating_rate(up, dn, p)
    assert h[0, 0] == pytest.approx(-9.80665 / (1004.64 * 100.0), rel=1e-15)
