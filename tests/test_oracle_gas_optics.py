"""Pins of the oracle's gas optics (get_col_dry, interpolation, major/minor
absorption, Planck fraction, Rayleigh) by closed forms and invariants."""
import copy

import numpy as np
import pytest


def _prof(kd, ncol=16, nlay=40, seed=1):
    from ccpprad import synthetic as syn
    return syn.make_profiles(ncol, nlay, seed=seed, clouds=False, kd=kd)


def test_col_dry_is_the_hydrostatic_mass(kd_lw):
    """Dry air column over the whole profile equals (p_sfc - p_top)/g of
    dry-air mass: sum(col_dry) * m_dry / N_A * 1e4 [kg/m2]."""
    from oracle import oracle as O
    pr = _prof(kd_lw)
    for g in pr.vmr:
        pr.vmr[g] = np.zeros_like(pr.vmr[g])
    tot = 0.0
    for l in range(pr.nlay):
        col = np.zeros(kd_lw.ngas + 1)
        out = [np.zeros(1, dtype=np.int32) for _ in range(3)] + [np.zeros(1), np.zeros(1)]
        import ctypes as C
        je = np.zeros(2 * kd_lw.nflav, dtype=np.int32)
        cm = np.zeros(2 * kd_lw.nflav)
        fm = np.zeros(4 * kd_lw.nflav)
        from ccpprad import capi
        h = capi.kdist_struct(kd_lw)
        dp = pr.p_lev[l, 0] - pr.p_lev[l + 1, 0]
        O.lib().orc_interp_layer(C.byref(h.struct), pr.p_lay[l, 0], pr.t_lay[l, 0], dp,
                                 np.zeros(kd_lw.ngas).ctypes.data_as(O.P_D), *[O._p(o, O.P_I) for o in out[:3]],
                                 O._p(out[3]), O._p(out[4]), O._p(col), O._p(je, O.P_I), O._p(cm), O._p(fm))
        tot += col[0]
    mass = tot * 0.028964 / 6.02214076e23 * 1e4
    assert mass == pytest.approx((pr.p_lev[0, 0] - pr.p_lev[-1, 0]) / 9.80665, rel=1e-12)


def _interp(kd, p, T, vmr, dp=1000.0):
    import ctypes as C
    from oracle import oracle as O
    from ccpprad import capi
    h = capi.kdist_struct(kd)
    ji = [np.zeros(1, dtype=np.int32) for _ in range(3)]
    ft, fp = np.zeros(1), np.zeros(1)
    col = np.zeros(kd.ngas + 1)
    je = np.zeros(2 * kd.nflav, dtype=np.int32)
    cm = np.zeros(2 * kd.nflav)
    fm = np.zeros(4 * kd.nflav)
    O.lib().orc_interp_layer(C.byref(h.struct), p, T, dp, np.ascontiguousarray(vmr, dtype=np.float64).ctypes.data_as(O.P_D),
                             *[O._p(o, O.P_I) for o in ji], O._p(ft), O._p(fp), O._p(col), O._p(je, O.P_I),
                             O._p(cm), O._p(fm))
    return dict(jtemp=int(ji[0][0]), jpress=int(ji[1][0]), itropo=int(ji[2][0]), ftemp=ft[0], fpress=fp[0],
                col=col, jeta=je.reshape(-1, 2), col_mix=cm.reshape(-1, 2), fminor=fm.reshape(-1, 2, 2))


def test_interpolation_indices_brute_force(kd_lw):
    """jtemp/jpress bracket T and ln p on the reference grids (searchsorted),
    weights lie in [0,1], and the eight fmajor weights of a flavor sum to 1."""
    rng = np.random.default_rng(2)
    for _ in range(300):
        p = float(np.exp(rng.uniform(np.log(2.0), np.log(1.05e5))))
        T = float(rng.uniform(161.0, 354.0))
        vmr = rng.uniform(1e-7, 1e-2, kd_lw.ngas)
        r = _interp(kd_lw, p, T, vmr)
        jt = np.searchsorted(kd_lw.temp_ref, T, side="right")       # temp_ref[jt-1] <= T < temp_ref[jt]
        assert r["jtemp"] == min(max(jt, 1), kd_lw.ntemp - 1)
        assert 0.0 <= r["ftemp"] <= 1.0 + 1e-12
        lp = -np.log(kd_lw.press_ref)                                # increasing
        jp = np.searchsorted(lp, -np.log(p), side="right")
        assert r["jpress"] == min(max(jp, 1), kd_lw.npres - 1)
        assert -1e-12 <= r["fpress"] <= 1.0 + 1e-12
        assert r["itropo"] == (1 if np.log(p) > np.log(kd_lw.press_ref_trop) else 2)
        w = r["fminor"].sum(axis=(1, 2))                             # sum over T corners and eta
        assert np.allclose(w, 1.0, atol=1e-12)
        assert (r["jeta"] >= 1).all() and (r["jeta"] <= kd_lw.neta - 1).all()


def _linear_kd(kd, A, Be, Cp, Dt):
    """k(g, eta, p-row, T) = A + Be*e + Cp*p + Dt*t (0-based node indices):
    exactly trilinear, so interpolation reproduces the continuous value."""
    k2 = copy.deepcopy(kd)
    t = np.arange(kd.ntemp)[:, None, None, None]
    p = np.arange(kd.npres + 1)[None, :, None, None]
    e = np.arange(kd.neta)[None, None, :, None]
    k2.kmajor = np.ascontiguousarray(np.broadcast_to(A + Be * e + Cp * p + Dt * t,
                                                     kd.kmajor.shape).astype(np.float64))
    vr = np.repeat(k2.vmr_ref[:1], kd.ntemp, axis=0)   # T-independent reference ratios
    k2.vmr_ref = np.ascontiguousarray(vr)
    for m in (k2.minor_lower, k2.minor_upper):
        m["kminor"] = np.zeros_like(m["kminor"])
    return k2


def test_major_absorption_exact_on_trilinear_table(kd_lw):
    from oracle import oracle as O
    from ccpprad import synthetic as syn
    A, Be, Cp, Dt = 2e-24, 3e-25, 5e-26, 7e-26
    kd = _linear_kd(kd_lw, A, Be, Cp, Dt)
    pr = _prof(kd, ncol=4, nlay=30, seed=7)
    tau, _ = O.gas_optics(kd, pr)
    v = np.stack([pr.vmr[g] for g in syn.GAS_NAMES])
    for c in range(pr.ncol):
        for l in range(pr.nlay):
            p, T = pr.p_lay[l, c], pr.t_lay[l, c]
            r = _interp(kd, p, T, v[:, l, c], dp=pr.p_lev[l, c] - pr.p_lev[l + 1, c])
            it = r["itropo"]
            tcont = (T - kd.temp_ref[0]) / (kd.temp_ref[1] - kd.temp_ref[0])
            dlp = (np.log(kd.press_ref[0]) - np.log(kd.press_ref[-1])) / (kd.npres - 1)
            pcont = (np.log(kd.press_ref[0]) - np.log(p)) / dlp + (it - 1)
            for g in (0, 37, 77, 127):
                f = kd.gpoint_flavor[g, it - 1] - 1
                g1, g2 = kd.flavor[f]
                ratio = kd.vmr_ref[0, g1, it - 1] / kd.vmr_ref[0, g2, it - 1]
                cm = r["col"][g1] + ratio * r["col"][g2]
                eta = r["col"][g1] / cm
                # jtemp clamps at the table edge: the linear form stays exact
                expected = cm * (A + Be * eta * (kd.neta - 1) + Cp * pcont + Dt * tcont)
                assert tau[c, l, g] == pytest.approx(expected, rel=1e-11)


def test_minor_absorption_constant_tables(kd_lw):
    """kmajor = 0 and kminor = K: tau(g) = K * sum over the intervals that
    cover g of the absorber amount with its density / scaling-gas factors."""
    from oracle import oracle as O
    from ccpprad import synthetic as syn
    kd = copy.deepcopy(kd_lw)
    kd.kmajor = np.zeros_like(kd.kmajor)
    K = 1.0e-27
    for m in (kd.minor_lower, kd.minor_upper):
        m["kminor"] = np.full_like(m["kminor"], K)
    pr = _prof(kd, ncol=3, nlay=60, seed=11)
    tau, _ = O.gas_optics(kd, pr)
    v = np.stack([pr.vmr[g] for g in syn.GAS_NAMES])
    for c in range(pr.ncol):
        for l in range(pr.nlay):
            p, T = pr.p_lay[l, c], pr.t_lay[l, c]
            dp = pr.p_lev[l, c] - pr.p_lev[l + 1, c]
            r = _interp(kd, p, T, v[:, l, c], dp=dp)
            col = r["col"]
            exp = np.zeros(kd.ngpt)
            m = kd.minor_lower if r["itropo"] == 1 else kd.minor_upper
            for i in range(m["n"]):
                s = col[m["idx_minor"][i]]
                if m["scales_with_density"][i]:
                    s *= p / 100.0 / T                         # hPa / K
                    if m["idx_scaling"][i] > 0:
                        x = col[m["idx_scaling"][i]] / (col[0] + col[1])   # fraction of moist air
                        s *= (1 - x) if m["scale_by_complement"][i] else x
                gs, ge = m["limits_gpt"][i]
                exp[gs - 1:ge] += K * s
            assert np.allclose(tau[c, l], exp, rtol=1e-12, atol=0)


def test_planck_fraction_sums_to_one_per_band(kd_lw):
    from oracle import oracle as O
    pr = _prof(kd_lw, ncol=8, nlay=50, seed=3)
    _, pf = O.gas_optics(kd_lw, pr)
    s = pf.reshape(pr.ncol, pr.nlay, kd_lw.nband, 8).sum(axis=-1)
    assert np.allclose(s, 1.0, atol=1e-13)


def test_humid_col_dry_is_the_hydrostatic_mass(kd_lw):
    """With water vapour: the layer's dry-air and water columns together
    weigh the layer, (col_dry m_dry + col_h2o m_h2o) / N_A 1e4 = dp / g,
    and col_h2o = vmr_h2o col_dry (get_col_dry's moist-air molar mass)."""
    vmr = np.zeros(kd_lw.ngas)
    for vh in (0.0, 0.005, 0.03):
        vmr[0] = vh
        r = _interp(kd_lw, 50000.0, 270.0, vmr, dp=2500.0)
        cd, ch = r["col"][0], r["col"][1]
        mass = (cd * 0.028964 + ch * 0.018016) / 6.02214076e23 * 1e4
        assert mass == pytest.approx(2500.0 / 9.80665, rel=1e-12)
        assert ch == pytest.approx(vh * cd, rel=1e-15, abs=0.0)


def test_rayleigh_uses_the_moist_column(kd_sw):
    """Rayleigh depth with a constant krayl K is K (col_dry + col_h2o): in
    a humid layer the ratio tau_ray / (K col_dry) is 1 + vmr_h2o (so neither
    col_dry alone nor the total including other gases), in a dry layer it
    is 1.  col_dry is the oracle's interpolation column, itself pinned by
    the hydrostatic mass above."""
    from oracle import oracle as O
    kd = copy.deepcopy(kd_sw)
    K = 3e-27
    kd.krayl_lower = np.full_like(kd.krayl_lower, K)
    kd.krayl_upper = np.full_like(kd.krayl_upper, K)
    pr = _prof(kd, ncol=3, nlay=20, seed=5)
    pr.vmr["h2o"] = pr.vmr["h2o"].copy()
    pr.vmr["h2o"][:, 0] = 0.0            # a dry column
    pr.vmr["h2o"][:, 1] = 0.03           # a very humid column
    _, ray = O.gas_optics(kd, pr)
    vs = O.vmr_stack(pr)
    for c in range(pr.ncol):
        for l in range(pr.nlay):
            dp = abs(pr.p_lev[l, c] - pr.p_lev[l + 1, c])
            r = _interp(kd, pr.p_lay[l, c], pr.t_lay[l, c], vs[:, l, c], dp=dp)
            ratio = ray[c, l] / (K * r["col"][0])
            assert np.allclose(ratio, 1.0 + pr.vmr["h2o"][l, c], rtol=1e-13, atol=0)


def test_cmp_tlev_matches_reference_fortran():
    """synthetic.cmp_tlev restates radiation_tools.F90:16-84; compare with the
    reference's own routine built in oracle/_ref."""
    import ctypes as C
    from oracle import oracle as O
    from ccpprad import synthetic as syn
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    pr = syn.make_profiles(7, 33, seed=9, clouds=False)
    ncol, nlay = 7, 33
    ref = np.zeros((nlay + 1, ncol))
    P = C.POINTER(C.c_double)
    R.ref_cmp_tlev(ncol, nlay, 1.0, pr.p_lay.ctypes.data_as(P), pr.t_lay.ctypes.data_as(P),
                   pr.p_lev.ctypes.data_as(P), np.ascontiguousarray(pr.tsfc).ctypes.data_as(P),
                   ref.ctypes.data_as(P))
    ours = syn.cmp_tlev(pr.p_lay, pr.t_lay, pr.p_lev, pr.tsfc, 1.0)
    assert np.allclose(ours, ref, rtol=0, atol=1e-12)
