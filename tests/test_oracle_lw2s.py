"""Pins of the oracle's LW two-stream solver with scattering (the
doGP_lwscat path, rte's lw_solver_2stream): against the two-stream equations
integrated independently (matrix exponentials per layer, boundary conditions
solved as one linear system, no adding), and against the no-scattering solver
at the diffusivity angle when nothing scatters."""
import numpy as np
import pytest
from scipy.linalg import expm

D = 1.66  # diffusivity secant of lw_two_stream


def _two_stream_bvp(tau, ssa, g, levs_td, emis, sfc):
    """Fluxes of the hemispheric two-stream equations, top-down level order.

    With tau' increasing downward and the Planck source linear in tau' in
    each layer (pi B from the level values):
      dF+/dtau' =  g1 F+ - g2 F- - pi (g1 - g2) B
      dF-/dtau' =  g2 F+ - g1 F- + pi (g1 - g2) B
    g1 = D (1 - w (1 + g) / 2), g2 = D w (1 - g) / 2.  Each layer's map
    (F+, F-, B, 1)(top) -> (bottom) is the exponential of the augmented
    constant matrix; F-(top of atmosphere) = 0 and
    F+(surface) = (1 - emis) F-(surface) + pi emis sfc close the system.
    """
    nl = len(tau)
    nv = nl + 1
    n = 2 * nv
    A = np.zeros((n, n))
    rhs = np.zeros(n)
    row = 0
    for l in range(nl):
        g1 = D * (1 - 0.5 * ssa[l] * (1 + g[l]))
        g2 = D * 0.5 * ssa[l] * (1 - g[l])
        slope = (levs_td[l + 1] - levs_td[l]) / tau[l]
        M = np.array([[g1, -g2, -np.pi * (g1 - g2), 0.0],
                      [g2, -g1, np.pi * (g1 - g2), 0.0],
                      [0.0, 0.0, 0.0, slope],
                      [0.0, 0.0, 0.0, 0.0]])
        E = expm(M * tau[l])
        # (F+, F-)(bottom) = E[:2,:2] (F+, F-)(top) + E[:2,2] B_top + E[:2,3]
        for i in range(2):
            A[row, 2 * (l + 1) + i] = 1.0
            A[row, 2 * l + 0] -= E[i, 0]
            A[row, 2 * l + 1] -= E[i, 1]
            rhs[row] = E[i, 2] * levs_td[l] + E[i, 3]
            row += 1
    A[row, 1] = 1.0                     # F-(top) = 0
    row += 1
    A[row, 2 * nl] = 1.0                # F+(sfc) - (1 - emis) F-(sfc) = pi emis sfc
    A[row, 2 * nl + 1] = -(1.0 - emis)
    rhs[row] = np.pi * emis * sfc
    x = np.linalg.solve(A, rhs)
    return x[0::2], x[1::2]


@pytest.mark.parametrize("top_at_1", [True, False])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lw2s_matches_two_stream_equations(top_at_1, seed):
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    nl = 14
    tau = 10 ** rng.uniform(-5, 0.3, nl)        # layers up to ~2, k tau within expm's comfort
    ssa = rng.uniform(0.0, 0.999, nl)
    ssa[::4] = 0.0
    g = rng.uniform(0.0, 0.9, nl)
    levs_td = 1.0 + rng.uniform(0.0, 3.0, nl + 1)   # level sources, top first
    emis, sfc = 0.93, 3.7
    up_td, dn_td = _two_stream_bvp(tau, ssa, g, levs_td, emis, sfc)
    if top_at_1:
        up, dn = O.lw2s_gpt(tau, ssa, g, levs_td, emis, sfc, top_at_1=True)
    else:
        up, dn = O.lw2s_gpt(tau[::-1], ssa[::-1], g[::-1], levs_td[::-1], emis, sfc, top_at_1=False)
        up, dn = up[::-1], dn[::-1]
    scale = np.abs(up_td).max()
    assert np.allclose(up, up_td, rtol=1e-9, atol=1e-10 * scale)
    assert np.allclose(dn, dn_td, rtol=1e-9, atol=1e-10 * scale)


def test_lw2s_conservative_isothermal_and_black_limits():
    """Non-emitting conservative layers over a black body: nothing is
    absorbed, so F+ - F- is the same at every level.  An opaque
    non-scattering layer emits pi B at its top."""
    from oracle import oracle as O
    nl = 6
    tau = np.full(nl, 0.7)
    up, dn = O.lw2s_gpt(tau, np.full(nl, 1.0), np.full(nl, 0.3), np.zeros(nl + 1), 1.0, 2.0)
    net = up - dn
    assert np.allclose(net, net[0], rtol=1e-12)
    assert dn[0] == 0.0
    up, dn = O.lw2s_gpt(np.array([60.0]), np.zeros(1), np.zeros(1), np.array([2.5, 2.5]), 1.0, 9.0)
    assert up[0] == pytest.approx(np.pi * 2.5, rel=1e-12)
    assert dn[1] == pytest.approx(np.pi * 2.5, rel=1e-12)


@pytest.mark.parametrize("top_at_1", [False, True])
def test_no_scattering_equals_noscat_solver_at_diffusivity(kd_lw, top_at_1):
    """ssa = 0: lw_solver_2stream is the linear-in-tau no-scattering solution
    at the secant 1.66 with weight 1/2, i.e. lw_solver_noscat with lw_Ds = 1.66
    and a layer source halfway between its level sources.  Planck fractions
    uniform over layers make lw_combine_sources' geometric mean exact."""
    from oracle import oracle as O
    rng = np.random.default_rng(7)
    nc, nl, ng, nb = 3, 12, kd_lw.ngpt, kd_lw.nband
    tau = 10 ** rng.uniform(-4, 1, (nc, nl, ng))
    pf = rng.uniform(0.01, 0.2, (nc, 1, ng)) * np.ones((1, nl, 1))
    Blev = rng.uniform(1.0, 5.0, (nc, nl + 1, 1))
    inc = pf * Blev[:, 1:]
    dec = pf * Blev[:, :-1]
    lay = 0.5 * (inc + dec)
    sfc = rng.uniform(0.1, 1.0, (nc, ng))
    src = dict(lay_src=lay, lev_inc=inc, lev_dec=dec, sfc_src=sfc, sfc_jac=np.zeros_like(sfc))
    emis = rng.uniform(0.85, 1.0, (nc, nb)).T.copy()
    up1, dn1, _ = O.lw_solve(kd_lw, nc, nl, tau, src, emis, nang=1, Ds=np.full((nc, ng), 1.66),
                             top_at_1=top_at_1)
    z = np.zeros_like(tau)
    up2, dn2 = O.lw_solve_2str(kd_lw, nc, nl, tau, z, z, src, emis, top_at_1=top_at_1)
    assert np.allclose(up2, up1, rtol=1e-8, atol=1e-9)
    assert np.allclose(dn2, dn1, rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("top_at_1", [True, False])
def test_combined_level_sources_at_unequal_planck_fractions(kd_lw, top_at_1):
    """lw_combine_sources where the Planck fraction changes from layer to
    layer: the two-stream solver's level source between two layers is the
    geometric mean of the two layers' sources at that level (each layer's
    Planck fraction times the level's B), and the one layer's at the top and
    bottom.  orc_lw_solve_2str (inc/dec per layer in, band and broadband
    sums out) equals the two-stream equations integrated with those level
    sources g-point by g-point; a swapped inc/dec or a level off by one
    fails (the fractions differ by up to 5x between layers)."""
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    nc, nl, ng, nb = 1, 10, kd_lw.ngpt, kd_lw.nband
    tau = 10 ** rng.uniform(-3, 0.2, (nc, nl, ng))
    ssa = rng.uniform(0.0, 0.95, (nc, nl, ng))
    g = rng.uniform(0.0, 0.9, (nc, nl, ng))
    pf = rng.uniform(0.02, 0.1, (nc, nl, ng))            # differs between layers
    B_td = rng.uniform(1.0, 5.0, nl + 1)                  # level Planck function, top first
    B = B_td if top_at_1 else B_td[::-1]                  # array order
    inc = pf * B[None, 1:, None]                          # layer l's source at array level l+1
    dec = pf * B[None, :-1, None]                         # ... at array level l
    sfc = rng.uniform(0.1, 1.0, (nc, ng))
    emis = np.full((nc, nb), 0.95)
    src = dict(lev_inc=inc, lev_dec=dec, sfc_src=sfc)
    up, dn = O.lw_solve_2str(kd_lw, nc, nl, tau, ssa, g, src, emis, top_at_1=top_at_1)
    up_ref = np.zeros(nl + 1)
    dn_ref = np.zeros(nl + 1)
    for gp in range(ng):
        p = pf[0, :, gp]
        levs = np.empty(nl + 1)                            # array order
        levs[0] = p[0] * B[0]
        levs[nl] = p[nl - 1] * B[nl]
        levs[1:nl] = np.sqrt(p[1:] * B[1:nl] * p[:-1] * B[1:nl])
        t, s, gg = tau[0, :, gp], ssa[0, :, gp], g[0, :, gp]
        if top_at_1:
            u, d = _two_stream_bvp(t, s, gg, levs, 0.95, sfc[0, gp])
        else:
            u, d = _two_stream_bvp(t[::-1], s[::-1], gg[::-1], levs[::-1], 0.95, sfc[0, gp])
            u, d = u[::-1], d[::-1]
        up_ref += u
        dn_ref += d
    scale = np.abs(up_ref).max()
    assert np.allclose(up[:, 0], up_ref, rtol=1e-9, atol=1e-10 * scale)
    assert np.allclose(dn[:, 0], dn_ref, rtol=1e-9, atol=1e-10 * scale)
    # the arithmetic mean would differ visibly: the geometric mean is what is pinned
    assert np.abs(pf[0, 1:] - pf[0, :-1]).max() > 0.02
