"""Pins of the oracle's Pade cloud optics (rte-rrtmgp's compute_all_from_pade,
whose source is absent from the reference tree: parity unpinned beyond these
closed forms).

- Constant approximants (only the numerator's constant term set) give
  tau = path c_ext, ssa = 1 - max(0, c_coa) and g = c_asy, each from the size
  regime rte's expression min(floor((re - b2) / b3) + 2, 3) selects: radii on
  either side of the two regime edges pick different constants.
- A rational approximant a re / (1 + b re^2) (extinction) is evaluated to
  rounding at arbitrary radii.
- A negative co-albedo approximant is clamped to 0 (ssa = 1)."""
import dataclasses

import numpy as np
import pytest


def _constant_pade(is_lw, ext, coa, asy):
    """ext, coa, asy: (3,) per-regime constants, the same for every band and
    both phases."""
    from ccpprad import synthetic as syn
    lut = syn.make_cloud_pade(is_lw)
    nb, nr = lut.nband, lut.nrghice
    P = dict(lut.pade)
    for ph, sfx in ((0, "liq"), (1, "ice")):
        e = np.zeros((6, 3, nb)); s = np.zeros((5, 3, nb)); a = np.zeros((5, 3, nb))
        e[0] = np.asarray(ext)[:, None]; s[0] = np.asarray(coa)[:, None]; a[0] = np.asarray(asy)[:, None]
        if ph:
            e, s, a = (np.stack([x] * nr) for x in (e, s, a))
        P[f"pade_ext{sfx}"], P[f"pade_ssa{sfx}"], P[f"pade_asy{sfx}"] = e, s, a
    return dataclasses.replace(lut, pade=P)


def test_constant_approximants_pick_the_regime():
    from oracle import oracle as O
    lut = _constant_pade(True, [1.0, 2.0, 3.0], [0.1, 0.2, 0.3], [0.7, 0.8, 0.9])
    # liquid bounds 2.5, 6.5, 11, 21.5: regime edges at re = 6.5 and 6.5 + 11 = 17.5
    rl = np.array([[3.0, 6.4, 6.6, 17.4, 17.6, 21.5]])
    cases = [0, 0, 1, 1, 2, 2]
    lwp = np.full_like(rl, 2.0)
    z = np.zeros_like(rl)
    t, s, g = O.cloud_optics(lut, lwp, z, rl, np.full_like(rl, 50.0))
    for j, r in enumerate(cases):
        assert t[0, 0, j] == 2.0 * [1.0, 2.0, 3.0][r]
        assert abs(s[0, 0, j] - (1.0 - [0.1, 0.2, 0.3][r])) < 1e-15
        assert abs(g[0, 0, j] - [0.7, 0.8, 0.9][r]) < 1e-15
    # ice bounds 10, 30, 70, 180: edges at 30 and 100
    ri = np.array([[12.0, 29.0, 31.0, 99.0, 101.0, 179.0]])
    t, s, g = O.cloud_optics(lut, z, lwp, np.full_like(ri, 10.0), ri)
    for j, r in enumerate(cases):
        assert t[0, 0, j] == 2.0 * [1.0, 2.0, 3.0][r]


def test_rational_extinction_and_clamped_coalbedo():
    from ccpprad import synthetic as syn
    from oracle import oracle as O
    lut = _constant_pade(False, [0.0, 0.0, 0.0], [-0.2, -0.2, -0.2], [0.85, 0.85, 0.85])
    P = dict(lut.pade)
    e = np.zeros((6, 3, lut.nband))
    e[1], e[4] = 1.7, 0.9            # 1.7 re / (1 + 0.9 re^2)
    P["pade_extliq"] = e
    lut = dataclasses.replace(lut, pade=P)
    rng = np.random.default_rng(2)
    rl = rng.uniform(2.5, 21.5, (4, 7))
    lwp = rng.uniform(1.0, 100.0, (4, 7))
    t, s, g = O.cloud_optics(lut, lwp, np.zeros_like(rl), rl, np.full_like(rl, 50.0))
    ref = lwp * 1.7 * rl / (1.0 + 0.9 * rl * rl)
    np.testing.assert_allclose(t[3], ref, rtol=1e-14)
    np.testing.assert_array_equal(s[3], np.ones_like(rl))      # co-albedo -0.2 clamped: ssa = 1
    np.testing.assert_allclose(g[3], 0.85, rtol=1e-15)


def test_pade_radius_out_of_range_is_an_error():
    from ccpprad import synthetic as syn
    from oracle import oracle as O
    lut = syn.make_cloud_pade(True)
    with pytest.raises(ValueError):
        O.cloud_optics(lut, np.ones((1, 1)), np.zeros((1, 1)), np.full((1, 1), 25.0), np.full((1, 1), 50.0))


def test_library_parses_pade_cloud_file(tmp_path):
    """The library reads a Pade-only netCDF cloud file up to the device
    upload (no GPU here)."""
    import ctypes as C
    import torch
    from ccpprad import capi, lib, synthetic as syn
    from nc_writer import write_kdist_nc, write_pade_nc
    if torch.cuda.is_available():
        pytest.skip("covered by the GPU tests")
    L = lib.load()
    write_kdist_nc(tmp_path / "k.nc", syn.make_kdist(True))
    write_pade_nc(tmp_path / "c.nc", syn.make_cloud_pade(True))
    names = (C.c_char_p * 6)(*[s.encode() for s in syn.GAS_NAMES])
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    rc = L.ccpprad_lw_init_files(str(tmp_path / "k.nc").encode(), str(tmp_path / "c.nc").encode(), 0, names, 6,
                                 127, 1, C.byref(ctx), buf, n)
    msg = buf.value.decode()
    assert rc != 0 and "upload" in msg, msg
