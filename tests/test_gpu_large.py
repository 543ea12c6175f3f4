"""The full-size workloads through the multi-launch path: LW+SW all sky at
262,144 columns x 127 layers (four launches of 65,536 columns) and the
learned LW gas optics at the same size, checked against the FP64 oracle on a
seeded 4,096-column subset made of whole 64-column blocks drawn from every
launch (the last block included), and bit for bit against a run with
131,072 columns per launch (other launch boundaries, same columns).

The oracle runs the subset as its own call; the McICA seeds of the full call
(isubc 1: ngpt + global column) are handed to it as isubc 2 seeds, so the
two draw the same streams.  Tolerances: the plain-path bar (0.05 W/m2,
0.01 K/day max-abs) for the table path, HR RMSE 0.1 K/day for the
emulator."""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NCOL, NLAY, BLK = 262_144, 127, 64
NSUB = 4096


@pytest.fixture(scope="module")
def big(kd_lw):
    from ccpprad import synthetic as syn, capi
    pr = syn.make_profiles(NCOL, NLAY, seed=20261205, clouds=True, kd=kd_lw)
    arr = capi.profile_fields(pr)
    rng = np.random.default_rng(17)
    nblk = NCOL // BLK
    blocks = np.sort(np.concatenate([rng.choice(nblk - 1, NSUB // BLK - 1, replace=False), [nblk - 1]]))
    sub = (blocks[:, None] * BLK + np.arange(BLK)[None, :]).reshape(-1)
    return pr, arr, sub


def _subset(arr, cols):
    out = {}
    for k, v in arr.items():
        if v.ndim == 1:
            out[k] = np.ascontiguousarray(v[cols])
        elif v.ndim == 2:
            out[k] = np.ascontiguousarray(v[:, cols])
    return out


def _to_dev(arr):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in arr.items()}


def _check(name, got, ref, tol, unit, scale=1.0):
    mx, rms = report(name, got * scale, ref * scale, unit)
    assert mx < tol, name
    return mx, rms


def test_lw_262144_subset_and_launch_boundaries(big, kd_lw, lut_lw):
    from ccpprad import lib, capi, device
    from oracle import oracle as O
    pr, arr, sub = big
    ints = capi.lw_default_ints(NCOL, NLAY, blksz=BLK, nangle=3, clrsky=True)
    s = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=BLK)
    assert s.set_launch_columns(65536) == 65536
    dev = _to_dev(arr)
    outs = []
    for launch in (65536, 131072):
        s.set_launch_columns(launch)
        o = device.lw_outputs(NCOL, NLAY)
        s.run_dev(ints, {**dev, **o})
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in o.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    sa = _subset(arr, sub)
    sa["icseed_lw"] = (kd_lw.ngpt + sub + 1).astype(np.int32)
    ref = dict(sa)
    ref.update(O.lw_outputs(NSUB, NLAY))
    O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(NSUB, NLAY, blksz=BLK, nangle=3, clrsky=True, isubc=2), ref)
    g = outs[0]
    for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky"):
        _check(n, g[n][:, sub], ref[n], 0.05, "W/m2")
    for n in ("htrlw", "htrlwc"):
        _check(n, g[n][:, sub], ref[n], 0.01, "K/day", 86400.0)


def test_sw_262144_subset_and_launch_boundaries(big, kd_sw, lut_sw):
    """SW with three quarters of the columns daylit (scattered): 196,608
    daylit columns in launches of 65,536; the subset is whole daylit blocks."""
    from ccpprad import lib, capi, device
    from oracle import oracle as O
    pr, arr, _ = big
    rng = np.random.default_rng(23)
    nday = 3 * NCOL // 4
    idx = np.sort(rng.choice(NCOL, nday, replace=False))
    nblk = nday // BLK
    blocks = np.sort(np.concatenate([rng.choice(nblk - 1, NSUB // BLK - 1, replace=False), [nblk - 1]]))
    pos = (blocks[:, None] * BLK + np.arange(BLK)[None, :]).reshape(-1)
    cols = idx[pos]
    a = dict(arr)
    a["idx"] = (idx + 1).astype(np.int32)
    ints = capi.sw_default_ints(NCOL, nday, NLAY, blksz=BLK, clrsky=True)
    s = lib.SWScheme(kd_sw, lut_sw, NLAY, blksz=BLK)
    dev = _to_dev(a)
    outs = []
    for launch in (65536, 131072):
        s.set_launch_columns(launch)
        o = device.sw_outputs(NCOL, NLAY)
        s.run_dev(ints, {**dev, **o})
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in o.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    sa = _subset(arr, cols)
    sa["idx"] = np.arange(1, NSUB + 1, dtype=np.int32)
    sa["icseed_sw"] = (kd_sw.ngpt + cols + 1).astype(np.int32)
    ref = dict(sa)
    ref.update(O.sw_outputs(NSUB, NLAY))
    O.sw_main_run(kd_sw, lut_sw, capi.sw_default_ints(NSUB, NSUB, NLAY, blksz=BLK, clrsky=True, isubc=2), ref)
    g = outs[0]
    for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky"):
        _check(n, g[n][:, cols], ref[n], 0.05, "W/m2")
    for n in ("htrsw", "htrswc"):
        _check(n, g[n][:, cols], ref[n], 0.01, "K/day", 86400.0)
    _check("scmpsw", g["scmpsw"][cols], ref["scmpsw"], 0.05, "W/m2")
    # cldtausw: only the last daylit block's columns (both runs end there)
    last = cols[-BLK:]
    _check("cldtausw (last block)", g["cldtausw"][:, last], ref["cldtausw"][:, -BLK:], 1e-4, "")
    night = np.ones(NCOL, dtype=bool)
    night[idx] = False
    for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "htrsw"):
        assert not g[n][:, night].any(), n
    assert not g["scmpsw"][night].any()


def test_emulator_262144_subset(big, kd_lw, lut_lw):
    """The learned LW gas optics (packaged weights) at 262,144 columns: HR
    RMSE against the FP64 table-path oracle on the subset, and the same
    result for two launch sizes."""
    from ccpprad import lib, capi, device, mlp
    from oracle import oracle as O
    pr, arr, sub = big
    ints = capi.lw_default_ints(NCOL, NLAY, blksz=BLK, nangle=3)
    s = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=BLK)
    emu = mlp.LWEmulator.load(mlp.DEFAULT_WEIGHTS)
    emu.attach(s)
    dev = _to_dev(arr)
    outs = []
    for launch in (65536, 131072):
        s.set_launch_columns(launch)
        o = device.lw_outputs(NCOL, NLAY)
        s.run_dev(ints, {**dev, **o})
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in o.items()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    sa = _subset(arr, sub)
    sa["icseed_lw"] = (kd_lw.ngpt + sub + 1).astype(np.int32)
    ref = dict(sa)
    ref.update(O.lw_outputs(NSUB, NLAY))
    O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(NSUB, NLAY, blksz=BLK, nangle=3, isubc=2), ref)
    g = outs[0]
    _, rms = report("htrlw (NN)", g["htrlw"][:, sub] * 86400, ref["htrlw"] * 86400, "K/day")
    assert rms <= 0.1
    _, frms = report("fluxlwUP_allsky (NN)", g["fluxlwUP_allsky"][:, sub], ref["fluxlwUP_allsky"], "W/m2")
    assert frms < 1.0
    emu.close()
