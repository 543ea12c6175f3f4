"""k-distributions at the device limits (16 flavors, 32 minor-absorber
intervals per regime, 8 of them covering one g-point, 128 g-points) and one
step beyond each: the limits are checked on the host at init, before any
device work, with a message naming the limit.  The tables at the limits are
built here and run on the GPU by test_gpu_table_limits.py."""
import copy
import ctypes as C
import itertools

import numpy as np
import pytest

# the device caps (csrc/ccpprad_dev.h): flavors, minor intervals per regime,
# intervals covering one g-point in one regime, g-points
MAX_FLAV, MAX_MINOR, MAX_MINOR_G, MAX_GPT = 16, 32, 8, 128


def _flavors(kd, nflav):
    """The synthetic table's flavors, then further gas pairs up to nflav;
    each band's lower and upper regime take flavors so that every one is
    used."""
    pairs = [tuple(p) for p in kd.flavor.tolist()]
    for p in itertools.product(range(1, kd.ngas + 1), repeat=2):
        if len(pairs) >= nflav:
            break
        if p not in pairs:
            pairs.append(p)
    gf = np.zeros((kd.ngpt, 2), dtype=np.int32)
    for b in range(kd.nband):
        gf[8 * b:8 * b + 8, 0] = b % nflav + 1
        gf[8 * b:8 * b + 8, 1] = (b + 7) % nflav + 1
    used = set(gf[:, 0].tolist()) | set(gf[:, 1].tolist())
    assert used == set(range(1, nflav + 1))
    return np.array(pairs[:nflav], dtype=np.int32), gf


def _minor_spec(kd, n, per_g, rng, first_band):
    """n intervals: per_g nested ones [g0, g0 + k] at the start of
    first_band (so its first g-point lies in per_g of them), the rest random
    sub-ranges of the other bands.  Every flag combination and every gas
    occurs."""
    flags = [(s, sc, cpl) for s in (0, 1) for sc, cpl in ((0, 0), (1, 0), (1, 1))]
    spec = []
    g0 = 8 * first_band + 1
    for i in range(n):
        s, sc, cpl = flags[i % len(flags)]
        gas = 1 + i % kd.ngas
        scaling = 1 + (i + 2) % kd.ngas if sc else 0
        if i < per_g:
            lo, hi = g0, g0 + min(i, 7)
        else:
            b = (first_band + 1 + (i - per_g)) % kd.nband
            lo = 8 * b + 1 + int(rng.integers(0, 4))
            hi = lo + int(rng.integers(0, 8 * b + 8 - lo + 1))
        spec.append((gas, lo, hi, s, scaling, cpl))
    return spec


def table_at_limits(is_lw, nflav=MAX_FLAV, nminor=MAX_MINOR, per_g=MAX_MINOR_G, seed=8):
    from ccpprad import synthetic as syn
    kd = copy.deepcopy(syn.make_kdist(is_lw))
    rng = np.random.default_rng(seed)
    kd.flavor, kd.gpoint_flavor = _flavors(kd, nflav)
    kd.nflav = nflav
    kd.minor_lower = syn._minor_tables(rng, _minor_spec(kd, nminor, per_g, rng, 0), kd.ntemp, kd.neta)
    kd.minor_upper = syn._minor_tables(rng, _minor_spec(kd, nminor, per_g, rng, 3), kd.ntemp, kd.neta)
    return kd


def _coverage(m, ngpt):
    cov = np.zeros(ngpt, dtype=int)
    for lo, hi in m["limits_gpt"]:
        cov[lo - 1:hi] += 1
    return cov


def test_tables_reach_the_limits():
    for lw in (True, False):
        kd = table_at_limits(lw)
        assert kd.nflav == MAX_FLAV
        for m in (kd.minor_lower, kd.minor_upper):
            assert m["n"] == MAX_MINOR
            assert _coverage(m, kd.ngpt).max() == MAX_MINOR_G


def _init_error(kd):
    from ccpprad import lib, capi
    L = lib.load()
    hk = capi.kdist_struct(kd)
    ctx = C.c_void_p()
    buf, n = capi.errbuf()
    init = L.ccpprad_lw_init if kd.is_lw else L.ccpprad_sw_init
    rc = init(C.addressof(hk.struct), None, None, 0, 127, 1, C.byref(ctx), buf, n)
    if rc == 0:
        L.ccpprad_finalize(ctx)
    return rc, buf.value


@pytest.mark.parametrize("lw", [True, False])
def test_beyond_the_limits_is_an_init_error(lw):
    """One flavor or one minor interval past the caps: rejected on the host
    before any device work (no GPU here)."""
    rc, msg = _init_error(table_at_limits(lw, nflav=MAX_FLAV + 1))
    assert rc != 0 and b"device limits" in msg
    rc, msg = _init_error(table_at_limits(lw, nminor=MAX_MINOR + 1))
    assert rc != 0 and b"too many minor intervals" in msg
