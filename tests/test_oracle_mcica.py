"""Pins of the oracle's McICA pieces.

MT19937: bit-exact against the reference's own Fortran
(physics/Radiation/mersenne_twister.f) through the committed golden vectors
and, where built, oracle/_ref directly.  sampled_mask: invariants that a wrong
index, a dropped condition or a swapped random stream would break."""
import ctypes as C
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mt19937_matches_reference_golden_vectors():
    from oracle import oracle as O
    g = json.load(open(os.path.join(HERE, "golden", "mt19937_draws.json")))
    for seed, vals in g["draws"].items():
        got = O.mt_draws(int(seed), len(vals))
        assert np.array_equal(got, np.array(vals)), seed
    for key, vals in g["chunked"].items():
        seed, chunk, n = (int(x) for x in key.split(":"))
        got = O.mt_draws(seed, chunk * n)   # chunked calls continue one stream
        assert np.array_equal(got, np.array(vals)), key


def test_mt19937_matches_reference_library():
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for seed in (3, 1000, 65536 + 128, -123456):
        a = np.zeros(2000)
        R.ref_mt_draws(seed, a.size, a.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.array_equal(O.mt_draws(seed, a.size), a)


def test_mt19937_negative_branch_and_single_precision_branch():
    """Draws whose signed value is negative go through FP64; non-negative ones
    through default REAL: the latter are multiples of float spacing."""
    from oracle import oracle as O
    d = O.mt_draws(4357, 5000)
    assert (d >= 0).all() and (d <= 1).all()
    lo = d[d < 0.5] * (2.0 ** 32 - 1)       # non-negative signed draws
    assert np.array_equal(lo, lo.astype(np.float32).astype(np.float64))


def _cf(ncol, nlay, rng, frac=0.6):
    cf = rng.uniform(0, 1, (nlay, ncol))
    cf[rng.uniform(0, 1, (nlay, ncol)) > frac] = 0.0
    return cf


def test_sampled_mask_clear_layers_false_and_frequency():
    from oracle import oracle as O
    rng = np.random.default_rng(1)
    nc, nl, ng = 400, 20, 128
    cf = _cf(nc, nl, rng)
    for iovr in ("iovr_maxrand", "iovr_rand", "iovr_max"):
        m = O.mcica_mask(np.arange(nc) + 129, cf, ng, iovr=iovr)  # (nc, ng, nl)
        assert not m.transpose(2, 0, 1)[cf == 0].any()
        freq = m.mean(axis=1).T                                   # (nl, nc)
        sel = cf > 0
        assert abs(freq[sel].mean() - cf[sel].mean()) < 0.01


def test_sampled_mask_max_random_shares_randoms_in_cloud_blocks():
    """Max-random: inside a contiguous cloudy block the same random vector is
    used, so with equal cloud fraction the masks are identical; a new block
    gets new randoms (masks differ)."""
    from oracle import oracle as O
    nl, ng = 8, 128
    cf = np.zeros((nl, 1))
    cf[1:4, 0] = 0.5      # block 1
    cf[5:7, 0] = 0.5      # block 2
    m = O.mcica_mask(np.array([4242]), cf, ng)[0]   # (ng, nl)
    assert np.array_equal(m[:, 1], m[:, 2]) and np.array_equal(m[:, 2], m[:, 3])
    assert np.array_equal(m[:, 5], m[:, 6])
    assert not np.array_equal(m[:, 3], m[:, 5])
    r = O.mt_draws(4242, ng * nl).reshape(nl, ng)
    assert np.array_equal(m[:, 1], r[1] > 0.5)      # block starts draw their own layer's randoms
    assert np.array_equal(m[:, 5], r[5] > 0.5)


def test_sampled_mask_maximum_overlap_is_nested():
    """iovr_max: one random vector for all layers, so masks are nested by
    cloud fraction."""
    from oracle import oracle as O
    cf = np.array([[0.2], [0.7], [0.4], [0.9]])
    m = O.mcica_mask(np.array([77]), cf, 128, iovr="iovr_max")[0]
    order = np.argsort(cf[:, 0])
    for a, b in zip(order[:-1], order[1:]):
        assert (m[:, a] <= m[:, b]).all()


def _per_layer_random_mask(seeds, cf, ng):
    """Random overlap proper: every layer draws its own vector."""
    from oracle import oracle as O
    nl, nc = cf.shape
    m = np.zeros((nc, ng, nl), dtype=np.uint8)
    for c in range(nc):
        r = O.mt_draws(int(seeds[c]), ng * nl).reshape(nl, ng)
        m[c] = (r > 1 - cf[:, c][:, None]).T
    return m


def test_iovr_rand_is_max_random_in_the_reference():
    """rrtmgp_lw_main.F90:470-471 sends iovr_rand down the same sampled_mask
    branch as iovr_maxrand: the two masks are identical."""
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    cf = _cf(20, 15, rng)
    seeds = np.arange(20) + 31
    assert np.array_equal(O.mcica_mask(seeds, cf, 128, iovr="iovr_rand"),
                          O.mcica_mask(seeds, cf, 128, iovr="iovr_maxrand"))


def test_sampled_mask_exponential_limits():
    """Exponential overlap with alpha = 1 equals max-random within a block;
    with alpha = 0 every cloudy layer uses its own random (random overlap)."""
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    nc, nl, ng = 50, 12, 128
    cf = rng.uniform(0.05, 1, (nl, nc))
    seeds = np.arange(nc) + 500
    mr = O.mcica_mask(seeds, cf, ng, iovr="iovr_maxrand")
    m1 = O.mcica_mask(seeds, cf, ng, iovr="iovr_exp", overlap=np.ones((nl, nc)))
    assert (m1 != mr).mean() < 1e-4          # (r - 0.5) + 0.5 may move r by an ulp
    m0 = O.mcica_mask(seeds, cf, ng, iovr="iovr_exp", overlap=np.zeros((nl, nc)))
    assert (m0 != _per_layer_random_mask(seeds, cf, ng)).mean() < 1e-4


def test_sampled_mask_decorrelation_limits():
    """Decorrelation-length overlap: alpha = 0 draws a fresh random per layer
    (random overlap); alpha = 1 reuses the layer above's random for every
    g-point of a cloudy layer."""
    from oracle import oracle as O
    rng = np.random.default_rng(4)
    nc, nl, ng = 30, 10, 112
    cf = rng.uniform(0.05, 1, (nl, nc))
    seeds = np.arange(nc) + 9
    m0 = O.mcica_mask(seeds, cf, ng, iovr="iovr_dcorr", overlap=np.zeros((nl, nc)))
    assert np.array_equal(_per_layer_random_mask(seeds, cf, ng), m0)
    m1 = O.mcica_mask(seeds, cf, ng, iovr="iovr_dcorr", overlap=np.ones((nl, nc)))
    for c in range(nc):
        r = O.mt_draws(int(seeds[c]), ng * nl).reshape(nl, ng)
        for l in range(1, nl):
            assert np.array_equal(m1[c, :, l], r[l - 1] > 1 - cf[l, c])
