"""The McICA cloud masks of the GPU bit for bit against the oracle's
sampled_mask (rrtmgp_sampling.F90:84-182) fed with the reference's own
MT19937 stream (mersenne_twister.f, pinned by tests/golden/mt19937_draws.json):
every overlap flag, isubc 1 (seed ngpt + column) and 2 (icseed), LW over
all columns and SW over scattered daylit columns (seeds by the column the
daylit index points at, rrtmgp_sw_main.F90:479-487), several launches per
call, and columns whose last cloudy layer is the first or the last layer.
Masks are integer work: the bar is equality."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

IOVRS = ["iovr_rand", "iovr_maxrand", "iovr_max", "iovr_dcorr", "iovr_exp", "iovr_exprand"]


def _clouds(ncol, nlay, seed):
    rng = np.random.default_rng(seed)
    cf = np.where(rng.uniform(size=(nlay, ncol)) < 0.35, rng.uniform(0.0, 1.0, (nlay, ncol)), 0.0)
    cf[:, 0] = 0.0
    cf[0, 0] = 0.8                # only the first layer cloudy
    cf[:, 1] = 0.0
    cf[nlay - 1, 1] = 0.6         # only the last layer cloudy
    cf[:, 2] = rng.uniform(0.05, 1.0, nlay)   # every layer cloudy
    cf[:, 3] = 0.0                # a clear column
    cf[:, 8:16] = 0.0             # a clear block (blocks of 8)
    cf[5, 9] = 5e-7               # below eps: the block stays clear
    return np.ascontiguousarray(cf)


def _gpu_bits(words, ngpt):
    """(n, nlay, 4) uint32 words -> (n, ngpt, nlay) 0/1."""
    w = words.astype(np.uint64)
    n, nl, _ = w.shape
    g = np.arange(ngpt)
    bits = (w[:, :, g // 32] >> (g % 32).astype(np.uint64)) & 1
    return np.ascontiguousarray(bits.transpose(0, 2, 1)).astype(np.uint8)


def _block_cloudy(cf_cols, blksz):
    """The reference's per-block flag: any column with total cover above eps
    (rrtmgp_lw_main.F90:353-364)."""
    z0 = np.clip(np.min(1.0 - cf_cols, axis=0), None, None)
    z0 = np.where(z0 <= 1e-12, 0.0, z0)
    z0 = np.where(z0 > 1.0 - 1e-6, 1.0, z0)
    col = (1.0 - z0) > 1e-6
    return np.repeat(col.reshape(-1, blksz).any(axis=1), blksz)


@pytest.mark.parametrize("isubc", [1, 2])
@pytest.mark.parametrize("iovr", IOVRS)
@pytest.mark.parametrize("lw", [True, False])
def test_mcica_mask_bit_exact(kd_lw, kd_sw, lut_lw, lut_sw, lw, iovr, isubc):
    from ccpprad import lib, capi
    from oracle import oracle as O
    kd, lut = (kd_lw, lut_lw) if lw else (kd_sw, lut_sw)
    ncol, nlay, blksz = 96, 127, 8
    ngpt = kd.ngpt
    cf = _clouds(ncol, nlay, seed=41 + isubc)
    rng = np.random.default_rng(43)
    ovl = np.ascontiguousarray(rng.uniform(0.0, 1.0, (nlay, ncol)))
    icseed = rng.integers(1, 2 ** 31 - 1, ncol).astype(np.int32)
    if lw:
        cols = np.arange(ncol)
        ints = capi.lw_default_ints(ncol, nlay, blksz=blksz, iovr=iovr, isubc=isubc)
        arrays = {"cld_frac": cf, "cloud_overlap_param": ovl, "icseed_lw": icseed}
        s = lib.LWScheme(kd, lut, nlay, blksz=blksz)
    else:
        nday = 64
        cols = np.sort(rng.choice(np.arange(4, ncol), nday - 4, replace=False))
        cols = np.concatenate([[0, 1, 2, 3], cols])   # the edge columns are daylit
        cols = np.sort(cols)
        ints = capi.sw_default_ints(ncol, nday, nlay, blksz=blksz, iovr=iovr, isubc=isubc)
        arrays = {"cld_frac": cf, "cloud_overlap_param": ovl, "icseed_sw": icseed,
                  "idx": (cols + 1).astype(np.int32)}
        s = lib.SWScheme(kd, lut, nlay, blksz=blksz)
    n = len(cols)
    s.set_launch_columns(24)   # several launches
    dev = {k: torch.from_numpy(v).cuda() for k, v in arrays.items()}
    words = torch.zeros((n, nlay, 4), dtype=torch.int32, device="cuda")
    s.mcica_mask_dev(ints, dev, words)
    torch.cuda.synchronize()
    got = _gpu_bits(words.cpu().numpy().view(np.uint32), ngpt)
    seeds = (ngpt + cols + 1) if isubc == 1 else icseed[cols]
    exp = O.mcica_mask(seeds.astype(np.int32), np.ascontiguousarray(cf[:, cols]), ngpt, iovr=iovr,
                       overlap=np.ascontiguousarray(ovl[:, cols]))
    exp = exp * _block_cloudy(cf[:, cols], blksz)[:, None, None].astype(np.uint8)
    assert exp[2].any() and exp[0, :, 0].any() or iovr in ("iovr_dcorr",), "the edge columns draw cloud"
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{len(bad)} mask bits differ, first (col, g, lay) {bad[:5].tolist()}"
    # the clear block and the clear column have no bits
    assert not got[3].any()
