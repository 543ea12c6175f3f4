"""The allocation-free LW McICA mask entry against the allocating one, bit
for bit (integer work), on two calls in a row with the second smaller (the
library's fixed block-flag array keeps the first call's flags, which the
second must not read), and its size limit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _case(ncol, nlay, seed):
    rng = np.random.default_rng(seed)
    cf = np.where(rng.uniform(size=(nlay, ncol)) < 0.3, rng.uniform(0.0, 1.0, (nlay, ncol)), 0.0)
    cf[:, 8:16] = 0.0   # a clear block (blocks of 8)
    ovl = rng.uniform(0.0, 1.0, (nlay, ncol))
    return {"cld_frac": torch.from_numpy(np.ascontiguousarray(cf)).cuda(),
            "cloud_overlap_param": torch.from_numpy(np.ascontiguousarray(ovl)).cuda()}


@pytest.mark.parametrize("iovr", ["iovr_maxrand", "iovr_exprand", "iovr_dcorr"])
def test_noalloc_mask_equals_allocating_entry(kd_lw, lut_lw, iovr):
    from ccpprad import lib, capi
    nlay, blk = 127, 8
    s = lib.LWScheme(kd_lw, lut_lw, nlay, blksz=blk)
    s.set_launch_columns(40)   # several launches per call
    for ncol, seed in ((256, 1), (96, 2)):   # the second call smaller
        arrays = _case(ncol, nlay, seed)
        ints = capi.lw_default_ints(ncol, nlay, blksz=blk, iovr=iovr)
        ref = torch.zeros((ncol, nlay, 4), dtype=torch.int32, device="cuda")
        got = torch.full((ncol, nlay, 4), -1, dtype=torch.int32, device="cuda")
        s.mcica_mask_dev(ints, arrays, ref)
        s.mcica_mask_noalloc_dev(ints, arrays, got)
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert r.any(), "the case draws cloud"
        assert not g[8:16].any(), "the clear block has no bits"
        assert np.array_equal(r, g), f"ncol={ncol}: {int((r != g).sum())} words differ"


def test_noalloc_mask_rejects_too_many_blocks(kd_lw, lut_lw):
    from ccpprad import lib, capi
    nlay = 127
    s = lib.LWScheme(kd_lw, lut_lw, nlay, blksz=1)
    ncol = (1 << 20) + 1
    ints = capi.lw_default_ints(ncol, nlay, blksz=1)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")   # never read: the size check comes first
    mask = torch.zeros(8, dtype=torch.int32, device="cuda")
    with pytest.raises(capi.CcppradError, match="allocation-free limit"):
        s.mcica_mask_noalloc_dev(ints, {"cld_frac": dummy}, mask)
