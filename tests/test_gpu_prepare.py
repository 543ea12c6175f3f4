"""GPU parity of the input preparation (GFS_rrtmgp_pre's bounds, cmp_tlev and
mixing ratios) against the FP64 oracle, through the C-ABI.  Everything is FP64
and bit-identical except t_lev, whose logarithms may differ from the host's
in the last bit (relative 1e-13)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("top", [False, True])
def test_prepare_inputs_matches_oracle(kd_lw, lut_lw, top):
    from ccpprad import lib
    from oracle import oracle as O
    from test_oracle_prepare import prep_inputs
    nl = 64
    x, _ = prep_inputs(kd_lw, ncol=300, nlay=nl, top=top, seed=5)
    ref, t1 = O.prepare(kd_lw, x["prsi"], x["prsl"], x["tgrs"], x["q"], x["o3"], x["tsfc"])
    nc = x["prsl"].shape[1]
    s = lib.LWScheme(kd_lw, lut_lw, nl)
    d = {k: torch.from_numpy(v).cuda() for k, v in x.items()}
    out = {k: torch.full(v.shape, -1.0, dtype=torch.float64, device="cuda") for k, v in ref.items()}
    flag = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    s.prepare_inputs_dev(nc, nl, d["prsi"], d["prsl"], d["tgrs"], d["q"], d["o3"], d["tsfc"], out, top_at_1=flag)
    torch.cuda.synchronize()
    assert flag.item() == int(top) == int(t1)
    for k, r in ref.items():
        g = out[k].cpu().numpy()
        if k == "t_lev":
            assert np.allclose(g, r, rtol=1e-13, atol=0), k
        else:
            assert np.array_equal(g, r), k


@pytest.mark.parametrize("nday", [1, 0])
def test_post_matches_oracle(nday):
    """GFS_rrtmgp_post's bookkeeping: the copies bit-identical; the fluxr
    sums to 1e-15 relative (the host and device compilers may contract
    different products into FMAs, and exp(-tau) may differ in the last
    bit)."""
    from ccpprad import lib
    from oracle import oracle as O
    from test_oracle_prepare import post_inputs
    ints, reals, arr, out = post_inputs(nc=1000, nl=64, nday=nday, seed=6)
    ref = {k: v.copy() for k, v in out.items()}
    a = dict(arr)
    a.update(ref)
    O.post(ints, reals, a)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in arr.items()}
    dout = {k: torch.from_numpy(v.copy()).cuda() for k, v in out.items()}
    d.update(dout)
    lib.post_dev(ints, reals, d)
    torch.cuda.synchronize()
    for k, r in ref.items():
        g = dout[k].cpu().numpy()
        if k in ("fluxr", "nirbmui", "nirdfui", "visbmui", "visdfui", "sfcnsw"):
            assert np.allclose(g, r, rtol=1e-15, atol=0), k
        else:
            assert np.array_equal(g, r), k
