"""What the trained LW gas-optics emulator gains over the table path.

The packaged network (ccpprad/data/mlp_lw_gas_optics.npz, trained on the GPU
by the library's own BF16 MFMA forward/backward) replaces the k-distribution
lookup in rrtmgp_lw_main_run.  On an all-sky workload of 16,384 columns x 127
layers this test checks, against the FP64 oracle on a 512-column subset:

  * the table path stays within the plain bar (HR 0.01 K/day max-abs),
  * the emulator path stays within HR RMSE 0.1 K/day,
  * the emulator's gas-optics stage (fused MLP inference plus the table
    fallback for out-of-hull layers), timed with HIP events on its stream,
    is faster than the table gas optics on the same columns.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NCOL, NLAY, NSUB = 16384, 127, 512


def _stage_ms(L, ctx, ids):
    import ctypes as C
    tot = 0.0
    for kid in ids:
        t, n = C.c_double(), C.c_int()
        L.ccpprad_timing_query(ctx, kid, C.byref(t), C.byref(n))
        tot += t.value
    return tot


def test_emulator_beats_table_within_hr_budget(kd_lw, lut_lw):
    from ccpprad import synthetic as syn, capi, lib, device, mlp
    from oracle import oracle as O
    assert os.path.exists(mlp.DEFAULT_WEIGHTS)
    pr = syn.make_profiles(NCOL, NLAY, seed=424242, clouds=True, kd=kd_lw)
    arr = capi.profile_fields(pr)
    ints = capi.lw_default_ints(NCOL, NLAY, blksz=64, nangle=3)
    L = lib.load()

    table = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=64)
    nn = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=64)
    emu = mlp.LWEmulator.load()
    emu.attach(nn)
    outs, stage = {}, {}
    for name, s, ids in (("table", table, (0, 6)), ("nn", nn, (0, 5, 6))):
        d = device.to_device(arr)
        d.update(device.lw_outputs(NCOL, NLAY))
        s.run_dev(ints, d)                      # warm-up and validation
        L.ccpprad_timing_enable(s.ctx, 1)
        for _ in range(3):
            s.run_dev(ints, d, validate=False)
        torch.cuda.synchronize()
        stage[name] = _stage_ms(L, s.ctx, ids) / 3
        L.ccpprad_timing_enable(s.ctx, 0)
        outs[name] = device.to_host(d)

    sub = pr.subset(np.arange(NSUB))
    sarr = capi.profile_fields(sub)
    ref = dict(sarr)
    ref.update(O.lw_outputs(NSUB, NLAY))
    O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(NSUB, NLAY, blksz=64, nangle=3), ref)
    href = ref["htrlw"] * 86400.0
    err = {k: outs[k]["htrlw"][:, :NSUB] * 86400.0 - href for k in outs}
    rmse = {k: float(np.sqrt(np.mean(v * v))) for k, v in err.items()}
    mx = {k: float(np.abs(v).max()) for k, v in err.items()}
    print("gas-optics stage ms (table, nn):", stage["table"], stage["nn"])
    print("HR RMSE K/day:", rmse, "max:", mx)
    assert mx["table"] < 0.01
    assert rmse["nn"] <= 0.1
    assert stage["nn"] < stage["table"]
    emu.close()
