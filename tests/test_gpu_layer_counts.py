"""LW and SW parity against the FP64 oracle at layer counts other than 127.

The kernels walk the column in pieces: the gas optics in 16-layer chunks, the
LW solver in 8-layer batches, the SW solver and the LW two-stream solver in
8-layer segments with checkpoints between them.  These counts put the
remainders at every edge: 2 layers (one piece, no checkpoint), 9 and 17 (one
layer past a piece), 65 (the GFS L65 grid), 200 and 260 (long columns, the
SW solver's LDS near its 64 KiB).  Both vertical orientations; clouds with
precipitation-free McICA max-random overlap; the plain-path bar (fluxes
0.05 W/m2, heating rate 0.01 K/day max-abs).
"""
import numpy as np
import pytest

from util import report

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FLUX_TOL = 0.05
HR_TOL = 0.01
NCOL = 32
NLAYS = [2, 9, 17, 65, 200, 260]


@pytest.mark.parametrize("nlay", NLAYS)
@pytest.mark.parametrize("top_at_1", [False, True])
def test_lw_layer_counts(kd_lw, lut_lw, nlay, top_at_1):
    from test_gpu_lw import _run_both
    ref, out = _run_both(kd_lw, lut_lw, NCOL, nlay, clouds=True, nangle=3, clrsky=True, jac=False,
                         top_at_1=top_at_1)
    for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky"):
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < FLUX_TOL, n
    for n in ("htrlw", "htrlwc"):
        mx, _ = report(n, out[n] * 86400, ref[n] * 86400, "K/day")
        assert mx < HR_TOL, n


@pytest.mark.parametrize("nlay", NLAYS)
def test_lw_scattering_layer_counts(kd_lw, lut_lw, nlay):
    from test_gpu_lw import _run_both
    ref, out = _run_both(kd_lw, lut_lw, NCOL, nlay, clouds=True, nangle=1, clrsky=False, jac=False, scat=True)
    for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky"):
        mx, _ = report(n, out[n], ref[n], "W/m2")
        assert mx < FLUX_TOL, n
    mx, _ = report("htrlw", out["htrlw"] * 86400, ref["htrlw"] * 86400, "K/day")
    assert mx < HR_TOL


@pytest.mark.parametrize("nlay", NLAYS)
@pytest.mark.parametrize("top_at_1", [False, True])
def test_sw_layer_counts(kd_sw, lut_sw, nlay, top_at_1):
    from test_gpu_sw import _run_both
    ref, out, cols = _run_both(kd_sw, lut_sw, NCOL, nlay, clouds=True, clrsky=True, top_at_1=top_at_1)
    for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky"):
        mx, _ = report(n, out[n][:, cols], ref[n][:, cols], "W/m2")
        assert mx < FLUX_TOL, n
    for n in ("htrsw", "htrswc"):
        mx, _ = report(n, out[n] * 86400, ref[n] * 86400, "K/day")
        assert mx < HR_TOL, n
    mx, _ = report("scmpsw", out["scmpsw"], ref["scmpsw"], "W/m2")
    assert mx < FLUX_TOL
