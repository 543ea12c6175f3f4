"""compute_heating_rate (mo_heating_rates, as GFS_rrtmgp_post calls it at
GFS_rrtmgp_post.F90:11,159-171,249-266) on the GPU in FP64, through the
host-pointer C entry, the device entry and the Fortran module, against the
oracle's FP64 formula (pinned by a hand case in test_oracle_lw.py)."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _fluxes(ncol, nlay, seed):
    """Level fluxes and pressures with thin layers near the top (the
    cancellation the FP64 step is there for): p_lev (nlay+1, ncol) ==
    Fortran (ncol, nlay+1), surface first."""
    rng = np.random.default_rng(seed)
    p = np.exp(np.linspace(np.log(1.0e5), np.log(1.0), nlay + 1))[:, None] * rng.uniform(0.97, 1.03, ncol)
    up = 250.0 + 150.0 * rng.uniform(size=(nlay + 1, ncol))
    dn = 300.0 * rng.uniform(size=(nlay + 1, ncol))
    c = np.ascontiguousarray
    return c(up), c(dn), c(p)


def test_heating_rate_host_and_device_entries(kd_lw):
    from ccpprad import lib
    from oracle import oracle as O
    ncol, nlay = 300, 127
    up, dn, p = _fluxes(ncol, nlay, 3)
    ref = O.heating_rate(up, dn, p)
    hr = lib.heating_rate(ncol, nlay, up, dn, p)
    rel = np.abs(hr - ref) / np.maximum(np.abs(ref), 1e-300)
    print("host entry: rel max", rel.max(), "abs max K/day", np.abs(hr - ref).max() * 86400)
    assert rel.max() < 1e-13
    d = {k: torch.from_numpy(v).cuda() for k, v in (("up", up), ("dn", dn), ("p", p))}
    htr = torch.zeros((nlay, ncol), dtype=torch.float64, device="cuda")
    lib.heating_rate_dev(ncol, nlay, d["up"], d["dn"], d["p"], htr)
    torch.cuda.synchronize()
    assert np.array_equal(htr.cpu().numpy(), hr)


def test_heating_rate_entry_errors():
    from ccpprad import lib, capi
    up, dn, p = _fluxes(4, 3, 1)
    with pytest.raises(capi.CcppradError, match="dimensions"):
        lib.heating_rate(4, 0, up, dn, p)


def test_fortran_mo_heating_rates(tmp_path):
    """The Fortran mo_heating_rates%compute_heating_rate called the way
    GFS_rrtmgp_post_run calls it: LW over every column, SW over the daylit
    rows idxday(1:nDay) (a vector-subscripted section, copied in by the
    compiler) scattered back into zeroed arrays; a mis-shaped call returns a
    nonblank error string."""
    from ccpprad import build, lib
    from oracle import oracle as O
    drv = build.HR_DRIVER
    assert os.path.exists(drv), "hr_driver not built"
    ncol, nlay = 200, 127
    lup, ldn, p = _fluxes(ncol, nlay, 4)
    sup, sdn, _ = _fluxes(ncol, nlay, 5)
    rng = np.random.default_rng(6)
    idx = np.sort(rng.choice(ncol, 120, replace=False)).astype(np.int32) + 1
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([ncol, nlay, len(idx)], dtype=np.int32).tofile(f)
        for a in (lup, ldn, sup, sdn, p):
            a.tofile(f)
        idx.tofile(f)
    r = subprocess.run([drv, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = open(tmp_path / "out.bin", "rb").read()
    flags = np.frombuffer(raw[:12], dtype=np.int32)
    assert flags.tolist() == [0, 0, 1], flags
    vals = np.frombuffer(raw[12:], dtype=np.float64)
    htrlw = vals[:ncol * nlay].reshape(nlay, ncol)
    htrsw = vals[ncol * nlay:].reshape(nlay, ncol)
    assert np.array_equal(htrlw, lib.heating_rate(ncol, nlay, lup, ldn, p))
    ref_lw = O.heating_rate(lup, ldn, p)
    assert np.abs(htrlw - ref_lw).max() * 86400 < 1e-9
    cols = idx - 1
    c = np.ascontiguousarray
    ref_sw = np.zeros((nlay, ncol))
    ref_sw[:, cols] = O.heating_rate(c(sup[:, cols]), c(sdn[:, cols]), c(p[:, cols]))
    assert np.abs(htrsw - ref_sw).max() * 86400 < 1e-9
    day = np.zeros(ncol, dtype=bool)
    day[cols] = True
    assert not htrsw[:, ~day].any()
