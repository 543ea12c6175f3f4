"""Concurrent scheme calls on one context.

The reference `_run` keeps no mutable module state, so a host may call it
from several OpenMP threads at once on different column blocks.  The library
keeps one workspace per stream (and a private stream per host thread for the
host-array entry), so concurrent calls must give exactly the results of the
same calls made one after another.
"""
import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NCOL, NLAY, NTHREADS = 256, 127, 4


def _inputs(kd, seed):
    from ccpprad import synthetic as syn, capi
    pr = syn.make_profiles(NCOL, NLAY, seed=seed, clouds=True, kd=kd)
    return capi.profile_fields(pr)


def test_host_entry_threads_match_sequential(kd_lw, lut_lw):
    from ccpprad import lib, capi
    from oracle import oracle as O
    s = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=64)
    ints = capi.lw_default_ints(NCOL, NLAY, blksz=64, nangle=3)
    jobs = []
    for t in range(NTHREADS):
        arr = _inputs(kd_lw, 5000 + t)
        seq = dict(arr)
        seq.update(O.lw_outputs(NCOL, NLAY))
        s.run(ints, seq)
        par = dict(arr)
        par.update(O.lw_outputs(NCOL, NLAY))
        jobs.append((seq, par))
    errors = []

    def work(par):
        try:
            for _ in range(3):
                s.run(ints, par)
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(p,)) for _, p in jobs]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for seq, par in jobs:
        for n in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "htrlw"):
            assert np.array_equal(seq[n], par[n]), n


def test_device_entry_streams_match_sequential(kd_sw, lut_sw):
    from ccpprad import lib, capi, device
    s = lib.SWScheme(kd_sw, lut_sw, NLAY, blksz=64)
    ints = capi.sw_default_ints(NCOL, NCOL, NLAY, blksz=64)
    jobs = []
    for t in range(NTHREADS):
        arr = _inputs(kd_sw, 6000 + t)
        arr["idx"] = np.arange(1, NCOL + 1, dtype=np.int32)
        seq = device.to_device(arr)
        seq.update(device.sw_outputs(NCOL, NLAY))
        s.run_dev(ints, seq)
        torch.cuda.synchronize()
        par = device.to_device(arr)
        par.update(device.sw_outputs(NCOL, NLAY))
        jobs.append((device.to_host(seq), par))
    streams = [torch.cuda.Stream() for _ in range(NTHREADS)]
    torch.cuda.synchronize()   # inputs and zeroed outputs are made on the default stream
    errors = []

    def work(par, st):
        try:
            for _ in range(3):
                s.run_dev(ints, par, stream=st.cuda_stream, validate=False)
            st.synchronize()
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(p, st)) for (_, p), st in zip(jobs, streams)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for seq, par in jobs:
        out = device.to_host(par)
        for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "htrsw", "scmpsw"):
            assert np.array_equal(seq[n], out[n]), n
