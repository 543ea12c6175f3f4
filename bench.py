#!/usr/bin/env python3
"""Throughput of the RRTMGP LW+SW radiation hot path on one MI355X.

A step is one rrtmgp_lw_main_run (all sky, 3 Gauss angles, 128 g-points)
plus one rrtmgp_sw_main_run (all sky, 112 g-points) over 262,144 columns x
127 layers, with every input already resident in HBM (device-pointer C-ABI
entries).  value = columns per second of the LW+SW pair.  The two schemes are
independent and run on two HIP streams, as a GPU-resident host would issue
them; per-kernel times come from a separate step that runs them in turn.

The JSON line also carries:
  roofline     : the dominant kernel's algorithmic bytes / its mean launch
                 time (HIP events on the launch stream) against 8 TB/s HBM
  cpu_baseline : the FP64 oracle timed on the host cores over a bounded
                 sample of the same workload, and the flux / heating-rate
                 error of the GPU result on that sample against it; also
                 on one core, and its FP32 build on all cores (with that
                 build's error against the FP64 run)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ccpp-physics_amd"))

HBM_PEAK_GBS = 8000.0
NCOL = 262_144
NLAY = 127
UNIQUE = 32_768        # distinct synthetic columns, tiled to NCOL on the device
METRIC = "LW+SW columns/s at 127 layers (table vs NN) @ HR RMSE<=0.1 K/day vs FP64"
# HBM bytes per 65,536-column launch measured with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE (separate runs of `bench.py --steps 1 --warmup 0 --no-cpu
# --no-nn --no-e2e --no-callers --copy-probe`), each scaled by the reading of a
# 1 GiB device copy (FETCH x 2.0, WRITE x 1.0).  Counters cannot be read from inside this
# process, so the last measurement of the current kernels is recorded here.
TRAFFIC_GB = {"sw_sw_solve": 19.366, "lw_lw_solve": 22.651, "lw_gas_optics": 9.131, "sw_gas_optics": 8.070,
              "lw_cloud_optics": 0.757, "sw_cloud_optics": 0.757, "lw_mcica": 0.670, "sw_mcica": 0.670,
              "lw_solve_2str": 24.285}
# sha256 (first 16 hex digits) of the kernel sources the figures above were
# measured with (kernel_source_sha()); on any other sources the bench line
# reports traffic null rather than a stale figure.
TRAFFIC_SRC_SHA = "d3d68faafdaccef1"


# Algorithmic bytes per column of each stage: the compulsory I/O of the
# reference's own data flow (every array it materialises, counted once; FP32
# optics, FP64 fluxes / heating rate), at 127 layers, LW 128 g / 16 bands, SW
# 112 g / 14 bands.  The solvers include the fused heating-rate stage
# (3,072 B in, 1,016 B out).
ALG_BYTES_PER_COLUMN = {
    "lw_lw_solve": 269_308 + 2_048 + 3_072 + 1_016,
    "sw_sw_solve": 193_000 + 3_120 + 3_072 + 1_016,
    "lw_gas_optics": 4_584 + 261_120,   # gas-optics inputs in, tau + Planck sources out
    "sw_gas_optics": 4_576 + 171_136,   # inputs in, tau_abs + tau_Rayleigh out
    "lw_cloud_optics": 2_540 + 8_128,
    "sw_cloud_optics": 2_540 + 21_336,
}


def kernel_source_sha():
    """sha256 over the HIP sources and headers of the library, in name order."""
    import hashlib
    d = os.path.join(REPO, "ccpp-physics_amd", "csrc")
    h = hashlib.sha256()
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".h")):
            h.update(f.encode())
            with open(os.path.join(d, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def file_sha(path):
    import hashlib
    with open(path, "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()[:16]


def host_info():
    """The host cores the CPU baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ncol", type=int, default=NCOL)
    p.add_argument("--cpu-sample", type=int, default=32768)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--mlp-weights", default=None,
                   help="trained emulator weights (.npz); default: the packaged ones")
    p.add_argument("--no-nn", action="store_true", help="skip the learned LW gas-optics leg")
    p.add_argument("--no-lwscat", action="store_true", help="skip the LW cloud-scattering (doGP_lwscat) leg")
    p.add_argument("--no-aer", action="store_true", help="skip the leg with aerosol optics")
    p.add_argument("--launch-columns", type=int, default=65536,
                   help="columns per kernel launch (the library's default and its maximum is 65,536)")
    p.add_argument("--no-callers", action="store_true",
                   help="skip the leg timing the device callers (pre, geometry, overlap, post)")
    p.add_argument("--no-e2e", action="store_true",
                   help="skip the end-to-end leg through the host-pointer entries (host<->device copies included)")
    p.add_argument("--e2e-steps", type=int, default=2)
    p.add_argument("--cpu-sample-1core", type=int, default=1024,
                   help="columns of the single-core CPU baseline sample")
    p.add_argument("--copy-probe", action="store_true",
                   help="after the bench, run one 1 GiB device copy: the reference dispatch that scales "
                        "FETCH_SIZE / WRITE_SIZE counter readings to bytes")
    return p.parse_args()


def build_inputs(kd_lw, kd_sw, ncol, torch):
    from ccpprad import synthetic as syn, capi, device
    uniq = min(UNIQUE, ncol)
    pr = syn.make_profiles(uniq, NLAY, seed=20261203, clouds=True, kd=kd_lw)
    arr = capi.profile_fields(pr)
    reps = (ncol + uniq - 1) // uniq
    dev = {}
    for k, v in arr.items():
        t = torch.from_numpy(np.ascontiguousarray(v)).cuda()
        if t.dim() == 1:
            t = t.repeat(reps)[:ncol].contiguous()
        else:
            t = t.repeat(1, reps)[:, :ncol].contiguous()
        dev[k] = t
    dev["idx"] = torch.arange(1, ncol + 1, dtype=torch.int32, device="cuda")
    return pr, arr, dev


def time_callers(a, torch, lw, dev, ncol):
    """The callers either side of the schemes on the device, at the workload
    size: GFS_rrtmgp_pre's input preparation and layer geometry,
    GFS_rrtmgp_cloud_overlap_run (Hogan length, exponential-random overlap,
    convective overlap on) and GFS_rrtmgp_post's bookkeeping with the
    diagnostics on.  HIP events on the launch stream, mean over the timed
    repeats.  Algorithmic bytes per column: prepare 8 (9 nl + 3 nv + 1),
    geometry 8 (7 nl + nv), overlap 8 (6 nl + 2) (every field read and
    written once); post is reported as time only."""
    from ccpprad import lib
    nl, nv = NLAY, NLAY + 1
    sp = torch.cuda.current_stream().cuda_stream
    f64 = dict(dtype=torch.float64, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(11)
    rnd = lambda *s, lo=0.0, hi=1.0: torch.rand(*s, generator=gen, **f64) * (hi - lo) + lo
    q = rnd(nl, ncol, hi=0.02)
    o3 = rnd(nl, ncol, hi=1e-5)
    prep = {k: torch.empty(nv if k in ("p_lev", "t_lev") else nl, ncol, **f64)
            for k in ("p_lay", "p_lev", "t_lay", "t_lev", "q_lay", "vmr_h2o", "vmr_o3")}
    geo = {k: torch.empty(nl, ncol, **f64) for k in ("tv", "dz", "dzc", "dp")}
    ovl_in = dict(lat=rnd(ncol, lo=-1.5, hi=1.5), deltaZc=geo["dzc"], cld_frac=dev["cld_frac"],
                  cld_cnv_frac=rnd(nl, ncol, hi=0.2))
    ovl_out = dict(de_lgth=torch.empty(ncol, **f64), cloud_overlap_param=torch.empty(nl, ncol, **f64),
                   cnv_cloud_overlap_param=torch.empty(nl, ncol, **f64),
                   precip_overlap_param=torch.empty(nl, ncol, **f64))
    ovl_ints = dict(nCol=ncol, nLev=nl, yearlen=365, doSWrad=1, doLWrad=1, literal_r8=1, idcor=1, idcor_con=0,
                    idcor_hogan=1, idcor_oreopoulos=2, iovr=5, iovr_convcld=5, iovr_dcorr=3, iovr_exp=4,
                    iovr_exprand=5, imfdeepcnv=2, imfdeepcnv_gf=3, imfdeepcnv_samf=2)
    ovl_reals = dict(julian=172.0, con_pi=np.pi, dcorr_con=2.5)
    post_in = {k: rnd(nv, ncol, hi=400.0) for k in ("fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky",
                                                    "fluxlwDOWN_clrsky", "fluxswUP_allsky", "fluxswDOWN_allsky",
                                                    "fluxswUP_clrsky", "fluxswDOWN_clrsky")}
    post_in.update(scmpsw=rnd(ncol, 6, hi=300.0), tsfa=rnd(ncol, lo=250.0, hi=300.0), coszen=rnd(ncol),
                   coszdg=rnd(ncol), htrlw=rnd(nl, ncol, hi=1e-5), p_lev=dev["p_lev"], cldsa=rnd(5, ncol),
                   cldtaulw=rnd(nl, ncol, hi=2.0), cldtausw=rnd(nl, ncol, hi=5.0), aerodp=rnd(6, ncol),
                   mbota=torch.full((3, ncol), 10, dtype=torch.int32, device="cuda"),
                   mtopa=torch.full((3, ncol), 40, dtype=torch.int32, device="cuda"))
    for k in ("sfc_alb_nir_dir", "sfc_alb_nir_dif", "sfc_alb_uvvis_dir", "sfc_alb_uvvis_dif"):
        post_in[k] = rnd(ncol, lo=0.05, hi=0.8)
    post_out = dict(topflw=torch.zeros(ncol, 2, **f64), sfcflw=torch.zeros(ncol, 4, **f64),
                    topfsw=torch.zeros(ncol, 3, **f64), sfcfsw=torch.zeros(ncol, 4, **f64),
                    htrlwu=torch.zeros(nl, ncol, **f64), fluxr=torch.zeros(45, ncol, **f64))
    for k in ("sfcdlw", "sfculw", "tsflw", "nirbmdi", "nirdfdi", "visbmdi", "visdfdi", "nirbmui", "nirdfui",
              "visbmui", "visdfui", "sfcnsw", "sfcdsw"):
        post_out[k] = torch.zeros(ncol, **f64)
    post_ints = dict(nCol=ncol, nLev=nl, nDay=ncol, iSFC=1, iTOA=nv, doLWrad=1, doSWrad=1, save_diag=1,
                     naerodp=6, nfxr=45)
    post_reals = dict(fhlwr=3600.0, fhswr=3600.0, raddt=1800.0)
    legs = {
        "prepare_inputs": (lambda: lw.prepare_inputs_dev(ncol, nl, dev["p_lev"], dev["p_lay"], dev["t_lay"], q, o3,
                                                         dev["tsfg"], prep, stream=sp), 8 * (9 * nl + 3 * nv + 1)),
        "layer_geometry": (lambda: lib.layer_geometry_dev(ncol, nl, prep["p_lev"], prep["p_lay"], prep["t_lay"],
                                                          prep["q_lay"], geo["tv"], geo["dz"], geo["dzc"], geo["dp"],
                                                          stream=sp), 8 * (7 * nl + nv)),
        "cloud_overlap": (lambda: lib.cloud_overlap_dev(ovl_ints, ovl_reals, {**ovl_in, **ovl_out}, stream=sp),
                          8 * (6 * nl + 2)),
        "post": (lambda: lib.post_dev(post_ints, post_reals, {**post_in, **post_out}, stream=sp), None),
    }
    res = {}
    reps = max(a.steps, 3)
    for name, (fn, bpc) in legs.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        r = {"ms": ms}
        if bpc is not None:
            gbs = bpc * ncol / (ms * 1e-3) / 1e9
            r.update(algorithmic_bytes_per_column=bpc, achieved_GBs=gbs, frac_of_hbm_peak=gbs / HBM_PEAK_GBS)
        res[name] = r
    print("# callers: " + ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in res.items()), file=sys.stderr, flush=True)
    return {"columns": ncol, "layers": nl, "kernels": res,
            "note": "thread per column (prepare, geometry, post) or per (column, layer) (overlap), FP64"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the path shards by columns across GPUs "
                                   "with no exchange, and a multi-GPU run is not launched", "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from ccpprad import synthetic as syn, capi, lib, device

    lib.load()
    lib_path = lib.loaded_path()
    kd_lw, kd_sw = syn.make_kdist(True), syn.make_kdist(False)
    lut_lw, lut_sw = syn.make_cloud_lut(True), syn.make_cloud_lut(False)
    ncol = a.ncol
    blksz = 64
    pr, arr, dev = build_inputs(kd_lw, kd_sw, ncol, torch)
    lw = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=blksz)
    sw = lib.SWScheme(kd_sw, lut_sw, NLAY, blksz=blksz)
    launch = min(lw.set_launch_columns(a.launch_columns), sw.set_launch_columns(a.launch_columns))
    lw_out = device.lw_outputs(ncol, NLAY)
    sw_out = device.sw_outputs(ncol, NLAY)
    lw_args = dict(dev)
    lw_args.update(lw_out)
    sw_args = dict(dev)
    sw_args.update(sw_out)
    lw_ints = capi.lw_default_ints(ncol, NLAY, blksz=blksz, nangle=3)
    sw_ints = capi.sw_default_ints(ncol, ncol, NLAY, blksz=blksz)
    # LW and SW are independent: each scheme gets its own stream (the library
    # keeps one workspace per stream) so their kernels overlap on the GPU.
    # The per-kernel timing step below runs them one after another.
    stream, stream_sw = torch.cuda.Stream(), torch.cuda.Stream()
    sp, sp_sw = stream.cuda_stream, stream_sw.cuda_stream

    def step(validate=False, serial=False):
        lw.run_dev(lw_ints, lw_args, stream=sp, validate=validate)
        sw.run_dev(sw_ints, sw_args, stream=sp if serial else sp_sw, validate=validate)

    print(f"# setup done: {ncol} columns x {NLAY} layers", file=sys.stderr, flush=True)
    step(validate=True)
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / a.steps
    value = ncol / (ms / 1e3)
    print(f"# timed {a.steps} steps: {ms:.3f} ms/step", file=sys.stderr, flush=True)

    # ---- per-kernel times with HIP events on the launch stream (one extra step)
    L = lib.load()
    for s in (lw, sw):
        L.ccpprad_timing_enable(s.ctx, 1)
    step(serial=True)
    kt = {}
    names = {0: "gas_optics", 1: "lw_solve", 2: "sw_solve", 3: "cloud_optics", 4: "mcica"}
    for tag, s in (("lw", lw), ("sw", sw)):
        for kid, nm in names.items():
            tot, cnt = C.c_double(), C.c_int()
            L.ccpprad_timing_query(s.ctx, kid, C.byref(tot), C.byref(cnt))
            if cnt.value:
                kt[f"{tag}_{nm}"] = (tot.value, cnt.value)
        L.ccpprad_timing_enable(s.ctx, 0)
    chunk = min(ncol, launch)   # columns per launch
    per_col = ALG_BYTES_PER_COLUMN
    src_sha = kernel_source_sha()
    dom = max(kt.items(), key=lambda kv: kv[1][0])
    dname, (dtot, dcnt) = dom
    avg_ms = dtot / dcnt
    bytes_launch = per_col.get(dname, 0) * chunk
    achieved = bytes_launch / (avg_ms / 1e3) / 1e9 if bytes_launch else None
    roof = {"kernel": dname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": (TRAFFIC_GB[dname] * 1e9 if dname in TRAFFIC_GB and chunk == 65536 and src_sha == TRAFFIC_SRC_SHA
                        else None),
            "traffic_measured_on_sources": TRAFFIC_SRC_SHA, "kernel_sources": src_sha,
            "traffic_unit": "bytes per launch",
            "avg_launch_ms": avg_ms, "bytes_per_column": per_col.get(dname), "bytes_per_launch": bytes_launch,
            "columns_per_launch": chunk,
            "kernel_ms_per_step": {k: v[0] for k, v in kt.items()}}

    # ---- learned LW gas optics: the same step with the MLP emulator in place
    # of the LW table lookup (out-of-hull layers fall back to the table)
    nn = None
    nn_out = None
    if not a.no_nn:
        from ccpprad import mlp
        wpath = a.mlp_weights or mlp.DEFAULT_WEIGHTS
        if os.path.exists(wpath):
            emu = mlp.LWEmulator.load(wpath)
            lw_nn = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=blksz)
            lw_nn.set_launch_columns(launch)
            emu.attach(lw_nn)
            nn_out = device.lw_outputs(ncol, NLAY)
            nn_args = dict(dev)
            nn_args.update(nn_out)

            def step_nn(validate=False):
                lw_nn.run_dev(lw_ints, nn_args, stream=sp, validate=validate)
                sw.run_dev(sw_ints, sw_args, stream=sp_sw, validate=validate)

            step_nn(validate=True)
            for _ in range(a.warmup):
                step_nn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(a.steps):
                step_nn()
            torch.cuda.synchronize()
            ms_nn = (time.perf_counter() - t0) * 1e3 / a.steps
            # LW alone, table vs NN
            def lw_only(s, args):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(a.steps):
                    s.run_dev(lw_ints, args, stream=sp, validate=False)
                torch.cuda.synchronize()
                return (time.perf_counter() - t0) * 1e3 / a.steps
            ms_lw_table, ms_lw_nn = lw_only(lw, lw_args), lw_only(lw_nn, nn_args)
            L.ccpprad_timing_enable(lw_nn.ctx, 1)
            lw_nn.run_dev(lw_ints, nn_args, stream=sp, validate=False)
            nkt = {}
            for kid, nm in {6: "gas_optics_fallback", 1: "lw_solve", 3: "cloud_optics", 4: "mcica",
                            5: "mlp_infer"}.items():
                tot, cnt = C.c_double(), C.c_int()
                L.ccpprad_timing_query(lw_nn.ctx, kid, C.byref(tot), C.byref(cnt))
                if cnt.value:
                    nkt[nm] = (tot.value, cnt.value)
            L.ccpprad_timing_enable(lw_nn.ctx, 0)
            # MFMA roofline of the fused inference kernel: 2 x (14x128 + 128x128
            # + 128x128 + 128x256) FLOP per layer row (layer 1 takes the inputs
            # as hi/lo BF16 pairs, K = 14)
            flop_row = 2 * (14 * 128 + 128 * 128 + 128 * 128 + 128 * 256)
            mt, mc = nkt.get("mlp_infer", (0.0, 0))
            mlp_avg = mt / mc if mc else None
            mlp_tf = flop_row * chunk * NLAY / (mlp_avg / 1e3) / 1e12 if mlp_avg else None
            # rows outside the training hull (table fallback), from the
            # features of the distinct columns
            F = mlp.features(pr)
            lo, hi = emu.norm["lo"].astype(np.float64), emu.norm["hi"].astype(np.float64)
            out_row = ((F < lo) | (F > hi)).any(axis=1)
            nn = {"value": ncol / (ms_nn / 1e3), "unit": "columns/s", "ms_per_step": ms_nn,
                  "out_of_hull": {"row_fraction": float(out_row.mean()),
                                  "column_fraction": float(out_row.reshape(-1, NLAY).any(axis=1).mean())},
                  "lw_only_columns_per_s": {"table": ncol / (ms_lw_table / 1e3), "nn": ncol / (ms_lw_nn / 1e3)},
                  "roofline": {"kernel": "mlp_infer", "bound": "mfma", "achieved": mlp_tf, "peak": 2500.0,
                               "unit": "TFLOP/s", "frac": (mlp_tf / 2500.0) if mlp_tf else None,
                               "avg_launch_ms": mlp_avg, "flop_per_launch": flop_row * chunk * NLAY},
                  "kernel_ms_per_step": {k: v[0] for k, v in nkt.items()},
                  "weights": os.path.relpath(wpath, REPO)}
            emu.close()

    # ---- LW with cloud scattering (doGP_lwscat): the LW scheme alone with the
    # two-stream adding solver in place of the no-scattering one
    lws = None
    lws_out = None
    if not a.no_lwscat:
        lw_s = lib.LWScheme(kd_lw, lut_lw, NLAY, blksz=blksz)
        lw_s.set_launch_columns(launch)
        lws_out = device.lw_outputs(ncol, NLAY)
        s_args = dict(dev)
        s_args.update(lws_out)
        s_ints = capi.lw_default_ints(ncol, NLAY, blksz=blksz, nangle=3, scat=True)
        lw_s.run_dev(s_ints, s_args, stream=sp, validate=True)

        def lw_alone(s, ints, args):
            for _ in range(a.warmup):
                s.run_dev(ints, args, stream=sp, validate=False)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(a.steps):
                s.run_dev(ints, args, stream=sp, validate=False)
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) * 1e3 / a.steps

        ms_scat, ms_noscat = lw_alone(lw_s, s_ints, s_args), lw_alone(lw, lw_ints, lw_args)
        L.ccpprad_timing_enable(lw_s.ctx, 1)
        lw_s.run_dev(s_ints, s_args, stream=sp, validate=False)
        skt = {}
        for kid, nm in {0: "gas_optics", 1: "lw_solve_2str", 3: "cloud_optics", 4: "mcica"}.items():
            tot, cnt = C.c_double(), C.c_int()
            L.ccpprad_timing_query(lw_s.ctx, kid, C.byref(tot), C.byref(cnt))
            if cnt.value:
                skt[nm] = tot.value
        L.ccpprad_timing_enable(lw_s.ctx, 0)
        lws = {"lw_only_columns_per_s": ncol / (ms_scat / 1e3), "ms_per_step": ms_scat,
               "lw_only_noscat_columns_per_s": ncol / (ms_noscat / 1e3),
               "kernel_ms_per_step": skt,
               "solver": "two-stream adding with cloud scattering (all sky), no-scattering clear sky off"}

    # ---- the same step with aerosol optics, which GFS hands both schemes on
    # every call (per band tau, ssa, g of every column, FP64, generated on
    # the device); the library packs them per launch before the solvers
    aer = None
    if not a.no_aer:
        gen = torch.Generator(device="cuda")
        gen.manual_seed(20261204)

        def afields(pre, nb):
            shp = (nb, NLAY, ncol)

            def u(lo, hi):
                return torch.rand(shp, generator=gen, dtype=torch.float64, device="cuda") * (hi - lo) + lo
            return {pre + "tau": torch.pow(10.0, u(-5.0, -1.0)), pre + "ssa": u(0.8, 1.0), pre + "g": u(0.5, 0.8)}
        # own outputs: the table-path results are checked against the oracle below
        lwa = dict(dev)
        lwa.update(device.lw_outputs(ncol, NLAY))
        lwa.update(afields("aerlw_", kd_lw.nband))
        swa = dict(dev)
        swa.update(device.sw_outputs(ncol, NLAY))
        swa.update(afields("aersw_", kd_sw.nband))

        def step_aer(validate=False, serial=False):
            lw.run_dev(lw_ints, lwa, stream=sp, validate=validate)
            sw.run_dev(sw_ints, swa, stream=sp if serial else sp_sw, validate=validate)

        step_aer(validate=True)
        for _ in range(a.warmup):
            step_aer()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step_aer()
        torch.cuda.synchronize()
        ms_aer = (time.perf_counter() - t0) * 1e3 / a.steps
        for s in (lw, sw):
            L.ccpprad_timing_enable(s.ctx, 1)
        step_aer(serial=True)
        akt = {}
        for tag, s in (("lw", lw), ("sw", sw)):
            for kid, nm in {**names, 7: "aer_pack"}.items():
                tot, cnt = C.c_double(), C.c_int()
                L.ccpprad_timing_query(s.ctx, kid, C.byref(tot), C.byref(cnt))
                if cnt.value:
                    akt[f"{tag}_{nm}"] = tot.value
            L.ccpprad_timing_enable(s.ctx, 0)
        aer = {"value": ncol / (ms_aer / 1e3), "unit": "columns/s", "ms_per_step": ms_aer,
               "kernel_ms_per_step": akt,
               "aerosols": "LW and SW, every column and band: tau 1e-5..1e-1 log-uniform, ssa 0.8..1, g 0.5..0.8"}
        print(f"# with aerosols: {ms_aer:.3f} ms/step", file=sys.stderr, flush=True)
        del lwa, swa

    callers = None if a.no_callers else time_callers(a, torch, lw, dev, ncol)

    # ---- end to end through the host-pointer entries (the ones the Fortran
    # scheme modules call): host arrays in pageable memory, as a Fortran
    # host holds them; every input copied in and every output copied out
    # each call (intent(inout) outputs are copied both ways, except those the
    # scheme writes in every column), synchronous.
    # LW then SW, one after the other.
    e2e = None
    if not a.no_e2e:
        reps = (ncol + UNIQUE - 1) // UNIQUE
        host = {}
        for k, v in arr.items():
            host[k] = np.ascontiguousarray(np.tile(v, reps)[:ncol] if v.ndim == 1 else np.tile(v, (1, reps))[:, :ncol])
        host["idx"] = np.arange(1, ncol + 1, dtype=np.int32)
        hl = dict(host)
        hl.update({k: np.zeros(tuple(v.shape), dtype=np.float64) for k, v in lw_out.items()})
        hs = dict(host)
        hs.update({k: np.zeros(tuple(v.shape), dtype=np.float64) for k, v in sw_out.items()})
        # bytes the host entries copy: each input of the scheme's argument
        # list once, each output back, and in as well unless the scheme writes
        # it in every column (the LW fluxes and heating rate)
        lw_whole = {"fluxlwUP_allsky", "fluxlwDOWN_allsky", "fluxlwUP_clrsky", "fluxlwDOWN_clrsky",
                    "fluxlwUP_radtime", "fluxlwDOWN_radtime", "htrlw"}
        moved = 0
        for d_, ins, outs, whole in ((hl, set(capi.LW_INPUT_FIELDS) | {"icseed_lw"}, set(lw_out), lw_whole),
                                     (hs, set(capi.SW_INPUT_FIELDS) | {"idx", "icseed_sw"}, set(sw_out), set())):
            for k, v in d_.items():
                if not isinstance(v, np.ndarray):
                    continue
                if k in outs:
                    moved += v.nbytes * (1 if k in whole else 2)
                elif k in ins:
                    moved += v.nbytes
        lw.run(lw_ints, hl)
        sw.run(sw_ints, hs)
        t0 = time.perf_counter()
        for _ in range(a.e2e_steps):
            lw.run(lw_ints, hl)
            sw.run(sw_ints, hs)
        ms_e2e = (time.perf_counter() - t0) * 1e3 / a.e2e_steps
        e2e = {"value": ncol / (ms_e2e / 1e3), "unit": "columns/s", "ms_per_step": ms_e2e, "steps": a.e2e_steps,
               "host_device_bytes_per_step": moved,
               "path": "ccpprad_lw_run + ccpprad_sw_run (host pointers, pageable memory, the entries the Fortran "
                       "rrtmgp_*_main_run call), copies in and out included, LW then SW"}
        print(f"# end to end: {ms_e2e:.1f} ms/step ({moved / 1e9:.1f} GB moved)", file=sys.stderr, flush=True)
        del host, hl, hs

    # ---- CPU baseline: FP64 oracle on a bounded sample, and GPU error on it
    cpu = None
    acc = None
    if not a.no_cpu:
        from oracle import oracle as O
        ns = min(a.cpu_sample, UNIQUE, ncol)
        sub = pr.subset(np.arange(ns))
        sarr = capi.profile_fields(sub)
        sarr["idx"] = np.arange(1, ns + 1, dtype=np.int32)
        nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        O.set_threads(nth)
        lw_ref = dict(sarr)
        lw_ref.update(O.lw_outputs(ns, NLAY))
        sw_ref = dict(sarr)
        sw_ref.update(O.sw_outputs(ns, NLAY))
        t0 = time.perf_counter()
        O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(ns, NLAY, blksz=blksz, nangle=3), lw_ref)
        O.sw_main_run(kd_sw, lut_sw, capi.sw_default_ints(ns, ns, NLAY, blksz=blksz), sw_ref)
        tc = time.perf_counter() - t0
        cpu = {"value": ns / tc, "unit": "columns/s", "cores": O.nthreads(), "kind": "port",
               "sample": f"FP64 oracle, LW+SW all-sky, first {ns} of the {UNIQUE} distinct columns x {NLAY} "
                         f"layers, {tc:.2f} s", "host": host_info()}
        # the same on one core, over a smaller sample
        n1 = min(a.cpu_sample_1core, ns)
        s1 = pr.subset(np.arange(n1))
        a1 = capi.profile_fields(s1)
        a1["idx"] = np.arange(1, n1 + 1, dtype=np.int32)
        O.set_threads(1)
        l1 = dict(a1)
        l1.update(O.lw_outputs(n1, NLAY))
        w1 = dict(a1)
        w1.update(O.sw_outputs(n1, NLAY))
        t0 = time.perf_counter()
        O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(n1, NLAY, blksz=blksz, nangle=3), l1)
        O.sw_main_run(kd_sw, lut_sw, capi.sw_default_ints(n1, n1, NLAY, blksz=blksz), w1)
        t1c = time.perf_counter() - t0
        O.set_threads(nth)
        cpu["one_core"] = {"value": n1 / t1c, "unit": "columns/s", "cores": 1,
                           "sample": f"first {n1} distinct columns, {t1c:.2f} s"}
        # the same source built with FP32 per-g-point optics and transport,
        # all cores, the same sample; its error against the FP64 run
        O.lib32()
        O.set_threads(nth)
        l32 = dict(sarr)
        l32.update(O.lw_outputs(ns, NLAY))
        w32 = dict(sarr)
        w32.update(O.sw_outputs(ns, NLAY))
        t0 = time.perf_counter()
        O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(ns, NLAY, blksz=blksz, nangle=3), l32, fp32=True)
        O.sw_main_run(kd_sw, lut_sw, capi.sw_default_ints(ns, ns, NLAY, blksz=blksz), w32, fp32=True)
        t32 = time.perf_counter() - t0

        def mx(x, y, scale=1.0):
            return float(np.abs(x - y).max() * scale)

        cpu["fp32"] = {
            "value": ns / t32, "unit": "columns/s", "cores": O.lib32().orc_nthreads(),
            "sample": f"FP32 build of the oracle, same {ns} columns, {t32:.2f} s",
            "lw_flux_maxabs_vs_fp64_Wm2": max(mx(l32[k], lw_ref[k]) for k in ("fluxlwUP_allsky", "fluxlwDOWN_allsky")),
            "sw_flux_maxabs_vs_fp64_Wm2": max(mx(w32[k], sw_ref[k]) for k in ("fluxswUP_allsky", "fluxswDOWN_allsky")),
            "lw_hr_maxabs_vs_fp64_Kday": mx(l32["htrlw"], lw_ref["htrlw"], 86400.0),
            "sw_hr_maxabs_vs_fp64_Kday": mx(w32["htrsw"], sw_ref["htrsw"], 86400.0)}
        del l32, w32
        # GPU results on the same columns (the first ns columns of the device run)
        g = {k: v[..., :ns].cpu().numpy() if v.dim() == 2 else v for k, v in
             {"lwu": lw_out["fluxlwUP_allsky"], "lwd": lw_out["fluxlwDOWN_allsky"], "lwh": lw_out["htrlw"],
              "swu": sw_out["fluxswUP_allsky"], "swd": sw_out["fluxswDOWN_allsky"], "swh": sw_out["htrsw"]}.items()}

        def e(x, y):
            d = x - y
            return float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))

        if nn_out is not None:
            gl = {k: v[..., :ns].cpu().numpy() for k, v in
                  {"u": nn_out["fluxlwUP_allsky"], "d": nn_out["fluxlwDOWN_allsky"], "h": nn_out["htrlw"]}.items()}
            hu = e(gl["h"] * 86400, lw_ref["htrlw"] * 86400)
            nn["accuracy_vs_fp64_oracle"] = {
                "lw_flux_maxabs_Wm2": max(e(gl["u"], lw_ref["fluxlwUP_allsky"])[0], e(gl["d"], lw_ref["fluxlwDOWN_allsky"])[0]),
                "lw_flux_rmse_Wm2": max(e(gl["u"], lw_ref["fluxlwUP_allsky"])[1], e(gl["d"], lw_ref["fluxlwDOWN_allsky"])[1]),
                "lw_hr_maxabs_Kday": hu[0], "lw_hr_rmse_Kday": hu[1], "columns": ns}
        if lws_out is not None:
            # the scattering path on the first 4,096 sample columns
            n4 = min(4096, ns)
            s4 = pr.subset(np.arange(n4))
            sa4 = capi.profile_fields(s4)
            sref = dict(sa4)
            sref.update(O.lw_outputs(n4, NLAY))
            O.lw_main_run(kd_lw, lut_lw, capi.lw_default_ints(n4, NLAY, blksz=blksz, nangle=3, scat=True), sref)
            gs = {k: v[..., :n4].cpu().numpy() for k, v in
                  {"u": lws_out["fluxlwUP_allsky"], "d": lws_out["fluxlwDOWN_allsky"], "h": lws_out["htrlw"]}.items()}
            hs = e(gs["h"] * 86400, sref["htrlw"] * 86400)
            lws["accuracy_vs_fp64_oracle"] = {
                "lw_flux_maxabs_Wm2": max(e(gs["u"], sref["fluxlwUP_allsky"])[0], e(gs["d"], sref["fluxlwDOWN_allsky"])[0]),
                "lw_hr_maxabs_Kday": hs[0], "lw_hr_rmse_Kday": hs[1], "columns": n4}
        acc = {
            "lw_flux_maxabs_Wm2": max(e(g["lwu"], lw_ref["fluxlwUP_allsky"])[0], e(g["lwd"], lw_ref["fluxlwDOWN_allsky"])[0]),
            "sw_flux_maxabs_Wm2": max(e(g["swu"], sw_ref["fluxswUP_allsky"])[0], e(g["swd"], sw_ref["fluxswDOWN_allsky"])[0]),
            "lw_hr_maxabs_Kday": e(g["lwh"] * 86400, lw_ref["htrlw"] * 86400)[0],
            "lw_hr_rmse_Kday": e(g["lwh"] * 86400, lw_ref["htrlw"] * 86400)[1],
            "sw_hr_maxabs_Kday": e(g["swh"] * 86400, sw_ref["htrsw"] * 86400)[0],
            "sw_hr_rmse_Kday": e(g["swh"] * 86400, sw_ref["htrsw"] * 86400)[1],
            "columns": ns,
        }
    line = {
        "metric": METRIC, "value": value, "unit": "columns/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"LW+SW all-sky table path, {ncol} columns x {NLAY} layers "
                               f"(LW 16 bands/128 g, 3 Gauss angles; SW 14 bands/112 g; McICA max-random; "
                               f"FP32 optics/transport, FP64 reductions and heating rate)",
                   "ncol": ncol, "nlay": NLAY, "path": "table", "blksz": blksz},
        "roofline": roof, "cpu_baseline": cpu, "accuracy_vs_fp64_oracle": acc, "nn": nn,
        "lw_scattering": lws, "with_aerosols": aer, "callers": callers, "end_to_end": e2e,
        "library": {"path": os.path.relpath(lib_path, REPO) if lib_path.startswith(REPO) else lib_path,
                    "sha256_16": file_sha(lib_path)},
    }
    print(json.dumps(line), flush=True)
    if a.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("# copy probe: 1073741824 bytes read, 1073741824 bytes written", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
