"""Learned LW gas optics (host side of the MLP emulator in csrc/mlp.hip).

The network maps a layer's (T, ln p, ln vmr of H2O, O3, CO2, CH4, N2O) to
ln(tau_g / col_dry) and the Planck fraction of each of the 128 g-points.
The optical depth is per unit dry-air column because tau is proportional to
the layer's air mass, which the seven inputs do not determine.

Training samples come from seeded profiles drawn like the all-sky LW
workload (lapse-rate temperature with N(0, 3 K) noise, exponential water
vapour, an ozone layer, trace gases scaled by U(0.8, 1.2)). The k-distribution
gas optics computes the targets. That gas optics is the GPU table path,
checked against the FP64 restatement in the parity tests. Training, the
gradients and the optimiser all run in the HIP library on BF16 MFMA with
FP32 accumulation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi, synthetic as syn
from .lib import load, _dp

NIN, NOUT, NG = 7, 256, 128
FEATURES = ("t_lay", "ln_p_lay", "ln_vmr_h2o", "ln_vmr_o3", "ln_vmr_co2", "ln_vmr_ch4", "ln_vmr_n2o")
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "mlp_lw_gas_optics.npz")
AVOGAD, M_DRY, M_H2O, GRAV = 6.02214076e23, 0.028964, 0.018016, 9.80665


def _bind(L):
    vp, cp, ci = C.c_void_p, C.c_char_p, C.c_int
    if getattr(L, "_mlp_bound", False):
        return L
    L.ccpprad_mlp_create.argtypes = [C.POINTER(vp), cp, ci]
    L.ccpprad_mlp_destroy.argtypes = [vp]
    L.ccpprad_mlp_destroy.restype = None
    L.ccpprad_mlp_set_params.argtypes = [vp, vp, cp, ci]
    L.ccpprad_mlp_get_params.argtypes = [vp, vp, vp]
    L.ccpprad_mlp_set_norm.argtypes = [vp, vp, vp, vp, vp, vp, vp, cp, ci]
    L.ccpprad_mlp_forward_dev.argtypes = [vp, vp, ci, vp, vp, cp, ci]
    L.ccpprad_mlp_grads_dev.argtypes = [vp, vp, vp, ci, vp, vp, cp, ci]
    L.ccpprad_mlp_train_step_dev.argtypes = [vp, vp, vp, ci, C.c_float, vp, vp, cp, ci]
    L.ccpprad_mlp_finalize_training.argtypes = [vp, cp, ci]
    L.ccpprad_mlp_infer_dev.argtypes = [vp, ci, ci] + [vp] * 13 + [cp, ci]
    L.ccpprad_lw_set_gas_optics_mlp.argtypes = [vp, vp]
    L._mlp_bound = True
    return L


def col_dry(p_lev, vmr_h2o):
    """Dry-air column [molecules/cm2] per layer, (nlay, ncol) numpy convention."""
    dp = np.abs(p_lev[:-1] - p_lev[1:])
    fact = 1.0 / (1.0 + vmr_h2o)
    m_air = (M_DRY + M_H2O * vmr_h2o) * fact
    return 10.0 * dp * AVOGAD / (1000.0 * m_air * 100.0 * GRAV) * fact


def features(pr: syn.Profiles):
    """Raw features, rows ordered column-major like the device tau layout:
    row = c * nlay + l.  Returns (rows, 7) FP64."""
    f = [pr.t_lay, np.log(pr.p_lay)] + [np.log(pr.vmr[g]) for g in ("h2o", "o3", "co2", "ch4", "n2o")]
    return np.stack([a.T.reshape(-1) for a in f], axis=1)


class LWEmulator:
    """Owns a ccpprad_mlp (device parameters, Adam state, inference image)."""

    def __init__(self, seed=0):
        self.L = _bind(load())
        h = C.c_void_p()
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_create(C.byref(h), buf, n), buf)
        self.h = h
        self.nparams = self.L.ccpprad_mlp_nparams()
        rng = np.random.default_rng(seed)
        sizes = (NIN, 128, 128, 128, NOUT)
        p = []
        for k, m in zip(sizes[:-1], sizes[1:]):
            p.append(rng.normal(0.0, np.sqrt(2.0 / k), (k, m)).reshape(-1))
            p.append(np.zeros(m))
        self.set_params(np.concatenate(p).astype(np.float32))
        self.norm = None

    def close(self):
        if getattr(self, "h", None):
            self.L.ccpprad_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def set_params(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.nparams
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_set_params(self.h, flat.ctypes.data, buf, n), buf)

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        self.L.ccpprad_mlp_get_params(self.h, p.ctypes.data, g.ctypes.data)
        return p, g

    def set_norm(self, xmean, xstd, lo, hi, ymean, ystd):
        self.norm = dict(xmean=np.asarray(xmean, np.float32), xstd=np.asarray(xstd, np.float32),
                         lo=np.asarray(lo, np.float32), hi=np.asarray(hi, np.float32),
                         ymean=np.asarray(ymean, np.float32), ystd=np.asarray(ystd, np.float32))
        a = [np.ascontiguousarray(self.norm[k]) for k in ("xmean", "xstd", "lo", "hi", "ymean", "ystd")]
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_set_norm(self.h, *[x.ctypes.data for x in a], buf, n), buf)

    def save(self, path):
        p, _ = self.get_params()
        np.savez(path, params=p, **self.norm)

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS):
        d = np.load(path, allow_pickle=False)
        e = cls()
        e.set_params(d["params"])
        e.set_norm(d["xmean"], d["xstd"], d["lo"], d["hi"], d["ymean"], d["ystd"])
        return e

    # ---- device entry points (torch tensors)
    def forward(self, X, Y, stream=None):
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_forward_dev(self.h, _dp(X), X.shape[0], _dp(Y), stream, buf, n), buf)

    def grads(self, X, T, loss, stream=None):
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_grads_dev(self.h, _dp(X), _dp(T), X.shape[0], _dp(loss), stream, buf, n), buf)

    def train_step(self, X, T, lr, loss, stream=None):
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_train_step_dev(self.h, _dp(X), _dp(T), X.shape[0], C.c_float(lr), _dp(loss),
                                                      stream, buf, n), buf)

    def finalize(self):
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_finalize_training(self.h, buf, n), buf)

    def infer(self, ncol, nlay, d, tau, pfrac, oob, colflag, stream=None):
        """d: dict of device tensors p_lay, p_lev, t_lay, vmr_* (CCPP layout)."""
        buf, n = capi.errbuf()
        capi.check(self.L.ccpprad_mlp_infer_dev(
            self.h, ncol, nlay, _dp(d["p_lay"]), _dp(d["p_lev"]), _dp(d["t_lay"]), _dp(d["vmr_h2o"]),
            _dp(d["vmr_o3"]), _dp(d["vmr_co2"]), _dp(d["vmr_ch4"]), _dp(d["vmr_n2o"]), _dp(tau), _dp(pfrac),
            _dp(oob), _dp(colflag), stream, buf, n), buf)

    def attach(self, lw_scheme):
        """Make an LWScheme use this emulator for its gas optics (None: table)."""
        rc = self.L.ccpprad_lw_set_gas_optics_mlp(lw_scheme.ctx, self.h)
        if rc:
            raise capi.CcppradError("cannot attach the emulator to this context")


def detach(lw_scheme):
    L = _bind(load())
    L.ccpprad_lw_set_gas_optics_mlp(lw_scheme.ctx, None)


def make_training_set(kd, nsamples=4_000_000, nlay=127, seed=20261204, torch=None, chunk=16384):
    """Layer samples from seeded profiles drawn like the all-sky LW workload.
    Targets come from the GPU k-distribution gas optics.  Returns torch
    tensors (device): raw features (rows, 7) FP64, targets (rows, 256) FP32
    [ln(tau/col_dry) | pfrac]."""
    from . import lib, device
    ncols = (nsamples + nlay - 1) // nlay
    s = lib.LWScheme(kd, None, nlay)
    feats, targs = [], []
    for c0 in range(0, ncols, chunk):
        nc = min(chunk, ncols - c0)
        pr = syn.make_profiles(nc, nlay, seed=seed + c0, clouds=False, kd=kd)
        arr = capi.profile_fields(pr)
        d = device.to_device({k: arr[k] for k in ("p_lay", "p_lev", "t_lay")})
        v = torch.from_numpy(np.ascontiguousarray(np.stack([pr.vmr[g] for g in syn.GAS_NAMES]))).cuda()
        tau = torch.empty((nc, nlay, kd.ngpt), dtype=torch.float32, device="cuda")
        pf = torch.empty_like(tau)
        s.gas_optics_dev(nc, nlay, d["p_lay"], d["p_lev"], d["t_lay"], v, tau, pf)
        cd = torch.from_numpy(col_dry(pr.p_lev, pr.vmr["h2o"]).T.reshape(-1)).cuda()
        y = torch.cat([torch.log(tau.reshape(-1, kd.ngpt).double() / cd[:, None]).float(),
                       pf.reshape(-1, kd.ngpt)], dim=1)
        feats.append(torch.from_numpy(features(pr)).cuda())
        targs.append(y)
    X = torch.cat(feats)[:nsamples]
    Y = torch.cat(targs)[:nsamples]
    s.close()
    return X, Y


def fit_norm(X, Y, margin=1e-3):
    """Standardisation of features/targets and the training hull: per-feature
    min/max widened by margin x span, which absorbs the FP32 rounding of the
    bounds and of the on-device feature computation (a feature pinned at its
    extreme, like water vapour at its 1e-6 floor aloft, stays inside)."""
    xm = X.mean(0)
    xs = X.std(0).clamp_min(1e-6)
    ym = Y.mean(0)
    ys = Y.std(0).clamp_min(1e-6)
    lo = X.min(0).values
    hi = X.max(0).values
    span = hi - lo
    return dict(xmean=xm.float().cpu().numpy(), xstd=xs.float().cpu().numpy(),
                lo=(lo - margin * span).float().cpu().numpy(), hi=(hi + margin * span).float().cpu().numpy(),
                ymean=ym.float().cpu().numpy(), ystd=ys.float().cpu().numpy())


def train(emu: LWEmulator, X, Y, torch, epochs=20, batch=8192, lr=2e-3, lr_min=2e-5, seed=0, log=None):
    """Adam with a cosine learning-rate decay over shuffled mini-batches of
    standardised data.  X, Y: device tensors (raw features, targets)."""
    n = emu.norm
    dev = X.device
    Xs = ((X - torch.from_numpy(n["xmean"]).to(dev).double()) / torch.from_numpy(n["xstd"]).to(dev).double()).float()
    Ys = ((Y - torch.from_numpy(n["ymean"]).to(dev)) / torch.from_numpy(n["ystd"]).to(dev)).float().contiguous()
    rows = Xs.shape[0]
    steps = epochs * (rows // batch)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    hist = []
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(rows, device=dev, generator=gen)
        tot = 0.0
        for b in range(rows // batch):
            idx = perm[b * batch:(b + 1) * batch]
            xb = Xs[idx].contiguous()
            yb = Ys[idx].contiguous()
            a = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * k / max(1, steps)))
            emu.train_step(xb, yb, float(a), loss)
            k += 1
            if b % 64 == 0:
                tot = float(loss.item())
        hist.append(tot)
        if log:
            log(f"epoch {ep}: loss {tot:.5f}")
    emu.finalize()
    return hist
