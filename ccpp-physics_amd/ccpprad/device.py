"""torch plumbing: move scheme arguments to HBM and back.  torch provides
device memory and streams only; every computation runs in the HIP library."""
from __future__ import annotations

import numpy as np
import torch

from . import capi


def to_device(arrays: dict, device="cuda:0") -> dict:
    out = {}
    for k, v in arrays.items():
        if v is None:
            continue
        if isinstance(v, np.ndarray):
            t = torch.from_numpy(np.ascontiguousarray(v))
            out[k] = t.to(device)
        else:
            out[k] = v
    return out


def lw_outputs(ncol, nlay, device="cuda:0", hr=True):
    o = {n: torch.zeros((nlay + 1, ncol), dtype=torch.float64, device=device)
         for n in capi.LW_OUTPUT_FIELDS if n.startswith("flux")}
    if hr:
        o["htrlw"] = torch.zeros((nlay, ncol), dtype=torch.float64, device=device)
        o["htrlwc"] = torch.zeros((nlay, ncol), dtype=torch.float64, device=device)
    return o


def sw_outputs(ncol, nlay, device="cuda:0", hr=True):
    o = {n: torch.zeros((nlay + 1, ncol), dtype=torch.float64, device=device)
         for n in ("fluxswUP_allsky", "fluxswDOWN_allsky", "fluxswUP_clrsky", "fluxswDOWN_clrsky")}
    o["scmpsw"] = torch.zeros((ncol, 6), dtype=torch.float64, device=device)
    o["cldtausw"] = torch.zeros((nlay, ncol), dtype=torch.float64, device=device)
    if hr:
        o["htrsw"] = torch.zeros((nlay, ncol), dtype=torch.float64, device=device)
        o["htrswc"] = torch.zeros((nlay, ncol), dtype=torch.float64, device=device)
    return o


def to_host(d: dict) -> dict:
    return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in d.items()}
