"""Build recipes: the gfx950 HIP library, the Fortran scheme modules and the
oracle (test infrastructure).  Everything is built in-tree so the shared
objects travel with the source tree."""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # ccpp-physics_amd/
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIBDIR = os.path.join(PKG, "lib")
HIP_LIB = os.path.join(LIBDIR, "libccpprad_hip.so")
F_LIB = os.path.join(LIBDIR, "libccpprad_f.so")
F_DRIVER = os.path.join(LIBDIR, "ccpp_cap_driver")
HR_DRIVER = os.path.join(LIBDIR, "hr_driver")
CALLER_DRIVER = os.path.join(LIBDIR, "caller_driver")

HIP_SOURCES = ["capi.hip", "host.hip", "gas_optics.hip", "lw_solver.hip", "lw_scat.hip", "sw_solver.hip",
               "clouds.hip", "util.hip", "mlp.hip", "prep.hip", "nc_tables.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_hip(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-o", tmp] + srcs)
    os.replace(tmp, HIP_LIB)
    if verbose:
        print("built", HIP_LIB)
    return HIP_LIB


def build_fortran(verbose=False):
    """Fortran CCPP scheme modules (same names/argument lists as the
    reference) over ISO_C_BINDING, plus a cap-like test driver."""
    if not os.path.isdir(FSRC) or not os.path.exists(FLANG):
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    bdir = os.path.join(LIBDIR, "fmod")
    os.makedirs(bdir, exist_ok=True)
    order = ["mpi_f08_stub.F90", "radiation_flux_types.F90", "ccpprad_c.F90", "mo_heating_rates.F90",
             "rrtmgp_lw_main.F90", "rrtmgp_sw_main.F90", "GFS_rrtmgp_cloud_overlap.F90", "GFS_rrtmgp_post.F90"]
    objs = []
    for f in order:
        src = os.path.join(FSRC, f)
        if not os.path.exists(src):
            continue
        obj = os.path.join(bdir, f.replace(".F90", ".o"))
        _run([FLANG, "-O2", "-fPIC", "-module-dir", bdir, "-c", src, "-o", obj], cwd=bdir)
        objs.append(obj)
    _run([FLANG, "-shared", "-o", F_LIB] + objs + ["-L" + LIBDIR, "-lccpprad_hip",
                                                   "-Wl,-rpath,$ORIGIN"])
    for name, exe in (("ccpp_cap_driver.F90", F_DRIVER), ("hr_driver.F90", HR_DRIVER),
                      ("caller_driver.F90", CALLER_DRIVER)):
        drv = os.path.join(FSRC, name)
        if os.path.exists(drv):
            _run([FLANG, "-O2", "-I", bdir, "-module-dir", bdir, drv, "-o", exe, "-L" + LIBDIR,
                  "-lccpprad_f", "-lccpprad_hip", "-Wl,-rpath,$ORIGIN"], cwd=bdir)
    if verbose:
        print("built", F_LIB)
    return F_LIB


def build_oracle(verbose=False):
    odir = os.path.join(REPO, "oracle")
    _run(["make", "-s", "-C", odir, "liboracle.so", "liboracle32.so"])
    if os.path.isdir("/root/reference/physics"):
        try:
            _run(["make", "-s", "-C", odir, "ref"])
        except RuntimeError as e:  # reference build is optional
            if verbose:
                print("reference build skipped:", e)
    return os.path.join(odir, "liboracle.so")


def build_all(verbose=False):
    build_hip(verbose)
    build_fortran(verbose)
    build_oracle(verbose)
