"""MI355X-native CCPP RRTMGP radiation hot path (host side).

Modules: synthetic (tables/profiles), capi (ctypes mirror of include/ccpprad.h),
lib (the HIP library and the scheme classes), tables (table-file writer),
build (build recipes), device (torch plumbing for device-resident runs)."""
from . import synthetic, capi  # noqa: F401

__all__ = ["synthetic", "capi"]
