"""Shared helpers of the parity tests."""
import numpy as np


def err_stats(a, b):
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    return float(np.abs(d).max()), float(np.sqrt(np.mean(d * d)))


def report(name, a, b, unit):
    mx, rms = err_stats(a, b)
    print(f"{name}: max-abs {mx:.3e} {unit}, rmse {rms:.3e} {unit}")
    return mx, rms
