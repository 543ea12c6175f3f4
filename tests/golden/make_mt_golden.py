"""Writes tests/golden/mt19937_draws.json from the reference's own
physics/Radiation/mersenne_twister.f, compiled unmodified by
`make -C oracle ref` into oracle/_ref/libccpp_ref.so (run in the container
that has the reference tree; the fixture is then committed)."""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)


def main():
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libccpp_ref.so not built (make -C oracle ref)")
    out = {"source": "physics/Radiation/mersenne_twister.f random_setseed_t + random_number_t",
           "draws": {}, "chunked": {}}
    for seed in (1, 129, 4357, 12345, -7, 0, 2147483647):
        a = np.zeros(700)
        R.ref_mt_draws(seed, a.size, a.ctypes.data_as(C.POINTER(C.c_double)))
        out["draws"][str(seed)] = [float(x) for x in a]
    # the pattern of rrtmgp_lw_main.F90:458-461: one ngpt-vector per layer
    for seed, chunk, n in ((129 + 5, 128, 6), (112 + 77, 112, 7)):
        a = np.zeros(chunk * n)
        R.ref_mt_draws_chunked(seed, chunk, n, a.ctypes.data_as(C.POINTER(C.c_double)))
        out["chunked"][f"{seed}:{chunk}:{n}"] = [float(x) for x in a]
    with open(os.path.join(HERE, "mt19937_draws.json"), "w") as f:
        json.dump(out, f)
    print("wrote", len(out["draws"]), "seed vectors")


if __name__ == "__main__":
    main()
