"""Writes tests/golden/cfg1_longdouble_delta.npz: the extended-precision
trajectory of the cfg1 workload (HD 64x64x(64+25), d=5, C=25, RK2, dt=1e-3,
nu=2e-3, 100 steps from the seeded spectrum-peaked field), stored as its
difference from the float64 oracle's trajectory.

The oracle is run twice on the same discrete problem: in float64 (the
reference's arithmetic) and in longdouble (x86 80-bit, eps 1.1e-19; every
constant -- wavenumbers, dir = A*Q^T, filter factors -- formed in float64 as
the reference forms it, transforms and arithmetic in longdouble).  The
longdouble run is the exact trajectory of that problem to ~1e-16.  delta =
LD - F64 is about 3e-11 of the fields, so storing it in float16 (scaled by its
largest entry) loses nothing at the 1e-13 level; a test rebuilds LD as the
float64 oracle's result + delta.

    python tests/golden/make_cfg1_longdouble.py      (about 3 minutes)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from conftest import CFG1, TABLES  # noqa: E402
from oracle import specter_ref as O  # noqa: E402
from specter_amd import ic  # noqa: E402


def run(precision):
    nx, ny, nz, C, d = CFG1["shape"]
    prm = CFG1["prm"]
    g = O.make_grid(nx, ny, nz, C, d, TABLES, precision=precision)
    u = ic.spectrum_peaked_velocity(nx, ny, nz, C, CFG1["seed"])
    zero = np.zeros(g.spec_shape, g.cdtype)
    st = O.HDState(*(O.fftp3d_real_to_complex(g, a.astype(g.rdtype)) for a in u), zero.copy(),
                   zero.copy(), zero.copy(), zero.copy())
    for _ in range(CFG1["nsteps"]):
        O.hd_step(g, st, prm["ord"], prm["dt"], prm["nu"])
    return st


def main():
    ld = run("longdouble")
    f64 = run("float64")
    out = {}
    for n in ("vx", "vy", "vz"):
        a, b = getattr(ld, n), getattr(f64, n).astype(np.clongdouble)
        dl = (a - b).astype(np.complex128)
        s = float(np.abs(dl).max())
        out[f"{n}_scale"] = np.array(s)
        out[f"{n}_re"] = (dl.real / s).astype(np.float16)
        out[f"{n}_im"] = (dl.imag / s).astype(np.float16)
        out[f"{n}_rel"] = np.array(float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(a.ravel())))
        out[f"{n}_norm"] = np.array(float(np.linalg.norm(a.ravel())))
        print(n, "float64 oracle vs longdouble:", float(out[f"{n}_rel"]))
    np.savez_compressed(os.path.join(HERE, "cfg1_longdouble_delta.npz"), **out)


if __name__ == "__main__":
    main()
