"""Writes tests/golden/healpix_ref.npz from the reference's own HEALPix (src/Healpix_2.15a compiled as-is into
oracle/_ref by oracle/Makefile): (z, phi) of every NEST pixel and the NEST children of every pixel, orders 0-4, and
the NEST pixel (ang2pix_z_phi) of 4000 seeded random directions plus the pole and seam cases, orders 0-4.

Run in the development container, where /root/reference exists: `python tests/golden/make_healpix_golden.py`.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

ref = oracle.ref_healpix()
if ref is None:
    raise SystemExit("oracle/_ref/libhealpix_ref.so is missing: run `make -C oracle` with /root/reference present")
out = {}
for order in range(5):
    n = int(ref.ref_hp_npix(order))
    z, phi = np.zeros(n), np.zeros(n)
    ix, iy, face = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
    for p in range(n):
        a, b = C.c_double(), C.c_double()
        ref.ref_hp_pix2ang_z_phi(order, p, C.byref(a), C.byref(b))
        z[p], phi[p] = a.value, b.value
        i, j, f = C.c_int(), C.c_int(), C.c_int()
        ref.ref_hp_nest2xyf(order, p, C.byref(i), C.byref(j), C.byref(f))
        ix[p], iy[p], face[p] = i.value, j.value, f.value
    out[f"z{order}"], out[f"phi{order}"] = z, phi
    out[f"ix{order}"], out[f"iy{order}"], out[f"face{order}"] = ix, iy, face
rng = np.random.default_rng(20241)
az = np.concatenate([rng.uniform(-1, 1, 4000), [1.0, -1.0, 2.0 / 3.0, -2.0 / 3.0, 0.0, 0.0]])
aphi = np.concatenate([rng.uniform(0, 2 * np.pi, 4000), [0.3, 5.0, 1.0, 2.0, 0.0, np.pi / 2]])
out["a_z"], out["a_phi"] = az, aphi
for order in range(5):
    out[f"a_pix{order}"] = np.array([ref.ref_hp_ang2pix_z_phi(order, float(a), float(b)) for a, b in zip(az, aphi)],
                                    np.int64)
np.savez_compressed(os.path.join(HERE, "healpix_ref.npz"), **out)
print("wrote", os.path.join(HERE, "healpix_ref.npz"))
