"""The C-ABI library loads (without a GPU) and exports every function include/rln_estep.h declares."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import relion_amd as ra

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
HEADERS = sorted(os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(rln_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_symbols_exported():
    names = declared_functions()
    assert len(HEADERS) >= 2 and len(names) >= 45
    assert "rln_proposal_train" in names and "rln_estep_set_orientation_sets" in names
    lib = C.CDLL(ra.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header_layout():
    # rln_particle_meta: 9 doubles; rln_particle_result: 17 doubles + int64 + 2 ints
    assert C.sizeof(ra.ParticleMeta) == 9 * 8
    assert C.sizeof(ra.ParticleResult) == 17 * 8 + 8 + 2 * 4
    assert C.sizeof(ra.Wsum) == 3 * 8 + 6 * 8


def test_errors_are_reported_without_gpu():
    cfg = ra.make_config(64, 2.0, 2, 10.0, 2.0, 24)
    cfg.adaptive_oversampling = 2
    with pytest.raises(ra.RelionError):
        ra.ExpectationStep(cfg)  # unsupported option or no device: both raise, never a silent fallback
    bad = ra.make_config(64, 2.0, -1, 10.0, 2.0, 24)
    with pytest.raises(ra.RelionError):
        ra.sampling_counts(bad)
    assert ra.lib.rln_last_error()


def test_synthetic_generator_is_seeded_and_batch_independent():
    spec = ra.synth_spec(11, 32, 4.0, 12, snr=0.5)
    a = ra.synth_particles(spec, 6, 32)
    b = ra.synth_particles(spec, 3, 32, first=3)
    assert np.array_equal(a[0][3:], b[0]) and np.array_equal(a[1][3:], b[1])
    assert np.array_equal(a[3][3:], b[3])
    v1 = ra.synth_reference(spec, 0, 32)
    v2 = ra.synth_reference(spec, 0, 32)
    assert np.array_equal(v1, v2)
    # the padded reference is Hermitian on the x == 0 plane: V(0, -y, -z) = conj(V(0, y, z))
    xd, yd, zd, iy, iz, _ = ra.volume_dims(32)
    plane = v1[:, :, 0, 0] + 1j * v1[:, :, 0, 1]
    assert np.allclose(plane[::-1, ::-1], np.conj(plane), atol=1e-12)


def test_complex_images_become_float_pairs():
    # complex images are passed as interleaved (re, im) float32, never truncated to their real parts
    z = (np.arange(6) + 1j * (10 + np.arange(6))).astype(np.complex128).reshape(1, 2, 3)
    f = ra._float_pairs(z)
    assert f.dtype == np.float32 and f.size == 12
    assert np.array_equal(f.reshape(1, 2, 3, 2)[..., 0], z.real) and np.array_equal(f.reshape(1, 2, 3, 2)[..., 1], z.imag)
    r = np.ones((1, 2, 3, 2), np.float64)
    assert ra._float_pairs(r).shape == (1, 2, 3, 2)
