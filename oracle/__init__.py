"""TEST INFRASTRUCTURE ONLY -- Python binding of the parity oracle (liborc_estep.so, built from
oracle/estep_oracle.cpp by oracle/Makefile) and of the reference's own HEALPix (oracle/_ref, built from
/root/reference/src/Healpix_2.15a where the reference is present).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, to check the product's results.
The product (relion_amd) never imports, links or runs anything under oracle/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from relion_amd import EstepConfig, ParticleMeta, ParticleResult, RESULT_DTYPE, Wsum

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc_estep.so")
REF_HEALPIX = os.path.join(HERE, "_ref", "libhealpix_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    """Compile the oracle (and oracle/_ref when /root/reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OrcModel(C.Structure):
    _fields_ = [("volumes", C.POINTER(_dp)), ("pdf_class", _dp), ("pdf_direction", _dp), ("sigma2_noise", _dp),
                ("data_vs_prior", _dp)]


class OrcDebug(C.Structure):
    _fields_ = [("coarse_diff2", _dp), ("coarse_weight", _dp), ("coarse_significant", C.POINTER(C.c_ubyte)),
                ("fine_diff2", _dp), ("fine_weight", _dp)]


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_hp_pix2ang_z_phi": (None, [C.c_int, C.c_longlong, _dp, _dp]),
        "orc_hp_nest2xyf": (None, [C.c_int, C.c_longlong, _ip, _ip, _ip]),
        "orc_hp_xyf2nest": (C.c_longlong, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_euler_angles2matrix": (None, [C.c_double, C.c_double, C.c_double, _dp]),
        "orc_ctf_value": (C.c_double, [C.c_double] * 11),
        "orc_sampling_counts": (C.c_int, [C.POINTER(EstepConfig), _ip, _ip, _ip]),
        "orc_sampling_grid": (C.c_int, [C.POINTER(EstepConfig), _dp, _dp, _dp, _dp, _dp]),
        "orc_oversampled_orientations": (C.c_int, [C.POINTER(EstepConfig), C.c_int, C.c_int, _dp, _dp, _dp]),
        "orc_oversampled_translations": (C.c_int, [C.POINTER(EstepConfig), C.c_int, _dp, _dp]),
        "orc_sampling_local": (C.c_int, [C.POINTER(EstepConfig), C.c_double, C.c_double, C.c_double, _ip, _dp,
                                         C.c_int, _ip, _ip, _dp, C.c_int, _ip]),
        "orc_project": (None, [_dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp]),
        "orc_backproject": (None, [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp,
                                   _dp]),
        "orc_shift_image": (None, [_dp, C.c_int, C.c_double, C.c_double, C.c_double, _dp]),
        "orc_estep": (C.c_int, [C.POINTER(EstepConfig), C.POINTER(OrcModel), C.c_int, C.POINTER(ParticleMeta), _dp,
                                _dp, _dp, C.POINTER(ParticleResult), C.POINTER(Wsum), _dp, _dp, C.POINTER(OrcDebug),
                                C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def ref_healpix():
    """The reference's HEALPix (oracle/_ref) or None when it was not built."""
    if not os.path.exists(REF_HEALPIX):
        return None
    r = C.CDLL(REF_HEALPIX)
    r.ref_hp_pix2ang_z_phi.argtypes = [C.c_int, C.c_longlong, _dp, _dp]
    r.ref_hp_nest2xyf.argtypes = [C.c_int, C.c_longlong, _ip, _ip, _ip]
    r.ref_hp_xyf2nest.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    r.ref_hp_xyf2nest.restype = C.c_longlong
    r.ref_hp_ang2pix_z_phi.argtypes = [C.c_int, C.c_double, C.c_double]
    r.ref_hp_ang2pix_z_phi.restype = C.c_longlong
    r.ref_hp_npix.argtypes = [C.c_int]
    r.ref_hp_npix.restype = C.c_longlong
    return r


def _p(a):
    return a.ctypes.data_as(_dp)


def euler(rot, tilt, psi):
    A = np.zeros(9)
    lib.orc_euler_angles2matrix(rot, tilt, psi, _p(A))
    return A.reshape(3, 3)


def ctf_value(*args):
    return lib.orc_ctf_value(*args)


def sampling_counts(cfg):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert lib.orc_sampling_counts(C.byref(cfg), C.byref(a), C.byref(b), C.byref(c)) == 0
    return a.value, b.value, c.value


def sampling_grid(cfg):
    nd, npsi, nt = sampling_counts(cfg)
    rot, tilt, psi, tx, ty = np.zeros(nd), np.zeros(nd), np.zeros(npsi), np.zeros(nt), np.zeros(nt)
    assert lib.orc_sampling_grid(C.byref(cfg), _p(rot), _p(tilt), _p(psi), _p(tx), _p(ty)) == 0
    return rot, tilt, psi, tx, ty


def oversampled_orientations(cfg, idir, ipsi):
    r, t, p = np.zeros(8), np.zeros(8), np.zeros(8)
    assert lib.orc_oversampled_orientations(C.byref(cfg), idir, ipsi, _p(r), _p(t), _p(p)) == 0
    return r, t, p


def oversampled_translations(cfg, itrans):
    x, y = np.zeros(4), np.zeros(4)
    assert lib.orc_oversampled_translations(C.byref(cfg), itrans, _p(x), _p(y)) == 0
    return x, y


def sampling_local(cfg, prior_rot, prior_tilt, prior_psi):
    nd, npsi, _ = sampling_counts(cfg)
    di, dp = np.zeros(nd, np.int32), np.zeros(nd)
    pi, pp = np.zeros(npsi, np.int32), np.zeros(npsi)
    n1, n2 = C.c_int(), C.c_int()
    assert lib.orc_sampling_local(C.byref(cfg), prior_rot, prior_tilt, prior_psi, di.ctypes.data_as(_ip), _p(dp), nd,
                                  C.byref(n1), pi.ctypes.data_as(_ip), _p(pp), npsi, C.byref(n2)) == 0
    return di[:n1.value], dp[:n1.value], pi[:n2.value], pp[:n2.value]


def project(vol, r_max, padding_factor, A, s):
    """vol: (z, y, x, 2) float64 projector layout. Returns (s, s//2+1) complex128."""
    v = np.ascontiguousarray(vol, np.float64)
    zd, yd, xd = v.shape[:3]
    out = np.zeros((s, s // 2 + 1, 2))
    Aa = np.ascontiguousarray(A, np.float64).reshape(9)
    lib.orc_project(_p(v), xd, yd, zd, r_max, padding_factor, _p(Aa), s, _p(out))
    return out[..., 0] + 1j * out[..., 1]


def backproject(vol, weight, r_max, padding_factor, A, img, mweight=None):
    """Accumulates img (s, s//2+1) complex into vol (z,y,x,2) float64 and weight (z,y,x) float64 in place."""
    zd, yd, xd = weight.shape
    s = img.shape[0]
    im = np.ascontiguousarray(np.stack([img.real, img.imag], -1), np.float64)
    mw = None if mweight is None else np.ascontiguousarray(mweight, np.float64)
    Aa = np.ascontiguousarray(A, np.float64).reshape(9)
    assert vol.flags.c_contiguous and weight.flags.c_contiguous and vol.dtype == np.float64
    lib.orc_backproject(_p(vol), _p(weight), xd, yd, zd, r_max, padding_factor, _p(Aa), s, _p(im),
                        None if mw is None else _p(mw))


def shift_image(img, oridim, xshift, yshift):
    s = img.shape[0]
    im = np.ascontiguousarray(np.stack([img.real, img.imag], -1), np.float64)
    out = np.zeros_like(im)
    lib.orc_shift_image(_p(im), s, oridim, xshift, yshift, _p(out))
    return out[..., 0] + 1j * out[..., 1]


def estep(cfg, volumes, pdf_class, sigma2_noise, meta, fimg, fctf, fimg_nomask=None, pdf_direction=None,
          data_vs_prior=None, debug_particles=(), with_bp=True, threads=0):
    """Full CPU E-step of the particles in `meta` (ctypes ParticleMeta array).

    fimg: (n, s, s//2+1, 2) or complex; fctf: (n, s, s//2+1). Returns results (structured array), wsum dict,
    (bp_data, bp_weight) and a dict of dense intermediates for the particles in debug_particles."""
    n = len(meta)
    K = cfg.nr_classes
    nd, npsi, nt = sampling_counts(cfg)

    def as_pairs(a):
        a = np.asarray(a)
        if np.iscomplexobj(a):
            a = np.stack([a.real, a.imag], -1)
        return np.ascontiguousarray(a, np.float64)

    X = as_pairs(fimg)
    Xn = X if fimg_nomask is None else as_pairs(fimg_nomask)
    Cc = np.ascontiguousarray(fctf, np.float64)
    vols = [np.ascontiguousarray(v, np.float64) for v in volumes]
    vp = (_dp * K)(*[_p(v) for v in vols])
    pc = np.ascontiguousarray(pdf_class, np.float64)
    s2 = np.ascontiguousarray(sigma2_noise, np.float64)
    pd = np.ones(K * nd) if pdf_direction is None else np.ascontiguousarray(pdf_direction, np.float64).ravel()
    dv = None if data_vs_prior is None else np.ascontiguousarray(data_vs_prior, np.float64)
    model = OrcModel(vp, _p(pc), _p(pd), _p(s2), None if dv is None else _p(dv))
    res = (ParticleResult * max(n, 1))()
    nshell = cfg.ori_size // 2 + 1
    ws_s2, ws_pc, ws_pd = np.zeros(nshell), np.zeros(K), np.zeros(K * nd)
    ws = Wsum(_p(ws_s2), _p(ws_pc), _p(ws_pd), 0, 0, 0, 0, 0, 0)
    zd, yd, xd = vols[0].shape[:3]
    bp_data = np.zeros((K, zd, yd, xd, 2)) if with_bp else None
    bp_w = np.zeros((K, zd, yd, xd)) if with_bp else None
    dbg_arr = (OrcDebug * max(n, 1))()
    dbg_out = {}
    for ip in debug_particles:
        if cfg.local_search:
            di, _, pi, _ = sampling_local(cfg, meta[ip].prior_rot, meta[ip].prior_tilt, meta[ip].prior_psi)
            ncoarse = K * len(di) * len(pi) * nt
        else:
            ncoarse = K * nd * npsi * nt
        d = dict(coarse_diff2=np.zeros(ncoarse), coarse_weight=np.zeros(ncoarse),
                 coarse_significant=np.zeros(ncoarse, np.uint8), fine_diff2=np.zeros(ncoarse * 32),
                 fine_weight=np.zeros(ncoarse * 32))
        dbg_arr[ip] = OrcDebug(_p(d["coarse_diff2"]), _p(d["coarse_weight"]),
                               d["coarse_significant"].ctypes.data_as(C.POINTER(C.c_ubyte)), _p(d["fine_diff2"]),
                               _p(d["fine_weight"]))
        dbg_out[ip] = d
    rc = lib.orc_estep(C.byref(cfg), C.byref(model), n, meta, _p(X), _p(Xn), _p(Cc), res, C.byref(ws),
                       None if bp_data is None else _p(bp_data), None if bp_w is None else _p(bp_w), dbg_arr, threads)
    if rc != 0:
        raise RuntimeError(f"oracle E-step failed with status {rc}")
    results = np.frombuffer(res, dtype=RESULT_DTYPE, count=n).copy()
    wsum = dict(sigma2_noise=ws_s2, pdf_class=ws_pc, pdf_direction=ws_pd.reshape(K, nd),
                sigma2_offset=ws.sigma2_offset, sumw_group=ws.sumw_group, signal_product=ws.signal_product,
                reference_power=ws.reference_power, LL=ws.LL, ave_Pmax=ws.ave_Pmax)
    return results, wsum, (bp_data, bp_w), dbg_out
