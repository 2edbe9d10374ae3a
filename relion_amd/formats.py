"""STAR tables and MRC images through the C ABI of include/rln_formats.h.

Python view of the reference's MetaDataTable (src/metadata_table.cpp) and MRC reading and writing (src/rwMRC.h):
``StarFile`` holds every data block of a file, ``read_mrc`` / ``write_mrc`` / ``mrc_info`` handle images, stacks and
maps, ``StarBlock.read_images`` loads the images an ``rlnImageName`` column names, ``read_particles`` /
``write_particles`` handle a particle set (optics and particles tables) for the image preparation and the E-step.
"""
import ctypes as C

import numpy as np

from . import _check, _ptr, lib

_vp = C.c_void_p
_cp = C.c_char_p


class MrcInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("mode", C.c_int), ("nsymbt", C.c_int),
                ("swapped", C.c_int), ("pixel_size", C.c_double), ("pixel_size_z", C.c_double),
                ("amin", C.c_float), ("amax", C.c_float), ("amean", C.c_float), ("arms", C.c_float)]


_SIG = {
    "rln_star_read": (C.c_int, [_cp, C.POINTER(_vp)]),
    "rln_star_parse": (C.c_int, [_cp, C.c_size_t, C.POINTER(_vp)]),
    "rln_star_create": (C.c_int, [C.POINTER(_vp)]),
    "rln_star_destroy": (C.c_int, [_vp]),
    "rln_star_version": (C.c_int, [_vp]),
    "rln_star_block_count": (C.c_int, [_vp]),
    "rln_star_block_name": (_cp, [_vp, C.c_int]),
    "rln_star_find_block": (C.c_int, [_vp, _cp]),
    "rln_star_block_is_list": (C.c_int, [_vp, C.c_int]),
    "rln_star_rows": (C.c_long, [_vp, C.c_int]),
    "rln_star_columns": (C.c_int, [_vp, C.c_int]),
    "rln_star_label": (_cp, [_vp, C.c_int, C.c_int]),
    "rln_star_find_label": (C.c_int, [_vp, C.c_int, _cp]),
    "rln_star_value": (_cp, [_vp, C.c_int, C.c_long, C.c_int]),
    "rln_star_column_double": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "rln_star_column_long": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_long)]),
    "rln_star_add_block": (C.c_int, [_vp, _cp, C.c_int, C.POINTER(C.c_int)]),
    "rln_star_add_column": (C.c_int, [_vp, C.c_int, _cp, _cp, C.POINTER(C.c_int)]),
    "rln_star_add_rows": (C.c_int, [_vp, C.c_int, C.c_long]),
    "rln_star_set_value": (C.c_int, [_vp, C.c_int, C.c_long, C.c_int, _cp]),
    "rln_star_write": (C.c_int, [_vp, _cp]),
    "rln_mrc_read_info": (C.c_int, [_cp, C.POINTER(MrcInfo)]),
    "rln_mrc_read": (C.c_int, [_cp, C.c_long, C.c_long, C.POINTER(C.c_float)]),
    "rln_mrc_write": (C.c_int, [_cp, C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int, C.c_double]),
    "rln_star_read_images": (C.c_int, [_vp, C.c_int, C.c_int, C.c_long, C.c_long, _cp, C.c_int, C.c_int,
                                       C.POINTER(C.c_float)]),
}
for _name, (_res, _args) in _SIG.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def _b(s):
    return s.encode() if isinstance(s, str) else s


class StarFile:
    """Every data block of a STAR file (MetaDataTable::readAll, src/metadata_table.cpp:1275-1345)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def read(cls, path):
        h = _vp()
        _check(lib.rln_star_read(_b(str(path)), C.byref(h)))
        return cls(h)

    @classmethod
    def parse(cls, text):
        data = _b(text)
        h = _vp()
        _check(lib.rln_star_parse(data, len(data), C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls):
        h = _vp()
        _check(lib.rln_star_create(C.byref(h)))
        return cls(h)

    def close(self):
        if self._h:
            lib.rln_star_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def version(self):
        return lib.rln_star_version(self._h)

    @property
    def block_names(self):
        return [lib.rln_star_block_name(self._h, b).decode() for b in range(lib.rln_star_block_count(self._h))]

    def block(self, name_or_index):
        if isinstance(name_or_index, int):
            b = name_or_index
        else:
            b = lib.rln_star_find_block(self._h, _b(name_or_index))
            if b < 0:
                raise KeyError(f"no data_{name_or_index} block")
        return StarBlock(self, b)

    def add_block(self, name, is_list=False):
        b = C.c_int()
        _check(lib.rln_star_add_block(self._h, _b(name), int(is_list), C.byref(b)))
        return StarBlock(self, b.value)

    def write(self, path):
        _check(lib.rln_star_write(self._h, _b(str(path))))


class StarBlock:
    """One data block: a loop_ table or a label-value list (one row)."""

    def __init__(self, star, index):
        self.star, self.index = star, index

    @property
    def _h(self):
        return self.star._h

    @property
    def name(self):
        return lib.rln_star_block_name(self._h, self.index).decode()

    @property
    def is_list(self):
        return bool(_check_nonneg(lib.rln_star_block_is_list(self._h, self.index)))

    @property
    def labels(self):
        n = _check_nonneg(lib.rln_star_columns(self._h, self.index))
        return [lib.rln_star_label(self._h, self.index, c).decode() for c in range(n)]

    def __len__(self):
        return _check_nonneg(lib.rln_star_rows(self._h, self.index))

    @property
    def shape(self):
        return len(self), len(self.labels)

    def column_index(self, label):
        c = lib.rln_star_find_label(self._h, self.index, _b(label))
        if c < 0:
            raise KeyError(f"no label {label} in data_{self.name}")
        return c

    def __contains__(self, label):
        return lib.rln_star_find_label(self._h, self.index, _b(label)) >= 0

    def value(self, row, label):
        v = lib.rln_star_value(self._h, self.index, row, self.column_index(label))
        if v is None:
            raise IndexError(row)
        return v.decode()

    def strings(self, label):
        c = self.column_index(label)
        return [lib.rln_star_value(self._h, self.index, r, c).decode() for r in range(len(self))]

    def floats(self, label):
        out = np.zeros(len(self), np.float64)
        _check(lib.rln_star_column_double(self._h, self.index, self.column_index(label), _ptr(out, C.c_double)))
        return out

    def ints(self, label):
        out = np.zeros(len(self), np.int64)
        _check(lib.rln_star_column_long(self._h, self.index, self.column_index(label), _ptr(out, C.c_long)))
        return out

    def add_column(self, label, default=""):
        c = C.c_int()
        _check(lib.rln_star_add_column(self._h, self.index, _b(label), _b(str(default)), C.byref(c)))
        return c.value

    def add_rows(self, n):
        _check(lib.rln_star_add_rows(self._h, self.index, n))

    def set_value(self, row, label, value):
        _check(lib.rln_star_set_value(self._h, self.index, row, self.column_index(label), _b(str(value))))

    def set_column(self, label, values, fmt="{}"):
        if label not in self:
            self.add_column(label)
        c = self.column_index(label)
        for r, v in enumerate(values):
            _check(lib.rln_star_set_value(self._h, self.index, r, c, _b(fmt.format(v))))

    def read_images(self, box, label="rlnImageName", first=0, count=None, base_dir=None, threads=0):
        """The images the column names, as float32 [count][box][box] (rln_star_read_images)."""
        n = len(self) - first if count is None else count
        out = np.zeros((n, box, box), np.float32)
        _check(lib.rln_star_read_images(self._h, self.index, self.column_index(label), first, n,
                                        _b(str(base_dir)) if base_dir else None, box, threads,
                                        _ptr(out, C.c_float)))
        return out


def _check_nonneg(v):
    if v < 0:
        _check(v)
    return v


def mrc_info(path):
    """Header of an MRC file (parseMRCHeader, src/rwMRC.h:143-245)."""
    info = MrcInfo()
    _check(lib.rln_mrc_read_info(_b(str(path)), C.byref(info)))
    return info


def read_mrc(path, first=0, count=None):
    """Sections [first, first + count) as float32 [count][ny][nx]."""
    info = mrc_info(path)
    n = info.nz - first if count is None else count
    out = np.zeros((n, info.ny, info.nx), np.float32)
    _check(lib.rln_mrc_read(_b(str(path)), first, n, _ptr(out, C.c_float)))
    return out


def write_mrc(path, data, pixel_size=0.0):
    """A mode-2 MRC file of a 2-D image, a stack or a volume [nz][ny][nx] (writeMRC, src/rwMRC.h:280-500)."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    _check(lib.rln_mrc_write(_b(str(path)), _ptr(a, C.c_float), nx, ny, nz, float(pixel_size)))


# ------------------------------------------------------------------------------------------------ particle sets
# (optics label, key, default) and (particle label, key, default) read as numbers; the defaults are the reference's
# when a label is absent (MlOptimiser::getMetaAndImageDataSubset, src/ml_optimiser.cpp:10420-10520).
_OPTICS = (("rlnOpticsGroup", "group", None), ("rlnImagePixelSize", "pixel_size", None),
           ("rlnImageSize", "image_size", None), ("rlnVoltage", "voltage", None),
           ("rlnSphericalAberration", "cs", None), ("rlnAmplitudeContrast", "q0", None),
           ("rlnCtfDataAreCtfPremultiplied", "ctf_premultiplied", 0.0))
_PARTICLE = (("rlnDefocusU", "defocus_u", 0.0), ("rlnDefocusV", "defocus_v", None),
             ("rlnDefocusAngle", "defocus_angle", 0.0), ("rlnCtfBfactor", "bfactor", 0.0),
             ("rlnCtfScalefactor", "kfactor", 1.0), ("rlnPhaseShift", "phase_shift", 0.0),
             ("rlnNormCorrection", "norm_correction", 1.0), ("rlnOriginXAngst", "origin_x_A", 0.0),
             ("rlnOriginYAngst", "origin_y_A", 0.0), ("rlnAngleRot", "rot", 0.0), ("rlnAngleTilt", "tilt", 0.0),
             ("rlnAnglePsi", "psi", 0.0), ("rlnOpticsGroup", "optics_group", 1.0),
             ("rlnRandomSubset", "random_subset", 0.0), ("rlnOriginXPriorAngst", "origin_x_prior_A", 0.0),
             ("rlnOriginYPriorAngst", "origin_y_prior_A", 0.0))


class ParticleSet:
    """Particles of a STAR file (data_optics + data_particles, the format since RELION 3.1), their images and the
    per-particle metadata the image preparation and the E-step take."""

    def __init__(self, optics, columns, image_names, images):
        self.optics, self.columns, self.image_names, self.images = optics, columns, image_names, images

    def __len__(self):
        return len(self.image_names)

    def optics_of(self, i):
        return self.optics[int(self.columns["optics_group"][i])]

    def imgprep_metas(self):
        """ImageMeta dicts: offsets in pixels (rlnOrigin*Angst / pixel size), CTF, norm correction, part_id = row."""
        c = self.columns
        out = []
        for i in range(len(self)):
            px = self.optics_of(i)["pixel_size"]
            out.append(dict(norm_correction=c["norm_correction"][i], old_offset_x=c["origin_x_A"][i] / px,
                            old_offset_y=c["origin_y_A"][i] / px, defocus_u=c["defocus_u"][i],
                            defocus_v=c["defocus_v"][i], defocus_angle=c["defocus_angle"][i],
                            bfactor=c["bfactor"][i], kfactor=c["kfactor"][i], phase_shift=c["phase_shift"][i],
                            part_id=i))
        return out


def read_particles(path, base_dir=None, load_images=True, threads=0, particles_block="particles",
                   optics_block="optics"):
    """A ParticleSet from a STAR file: optics groups, per-particle CTF / offsets / angles (the reference's defaults
    for absent labels: rlnDefocusV = rlnDefocusU, scale factor 1, norm correction 1, others 0) and, with
    load_images, the images named by rlnImageName (every optics group's box must be the same)."""
    star = StarFile.read(path)
    names = star.block_names
    if optics_block not in names:
        raise ValueError(f"{path}: no data_{optics_block} block (STAR files before RELION 3.1 are not read)")
    ob, pb = star.block(optics_block), star.block(particles_block)
    optics = {}
    ocols = {key: (ob.floats(lab) if lab in ob else np.full(len(ob), dflt)) for lab, key, dflt in _OPTICS
             if lab in ob or dflt is not None}
    missing = [lab for lab, key, dflt in _OPTICS if key not in ocols]
    if missing:
        raise ValueError(f"{path}: optics table without {missing}")
    for r in range(len(ob)):
        g = {k: float(v[r]) for k, v in ocols.items()}
        g["image_size"] = int(g["image_size"])
        g["ctf_premultiplied"] = bool(g["ctf_premultiplied"])
        optics[int(g.pop("group"))] = g
    cols = {}
    for lab, key, dflt in _PARTICLE:
        if lab in pb:
            cols[key] = pb.floats(lab)
        elif dflt is not None:
            cols[key] = np.full(len(pb), dflt)
    if "defocus_v" not in cols:
        cols["defocus_v"] = cols["defocus_u"].copy()
    cols["optics_group"] = cols["optics_group"].astype(np.int64)
    unknown = set(cols["optics_group"].tolist()) - set(optics)
    if unknown:
        raise ValueError(f"{path}: particles refer to optics groups {sorted(unknown)} not in data_{optics_block}")
    image_names = pb.strings("rlnImageName")
    images = None
    if load_images:
        boxes = {optics[g]["image_size"] for g in set(cols["optics_group"].tolist())}
        if len(boxes) != 1:
            raise ValueError(f"{path}: optics groups of different image sizes {sorted(boxes)}")
        images = pb.read_images(boxes.pop(), base_dir=base_dir, threads=threads)
    return ParticleSet(optics, cols, image_names, images)


def write_particles(path, image_names, metas, optics, angles=None):
    """A STAR file with data_optics (one dict per group: group, pixel_size, image_size, voltage, cs, q0) and
    data_particles (rlnImageName, rlnOpticsGroup = 1 unless a meta gives 'optics_group', CTF, norm correction,
    origins in Angstrom from the ImageMeta offsets in pixels, angles)."""
    s = StarFile.create()
    o = s.add_block("optics")
    o.add_rows(len(optics))
    for lab, key, fmt in (("rlnOpticsGroupName", None, None), ("rlnOpticsGroup", "group", "{:d}"),
                          ("rlnImagePixelSize", "pixel_size", "{:.6f}"), ("rlnImageSize", "image_size", "{:d}"),
                          ("rlnImageDimensionality", None, None), ("rlnVoltage", "voltage", "{:.6f}"),
                          ("rlnSphericalAberration", "cs", "{:.6f}"), ("rlnAmplitudeContrast", "q0", "{:.6f}")):
        if key is None:
            vals = [f"opticsGroup{g['group']}" for g in optics] if lab == "rlnOpticsGroupName" else [2] * len(optics)
            o.set_column(lab, vals)
        else:
            o.set_column(lab, [int(g[key]) if fmt == "{:d}" else float(g[key]) for g in optics], fmt)
    px = {int(g["group"]): g["pixel_size"] for g in optics}
    p = s.add_block("particles")
    p.add_rows(len(image_names))
    p.set_column("rlnImageName", image_names)
    grp = [int(m.get("optics_group", 1)) for m in metas]
    p.set_column("rlnOpticsGroup", grp, "{:d}")
    for lab, key in (("rlnDefocusU", "defocus_u"), ("rlnDefocusV", "defocus_v"), ("rlnDefocusAngle", "defocus_angle"),
                     ("rlnCtfBfactor", "bfactor"), ("rlnCtfScalefactor", "kfactor"), ("rlnPhaseShift", "phase_shift"),
                     ("rlnNormCorrection", "norm_correction")):
        p.set_column(lab, [float(m[key]) for m in metas], "{!r}")  # shortest text that reads back exactly
    p.set_column("rlnOriginXAngst", [float(m["old_offset_x"] * px[g]) for m, g in zip(metas, grp)], "{!r}")
    p.set_column("rlnOriginYAngst", [float(m["old_offset_y"] * px[g]) for m, g in zip(metas, grp)], "{!r}")
    if angles is not None:
        for j, lab in enumerate(("rlnAngleRot", "rlnAngleTilt", "rlnAnglePsi")):
            p.set_column(lab, [float(a[j]) for a in angles], "{!r}")
    s.write(path)
