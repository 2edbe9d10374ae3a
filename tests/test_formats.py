"""STAR and MRC formats (include/rln_formats.h through relion_amd.formats), on the CPU.

Pins: the reference's own STAR test files (tests/golden/star, from src/tomography_python_programs/_tests, whose
test_metadata_model.py asserts a 41 x 29 tilt-series table and 5 tilt series) parsed and compared with an
independent whitespace split of their rows; the tokenizer's rules (quotes, comments, blocks, lists, errors) on
hand-written text; MRC headers decoded at the MRC2014 byte offsets with numpy; every supported mode converted by
numpy from files numpy wrote.
"""
import os

import numpy as np
import pytest

import relion_amd as ra
from relion_amd import formats as F

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "star")


def _rows_by_split(path):
    """Labels and rows of the single loop of a STAR file without quoted values, by str.split."""
    labels, rows, in_loop = [], [], False
    for line in open(path).read().split("\n"):
        t = line.strip()
        if t.startswith("loop_"):
            in_loop = True
            continue
        if not in_loop:
            continue
        if t.startswith("_"):
            labels.append(t[1:].split()[0])
        elif t:
            rows.append(t.split())
        elif rows:
            break
    return labels, rows


@pytest.mark.parametrize("name,block,shape", [("220728_TS_01.star", "220728_TS_01", (41, 29)),
                                              ("tilt_series.star", "global", (5, 14))])
def test_reference_star_files(name, block, shape):
    path = os.path.join(GOLDEN, name)
    s = F.StarFile.read(path)
    assert s.version == 30001
    assert s.block_names == [block]
    b = s.block(block)
    assert not b.is_list
    assert b.shape == shape
    labels, rows = _rows_by_split(path)
    assert b.labels == labels
    for j, lab in enumerate(labels):
        assert b.strings(lab) == [r[j] for r in rows]


def test_reference_tilt_series_set():
    s = F.StarFile.read(os.path.join(GOLDEN, "tilt_series.star"))
    g = s.block("global")
    assert g.strings("rlnTomoName") == ["220728_TS_01", "220728_TS_03", "220728_TS_43", "220728_TS_45",
                                        "220728_TS_54"]
    assert np.all(g.floats("rlnVoltage") == 300.0)
    assert np.all(g.ints("rlnTomoSizeZ") == 3000)
    ts = F.StarFile.read(os.path.join(GOLDEN, "220728_TS_01.star")).block(0)
    tilt = ts.floats("rlnTomoNominalStageTiltAngle")
    assert tilt[:4].tolist() == [0.001, 3.00113, -2.99863, -5.99876]
    assert np.array_equal(ts.ints("rlnTomoTiltMovieIndex"), np.arange(1, 42))
    with pytest.raises(ra.RelionError):
        ts.floats("rlnMicrographName")  # file names are not numbers


TEXT = """
# version 30001

data_optics

loop_
_rlnOpticsGroup #1
_rlnOpticsGroupName #2
_rlnVoltage #3
1 "group one"   300.0
2  'two  spaces' 200

  # a comment line after the table (a comment right under the rows would be a short row: an error)

data_general

_rlnImageSizeX      64
_rlnComment  "a b"   # trailing comment
# comment
_rlnEmpty ""

data_mixed
_rlnListValue 7
loop_
_rlnIgnored #1
1
2

data_particles
loop_
_rlnImageName #1
_rlnDefocusU #2
000001@stack.mrcs 10000.5
000002@stack.mrcs 12000
"""


def test_tokenizer_blocks_lists_and_quotes():
    s = F.StarFile.parse(TEXT)
    assert s.version == 30001
    assert s.block_names == ["optics", "general", "mixed", "particles"]
    o = s.block("optics")
    assert o.shape == (2, 3)
    # quoted values keep their text; the loop's line simplification collapses runs of spaces inside quotes
    assert o.strings("rlnOpticsGroupName") == ["group one", "two spaces"]
    assert o.floats("rlnVoltage").tolist() == [300.0, 200.0]
    g = s.block("general")
    assert g.is_list and g.shape == (1, 3)
    assert g.value(0, "rlnComment") == "a b" and g.value(0, "rlnEmpty") == ""
    m = s.block("mixed")  # a list followed by a loop keeps the list
    assert m.is_list and m.labels == ["rlnListValue"] and m.value(0, "rlnListValue") == "7"
    p = s.block("particles")
    assert p.strings("rlnImageName") == ["000001@stack.mrcs", "000002@stack.mrcs"]
    assert p.floats("rlnDefocusU").tolist() == [10000.5, 12000.0]
    with pytest.raises(KeyError):
        s.block("missing")


@pytest.mark.parametrize("text", [
    "data_a\nloop_\n_rlnA #1\n_rlnB #2\n1 2 3\n",            # more values than labels
    "data_a\nloop_\n_rlnA #1\n_rlnB #2\n_rlnC #3\n1 2\n",    # fewer values than three labels
    "data_a\nloop_\n_rlnA #1\n\"open quote\n",                # unclosed quote
    "data_a\r\nloop_\r\n_rlnA #1\r\n1\r\n",                   # CR+LF
    "data_a\n_rlnA\n",                                          # list label without a value
])
def test_star_errors(text):
    with pytest.raises(ra.RelionError):
        F.StarFile.parse(text)


def test_two_label_rows_may_be_short():
    # the reference lets rows of a two-label table miss values ("fn_mtf <empty>")
    b = F.StarFile.parse("data_a\nloop_\n_rlnA #1\n_rlnB #2\n1\n3 4\n").block(0)
    assert b.strings("rlnA") == ["1", "3"] and b.strings("rlnB") == ["", "4"]


def test_star_write_round_trip(tmp_path):
    s = F.StarFile.create()
    t = s.add_block("particles")
    t.add_column("rlnImageName")
    t.add_column("rlnDefocusU", "0")
    t.add_column("rlnComment")
    t.add_rows(3)
    names = ["000001@a.mrcs", "000002@a.mrcs", "000001@b c.mrcs"]
    t.set_column("rlnImageName", names)
    t.set_column("rlnDefocusU", [1.5, 2.25, 3.0])
    t.set_column("rlnComment", ["", "x", "'q"])
    lst = s.add_block("general", is_list=True)
    lst.add_column("rlnNrClasses", "4")
    path = tmp_path / "out.star"
    s.write(path)
    text = path.read_text()
    assert text.startswith("\n# version 50001\n\ndata_particles\n\nloop_ \n_rlnImageName #1 \n")
    assert "\n    1.5 " not in text and "       1.5 " in text  # right-aligned in 10 characters
    r = F.StarFile.read(path)
    assert r.block_names == ["particles", "general"]
    b = r.block("particles")
    assert b.strings("rlnImageName") == names
    assert b.floats("rlnDefocusU").tolist() == [1.5, 2.25, 3.0]
    assert b.strings("rlnComment") == ["", "x", "'q"]
    assert r.block("general").is_list and r.block("general").ints("rlnNrClasses").tolist() == [4]


def _header(path):
    raw = open(path, "rb").read(1024)
    i4 = np.frombuffer(raw, "<i4")
    f4 = np.frombuffer(raw, "<f4")
    return raw, i4, f4


def test_mrc_write_header_at_spec_offsets(tmp_path):
    rng = np.random.default_rng(0)
    vol = rng.normal(size=(5, 6, 7)).astype(np.float32)
    path = tmp_path / "v.mrc"
    F.write_mrc(path, vol, pixel_size=1.25)
    raw, i4, f4 = _header(path)
    assert i4[:4].tolist() == [7, 6, 5, 2]                       # nx ny nz mode
    assert i4[7:10].tolist() == [7, 6, 5]                        # mx my mz
    assert np.allclose(f4[10:13], [1.25 * 7, 1.25 * 6, 1.25 * 5])  # cell a b c
    assert f4[13:16].tolist() == [90.0, 90.0, 90.0]
    assert i4[16:19].tolist() == [1, 2, 3]                       # mapc mapr maps
    assert f4[19] == vol.min() and f4[20] == vol.max()
    assert np.isclose(f4[21], vol.mean(dtype=np.float64), rtol=1e-6)
    assert i4[23] == 0                                           # nsymbt
    assert raw[208:212] == b"MAP " and raw[212:214] == bytes([68, 65])
    assert np.isclose(f4[54], vol.std(ddof=1, dtype=np.float64), rtol=1e-6)
    assert i4[55] == 1
    data = np.frombuffer(open(path, "rb").read()[1024:], "<f4").reshape(5, 6, 7)
    assert np.array_equal(data, vol)
    info = F.mrc_info(path)
    assert (info.nx, info.ny, info.nz, info.mode, info.swapped) == (7, 6, 5, 2, 0)
    assert info.pixel_size == pytest.approx(1.25)
    assert np.array_equal(F.read_mrc(path), vol)
    assert np.array_equal(F.read_mrc(path, 2, 2), vol[2:4])


def _write_raw(path, nx, ny, nz, mode, payload, nsymbt=0, big=False, cell=None):
    e = ">" if big else "<"
    h = np.zeros(256, np.int32).view(np.uint8)
    ints = np.zeros(256, e + "i4")
    flts = ints.view(e + "f4")
    ints[:4] = [nx, ny, nz, mode]
    ints[7:10] = [nx, ny, nz]
    flts[10:13] = cell if cell is not None else [nx, ny, nz]
    ints[23] = nsymbt
    h = ints.tobytes()
    with open(path, "wb") as f:
        f.write(h)
        f.write(b"\x07" * nsymbt)
        f.write(payload)


@pytest.mark.parametrize("mode,dtype", [(0, "i1"), (1, "<i2"), (6, "<u2"), (12, "<f2")])
def test_mrc_modes(tmp_path, mode, dtype):
    rng = np.random.default_rng(mode)
    a = (rng.normal(size=(3, 4, 6)) * 50).astype(dtype)
    if mode == 6:
        a = np.abs(a.astype(np.int64)).astype(dtype)
    path = tmp_path / f"m{mode}.mrcs"
    _write_raw(path, 6, 4, 3, mode, a.tobytes(), nsymbt=96)
    out = F.read_mrc(path)
    assert np.array_equal(out, a.astype(np.float32))
    assert F.mrc_info(path).nsymbt == 96


def test_mrc_half_floats_special_values(tmp_path):
    vals = np.array([0.0, -0.0, 1.0, -2.5, 65504.0, 6.1e-5, 5.96e-8, np.inf, -np.inf], np.float16)
    path = tmp_path / "h.mrc"
    _write_raw(path, len(vals), 1, 1, 12, vals.tobytes())
    out = F.read_mrc(path)[0, 0]
    assert np.array_equal(out, vals.astype(np.float32))
    nan = np.array([np.nan, np.nan], np.float16)
    _write_raw(path, 2, 1, 1, 12, nan.tobytes())
    assert np.all(np.isnan(F.read_mrc(path)))


def test_mrc_4bit_and_big_endian(tmp_path):
    rng = np.random.default_rng(3)
    px = rng.integers(0, 16, size=(2, 4, 6)).astype(np.uint8)
    flat = px.reshape(-1)
    packed = (flat[0::2] | (flat[1::2] << 4)).astype(np.uint8)  # low nibble first
    path = tmp_path / "n.mrc"
    _write_raw(path, 6, 4, 2, 101, packed.tobytes())
    assert np.array_equal(F.read_mrc(path), px.astype(np.float32))
    _write_raw(path, 5, 3, 1, 101, b"\x00" * 8)
    with pytest.raises(ra.RelionError):
        F.read_mrc(path)  # odd nx * ny
    f = rng.normal(size=(2, 3, 5)).astype(">f4")
    _write_raw(path, 5, 3, 2, 2, f.tobytes(), big=True, cell=[10.0, 6.0, 4.0])
    info = F.mrc_info(path)
    assert info.swapped == 1 and (info.nx, info.ny, info.nz) == (5, 3, 2)
    assert info.pixel_size == pytest.approx(2.0)
    assert np.array_equal(F.read_mrc(path), f.astype(np.float32))
    s = rng.integers(-300, 300, size=(1, 3, 5)).astype(">i2")
    _write_raw(path, 5, 3, 1, 1, s.tobytes(), big=True)
    assert np.array_equal(F.read_mrc(path), s.astype(np.float32))


def test_mrc_refused(tmp_path):
    path = tmp_path / "c.mrc"
    _write_raw(path, 4, 4, 1, 4, b"\x00" * 128)
    with pytest.raises(ra.RelionError):
        F.read_mrc(path)  # complex mode
    _write_raw(path, 4, 4, 2, 2, b"\x00" * 64)
    with pytest.raises(ra.RelionError):
        F.read_mrc(path)  # file shorter than its header says
    F.write_mrc(path, np.zeros((2, 4, 4), np.float32))
    with pytest.raises(ra.RelionError):
        F.read_mrc(path, 1, 2)  # beyond the stack


def test_particle_images_from_star(tmp_path):
    rng = np.random.default_rng(5)
    box = 16
    os.makedirs(tmp_path / "Particles")
    stacks = {n: rng.normal(size=(k, box, box)).astype(np.float32) for n, k in (("s1", 5), ("s2", 3), ("s3", 1))}
    for n, a in stacks.items():
        F.write_mrc(tmp_path / "Particles" / f"{n}.mrcs", a, pixel_size=1.5)
    refs = [("s1", 3), ("s2", 1), ("s1", 4), ("s1", 5), ("s3", None), ("s2", 3), ("s1", 1)]
    s = F.StarFile.create()
    t = s.add_block("particles")
    t.add_column("rlnImageName")
    t.add_rows(len(refs))
    t.set_column("rlnImageName", [f"Particles/{n}.mrcs" if i is None else f"{i:06d}@Particles/{n}.mrcs"
                                  for n, i in refs])
    expect = np.stack([stacks[n][0 if i is None else i - 1] for n, i in refs])
    out = t.read_images(box, base_dir=tmp_path, threads=3)
    assert np.array_equal(out, expect)
    assert np.array_equal(t.read_images(box, first=2, count=3, base_dir=tmp_path), expect[2:5])
    with pytest.raises(ra.RelionError):
        t.read_images(box + 2, base_dir=tmp_path)  # wrong box
    t.set_value(0, "rlnImageName", "000009@Particles/s3.mrcs")
    with pytest.raises(ra.RelionError):
        t.read_images(box, base_dir=tmp_path)  # index beyond the stack
    t.set_value(0, "rlnImageName", "x@Particles/s3.mrcs")
    with pytest.raises(ra.RelionError):
        t.read_images(box, base_dir=tmp_path)


def test_particle_set_round_trip_and_defaults(tmp_path):
    rng = np.random.default_rng(8)
    box, n = 12, 6
    imgs = rng.normal(size=(n, box, box)).astype(np.float32)
    F.write_mrc(tmp_path / "p.mrcs", imgs, pixel_size=1.7)
    metas = [dict(norm_correction=rng.uniform(0.5, 2), old_offset_x=rng.uniform(-3, 3), old_offset_y=rng.uniform(-3, 3),
                  defocus_u=rng.uniform(8000, 20000), defocus_v=rng.uniform(8000, 20000),
                  defocus_angle=rng.uniform(0, 180), bfactor=rng.uniform(0, 50), kfactor=rng.uniform(0.8, 1.2),
                  phase_shift=rng.uniform(0, 20)) for _ in range(n)]
    optics = [dict(group=1, pixel_size=1.7, image_size=box, voltage=300.0, cs=2.7, q0=0.1)]
    names = [f"{i + 1:06d}@p.mrcs" for i in range(n)]
    F.write_particles(tmp_path / "p.star", names, metas, optics, angles=rng.uniform(0, 180, size=(n, 3)))
    ps = F.read_particles(tmp_path / "p.star", base_dir=tmp_path)
    assert np.array_equal(ps.images, imgs)
    assert ps.optics == {1: dict(pixel_size=1.7, image_size=box, voltage=300.0, cs=2.7, q0=0.1, ctf_premultiplied=False)}
    for m, r in zip(metas, ps.imgprep_metas()):
        for k, v in m.items():
            assert r[k] == pytest.approx(v, rel=1e-12, abs=1e-12), k
    # absent labels take the reference's defaults
    s = F.StarFile.parse("data_optics\nloop_\n_rlnOpticsGroup\n_rlnImagePixelSize\n_rlnImageSize\n_rlnVoltage\n"
                         "_rlnSphericalAberration\n_rlnAmplitudeContrast\n1 1.0 12 200 2.0 0.07\n\n"
                         "data_particles\nloop_\n_rlnImageName\n_rlnDefocusU\n000001@p.mrcs 15000\n")
    s.write(tmp_path / "d.star")
    d = F.read_particles(tmp_path / "d.star", base_dir=tmp_path).imgprep_metas()[0]
    assert d["defocus_v"] == 15000 and d["kfactor"] == 1.0 and d["norm_correction"] == 1.0
    assert d["bfactor"] == 0.0 and d["phase_shift"] == 0.0 and d["old_offset_x"] == 0.0
    with pytest.raises(ValueError):
        F.read_particles(_no_optics(tmp_path), base_dir=tmp_path)


def _no_optics(tmp_path):
    F.StarFile.parse("data_particles\nloop_\n_rlnImageName\n000001@p.mrcs\n").write(tmp_path / "old.star")
    return tmp_path / "old.star"
