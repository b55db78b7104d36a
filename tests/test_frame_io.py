"""Frame files of the reference (host-only entry points; no GPU needed).

* .xyz: the example of FSR/data_format.txt (header and the first particle
  lines, copied as data) parses to the FP32 values of its decimal strings;
  round trip of a generated frame; blank lines skipped; malformed input and
  capacity errors.
* .obj: the exact text layout of GPU/ReconstructorGPUOurs.cpp:20-100 (all v,
  then all vn, then 1-based "f i//i j//j k//k " faces, %g numbers).
* constants of the reconstructor classes.
"""
import os
import subprocess

import numpy as np
import pytest

from fluidsurfacereconstructiongpu_amd import _lib, frame_io

EXAMPLE = """-0.04 -0.04 -0.04
3.04 2.04 1.54
0.036
0.01
0.00403914
0.0199628 0.02 0.0199924 429.723
0.0397894 0.0199325 0.02 577.938
0.0598905 0.02 0.0199746 587.701
0.0799356 0.02 0.0199719 591.818
"""


def test_reference_example_parses(tmp_path):
    p = tmp_path / "frame_000000.xyz"
    p.write_text(EXAMPLE + "\n\n")
    head, pos, dens = frame_io.read_xyz(str(p))
    f = np.float32
    assert head["aabb_min"] == (f(-0.04), f(-0.04), f(-0.04))
    assert head["aabb_max"] == (f(3.04), f(2.04), f(1.54))
    assert head["kernel_radius"] == f(0.036)
    assert head["particle_radius"] == f(0.01)
    assert head["particle_mass"] == f(0.00403914)
    assert pos.shape == (4, 3)
    assert pos[0].tolist() == [f(0.0199628), f(0.02), f(0.0199924)]
    assert dens.tolist() == [f(429.723), f(577.938), f(587.701), f(591.818)]


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(5)
    pos = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    dens = rng.uniform(900, 1100, 500).astype(np.float32)
    p = str(tmp_path / "f.xyz")
    frame_io.write_xyz(p, (-1.1, -1.1, -1.1), (1.1, 1.1, 1.1), 0.04, 0.005, 1e-6, pos, dens)
    head, pos2, dens2 = frame_io.read_xyz(p)
    assert np.array_equal(pos, pos2) and np.array_equal(dens, dens2)
    assert head["kernel_radius"] == np.float32(0.04)


def test_errors(tmp_path):
    lib = _lib.load()
    bad = tmp_path / "bad.xyz"
    bad.write_text("0 0 0\n1 1 1\n0.1\n0.01\n")
    with pytest.raises(_lib.FsrError):
        frame_io.read_xyz(str(bad))
    bad.write_text(EXAMPLE + "1 2 x 4\n")
    with pytest.raises(_lib.FsrError):
        frame_io.read_xyz(str(bad))
    ok = tmp_path / "ok.xyz"
    ok.write_text(EXAMPLE)
    import ctypes
    h = _lib.XyzHeader()
    n = ctypes.c_uint32()
    pos = np.empty((2, 3), np.float32)
    dens = np.empty(2, np.float32)
    st = lib.fsr_xyz_read(str(ok).encode(), ctypes.byref(h), pos.ctypes.data, dens.ctypes.data,
                          2, ctypes.byref(n))
    assert st == 3  # FSR_ERR_CAPACITY
    assert lib.fsr_xyz_info(b"/nonexistent/x.xyz", ctypes.byref(h), ctypes.byref(n)) == 1


def test_obj_soup_layout(tmp_path):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.5, 0.25, 1e-7], [2, 2, 2], [3, 3, 3]],
                   np.float32)
    nrm = np.tile(np.array([[0, 0, 1]], np.float32), (6, 1))
    p = str(tmp_path / "m.obj")
    frame_io.write_obj_soup(p, pos, nrm)
    lines = open(p).read().splitlines()
    assert lines[0] == "v 0 0 0"
    assert lines[3] == "v 0.5 0.25 1e-07"
    assert lines[6] == "vn 0 0 1"
    assert lines[12] == "f 1//1 2//2 3//3 "
    assert lines[13] == "f 4//4 5//5 6//6 "
    assert len(lines) == 14


def test_obj_indexed_layout(tmp_path):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    idx = np.array([0, 1, 2, 2, 1, 3], np.uint32)
    p = str(tmp_path / "m.obj")
    frame_io.write_obj_indexed(p, pos, None, idx)
    lines = open(p).read().splitlines()
    assert lines[:4] == ["v 0 0 0", "v 1 0 0", "v 0 1 0", "v 1 1 0"]
    assert lines[4:] == ["f 1//1 2//2 3//3 ", "f 3//3 2//2 4//4 "]
    with pytest.raises(_lib.FsrError):
        frame_io.write_obj_indexed(p, pos, None, np.array([0, 1, 4], np.uint32))


def _g(x) -> str:
    # C's printf("%g") of the float's exact value, as `file << float` writes
    return "%g" % float(x)


def test_obj_text_matches_printf_g(tmp_path):
    """The .obj writers format on several threads with std::to_chars; their
    text must equal printf's "%g" (6 significant digits) line for line, over
    magnitudes from denormals to FLT_MAX, signed zeros and infinities, values
    that round up a digit (999999.5), and more lines than one thread takes."""
    rng = np.random.default_rng(5)
    n = 3 * 4000
    with np.errstate(over="ignore"):
        pos = (rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-45, 39, size=(n, 3)))
    pos = pos.astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-45, 3.4028235e38, 1.17549435e-38,
                        0.1, 123456.5, 999999.5, 1e6, 1e-5, 1e-4, 0.000123456789], np.float32)
    pos.ravel()[:special.size] = special
    nrm = rng.random((n, 3)).astype(np.float32)
    p = str(tmp_path / "s.obj")
    frame_io.write_obj_soup(p, pos, nrm)
    exp = ["v %s %s %s\n" % tuple(_g(x) for x in v) for v in pos]
    exp += ["vn %s %s %s\n" % tuple(_g(x) for x in v) for v in nrm]
    exp += ["f %d//%d %d//%d %d//%d \n" % (i, i, i + 1, i + 1, i + 2, i + 2)
            for i in range(1, n + 1, 3)]
    assert open(p).read() == "".join(exp)
    idx = rng.integers(0, n, size=3 * 3000).astype(np.uint32)
    frame_io.write_obj_indexed(p, pos, None, idx)
    exp = ["v %s %s %s\n" % tuple(_g(x) for x in v) for v in pos]
    for t in range(0, idx.size, 3):
        a, b, c = (int(idx[t + k]) + 1 for k in range(3))
        exp.append("f %d//%d %d//%d %d//%d \n" % (a, a, b, b, c, c))
    assert open(p).read() == "".join(exp)


def _nearest_f32(text: str) -> np.float32:
    """The float32 nearest to the decimal `text` (ties to even), by exact
    rational comparison of the three candidates around a first guess."""
    from fractions import Fraction

    neg = text.lstrip().startswith("-")
    x = abs(Fraction(text))
    big = np.finfo(np.float32).max
    with np.errstate(over="ignore"):
        g = np.float32(float(x)) if abs(x) <= Fraction(2) ** 128 else np.float32(np.sign(x) * big)
        cands = [np.nextafter(g, np.float32(-np.inf)), g, np.nextafter(g, np.float32(np.inf))]

    def exact(c):
        # infinity stands for the first value past FLT_MAX's binade, 2^128
        return Fraction(float(c)) if np.isfinite(c) else Fraction(int(np.sign(c)) * 2 ** 128)

    def key(c):
        lsb = int(np.frombuffer(np.float32(c).tobytes(), np.uint32)[0]) & 1
        return (abs(exact(c) - x), lsb)

    best = np.float32(min(cands, key=key))
    return -best if neg else best  # the sign survives a zero result


def test_xyz_numbers_round_correctly(tmp_path):
    """The reader's numbers are the correctly rounded float32 of their decimal
    text (as `stream >> float` gives): 9-digit round trips of random floats
    over many binades, and strings near rounding boundaries: exact midpoints
    of two floats (ties to even), decimals a hair above and below a midpoint,
    more than 19 significant digits, large exponents, subnormals, values that
    underflow to zero or overflow to infinity (from_chars leaves those unset;
    the reader takes strtof's value), signed zeros, leading zeros and a bare
    exponent marker."""
    from fractions import Fraction

    rng = np.random.default_rng(11)
    vals = (rng.normal(size=4000) * 10.0 ** rng.integers(-30, 30, size=4000)).astype(np.float32)
    texts = ["%.9g" % float(v) for v in vals]
    expect = list(vals)
    for v in vals[:300]:
        v = np.float32(abs(v)) if v != 0 else np.float32(1.5)
        nxt = np.nextafter(v, np.float32(np.inf))
        mid = (Fraction(float(v)) + Fraction(float(nxt))) / 2
        exact = format(mid.numerator / mid.denominator, ".40g")  # a long decimal near it
        for t in (exact, "%.17g" % float(mid), "%.12g" % float(mid)):
            texts.append(t)
            expect.append(_nearest_f32(t))
        # exact midpoint written out in full: a tie
        num, den = mid.numerator, mid.denominator
        k = den.bit_length() - 1  # den = 2^k
        digits = str(num * 5 ** k)
        t = digits[:-k] + "." + digits[-k:] if k < len(digits) else "0." + digits.zfill(k)
        texts.append(t)
        expect.append(_nearest_f32(t))
    for t in ["0.000000000000000000000000000000000001234567", "1e-40", "-1e-45", "3.4028235e38",
              "3.4028236e38", "1.17549435e-38", "1.1754942e-38", "-0.0", "0", "000123.4500",
              "12345678901234567890123", "0.1234567890123456789012", "9007199254740993",
              "1e23", "1e22", "4.9406564584124654e-324", "2.5e-1", "-7.00000000000000000001"]:
        texts.append(t)
        expect.append(_nearest_f32(t))
    n = len(texts)
    while n % 4:
        texts.append("1")
        expect.append(np.float32(1))
        n += 1
    rows = [" ".join(texts[i:i + 4]) for i in range(0, n, 4)]
    p = tmp_path / "t.xyz"
    p.write_text("0 0 0\n1 1 1\n0.1\n0.01\n0.001\n" + "\n".join(rows) + "\n1.5e 2 3 4\n")
    try:
        frame_io.read_xyz(str(p))
        raise AssertionError("a line with a bare exponent marker must not parse")
    except _lib.FsrError:
        pass
    p.write_text("0 0 0\n1 1 1\n0.1\n0.01\n0.001\n" + "\n".join(rows) + "\n")
    _, pos, dens = frame_io.read_xyz(str(p))
    got = np.column_stack([pos, dens]).ravel()
    want = np.array(expect, np.float32)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, [(texts[i], got[i], want[i]) for i in bad[:5]]


def test_reconstructor_constants():
    head = dict(aabb_min=(0, 0, 0), aabb_max=(1, 1, 1), kernel_radius=0.036,
                particle_radius=0.01, particle_mass=0.004)
    f = np.float32
    z = frame_io.sim_param_for("zb05", head)
    assert z.iso_value == f(-1e-4) and z.sc_sp_grid_res_ratio == 2
    assert z.spatial_cell_size_scale == 1.0 and z.expand_extent == 4
    assert z.smoothing_radius_sq == f(f(0.036) * f(0.036))
    assert z.smoothing_radius_inv == f(1.0 / float(f(0.036)))
    a = frame_io.sim_param_for("akinci12", head)
    assert a.iso_value == f(-1e-4) and a.sc_sp_grid_res_ratio == 2
    y = frame_io.sim_param_for("yu13", head)
    assert y.iso_value == 0.0 and y.sc_sp_grid_res_ratio == 3 and y.expand_extent == 3
    assert y.spatial_cell_size_scale == f(1.3) and y.lambda_for_smoothed == f(0.5)
    assert y.min_num_neighbors == 25 and y.anisotropic_radius == f(f(0.036) * 2)


def test_driver_usage():
    if not os.path.exists(frame_io.DRIVER):
        pytest.fail("fsr_reconstruct is not built (run __graft_entry__.build())")
    r = subprocess.run([frame_io.DRIVER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
