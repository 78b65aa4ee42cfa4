"""Reference case files through the C ABI (include/mpm/mpm_case.h): host-only
reading and writing, no GPU call.

The beam block of the reference generator's input.txt (tests/golden/
beam_input.txt) is turned into griddata.txt / spdata.txt by an independent
Python restatement of the generator (tests/golden/beam_case.py). The library
must read back what the generator's formats hold and write byte-identical
files."""
import os
import sys

import numpy as np
import pytest

from mpm_mi355x import capi

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import beam_case  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="engine library not built")


@pytest.fixture()
def beam(tmp_path):
    c, b = beam_case.make(str(tmp_path))
    case = capi.Case(beam_case.BEAM_INPUT, str(tmp_path / "griddata.txt"), str(tmp_path / "spdata.txt"))
    return case, c, b, tmp_path


def test_beam_parameters_read_as_written(beam):
    case, c, b, _ = beam
    d = case.desc()
    assert d.step.scheme == capi.SCHEME_FLIP
    assert (d.rstflag, d.step.nlstep, d.ista, d.iend, d.iout) == (2, 1, 1, 2, 1)
    assert (d.step.dt, d.step.mtol) == (1.0e-3, 1.0e-12)
    assert tuple(d.grid.xmin) == (0.0, 0.0, 0.0) and tuple(d.grid.xmax) == (5.0, 0.5, 5.0)
    assert tuple(d.grid.nelem) == (10, 1, 10) and tuple(d.grid.idimc) == (2, 1, 2)
    assert tuple(d.npxye) == (4, 1, 4) and d.rho == 2500.0 and d.nmat == 1
    assert tuple(d.step.bb) == (0.0, 0.0, -9.81)
    assert case.materials() == [(capi.MODEL_NEOHOOKEAN, [1.0e7, 0.3])]


def test_beam_points_and_bcs_closed_form(beam):
    case, c, b, _ = beam
    d = case.desc()
    # Cells i = 0..7 (x <= 4) and k = 7, 8 (z in [3.5, 4.5]), 16 points each.
    assert d.n == 8 * 2 * 16
    # nodec = 12 x 2 x 12 = 288; u and w fix the x and z boundary planes
    # (288 - 10*2*10 interior = 88), v fixes k == 0 and both y planes (all).
    assert tuple(d.ibc) == (88, 288, 88)
    p = case.particles()
    volp = 0.5 * 0.5 * 0.5 / 16
    assert np.all(p.vol0 == float("%.6e" % volp)) and np.all(p.mass == float("%.6e" % (2500.0 * volp)))
    assert np.all(p.matid == 0)
    assert p.coord[:, 0].min() > 0 and p.coord[:, 0].max() < 4.0
    assert p.coord[:, 2].min() > 3.5 and p.coord[:, 2].max() < 4.5
    # the file holds 7 significant digits; the reader returns exactly those
    assert np.array_equal(p.coord, np.array([[float("%.6e" % v) for v in r] for r in b["coord"]]))
    bc = case.bcs()
    for comp in range(3):
        assert np.array_equal(bc[comp][0], np.array(b["bc"][comp]))
        assert np.all(bc[comp][1] == 0.0)


def test_writer_is_byte_identical_to_generator_format(beam):
    case, c, b, tmp = beam
    out = tmp / "out"
    out.mkdir()
    case.write(str(out / "griddata.txt"), str(out / "spdata.txt"))
    for f in ("griddata.txt", "spdata.txt"):
        assert (out / f).read_bytes() == (tmp / f).read_bytes(), f


def test_round_trip_through_written_files(beam):
    case, c, b, tmp = beam
    out = tmp / "rt"
    out.mkdir()
    case.write(str(out / "griddata.txt"), str(out / "spdata.txt"))
    again = capi.Case(beam_case.BEAM_INPUT, str(out / "griddata.txt"), str(out / "spdata.txt"))
    p, q = case.particles(), again.particles()
    assert np.array_equal(p.coord, q.coord) and np.array_equal(p.mass, q.mass)
    assert tuple(case.desc().ibc) == tuple(again.desc().ibc)


def test_scene_view_matches_case(beam):
    case, c, b, _ = beam
    s = case.scene()
    assert s.idimc == (2, 1, 2) and s.npxye == (4, 1, 4) and s.nelem == (10, 1, 10)
    assert s.parts.n == 256 and s.scheme == capi.SCHEME_FLIP
    assert s.materials == [(capi.MODEL_NEOHOOKEAN, [1.0e7, 0.3])]


def test_missing_and_malformed_files_are_errors(tmp_path):
    with pytest.raises(capi.MPMError) as e:
        capi.Case(str(tmp_path / "nope.txt"))
    assert e.value.code == capi.MPM_ERR_INVALID_ARGUMENT and "cannot open" in str(e.value)
    bad = tmp_path / "spdata.txt"
    bad.write_text("         3\n 1.0 2.0 3.0\n")
    with pytest.raises(capi.MPMError):
        capi.Case(None, None, str(bad))
    gd = tmp_path / "griddata.txt"
    gd.write_text("2 1 2\n0 0 0\n1 1 1\n1 1 1 1\n8 2 2 2\n5 3 2 3\n")  # nodec 3x2x3 inconsistent with idimc
    with pytest.raises(capi.MPMError):
        capi.Case(None, str(gd), None)


def test_transfer_scheme_names_and_bad_matid(tmp_path):
    src = open(beam_case.BEAM_INPUT).read().splitlines()
    f = tmp_path / "input.txt"
    for name, code in (("APIC", capi.SCHEME_APIC), ("TPIC", capi.SCHEME_TPIC), ("PIC", capi.SCHEME_PIC)):
        src[1] = name + "  0  2  1"
        f.write_text("\n".join(src) + "\n")
        assert capi.Case(str(f)).desc().step.scheme == code
    src[1] = "MUSL  0  2  1"   # ParseMapScheme rejects unknown names
    f.write_text("\n".join(src) + "\n")
    with pytest.raises(capi.MPMError) as e:
        capi.Case(str(f))
    assert e.value.code == capi.MPM_ERR_INVALID_ARGUMENT and "MUSL" in str(e.value)
    sp = tmp_path / "spdata.txt"
    sp.write_text("%10d\n%15.6e%15.6e%15.6e\n%10d%10d%15.6e%15.6e\n" % (1, 1.0, 0.25, 4.0, 0, 3, 1.0, 1.0))
    with pytest.raises(capi.MPMError):
        capi.Case(beam_case.BEAM_INPUT, None, str(sp))


# ------------------------------------------------------------------ restart files (solid_data_io.cpp:61-141)
def _state(n=7, seed=4):
    rng = np.random.default_rng(seed)
    return capi.Particles(n, coord=rng.random((n, 3)), vel=rng.normal(size=(n, 3)), mass=1 + rng.random(n),
                          vol0=1e-3 * (1 + rng.random(n)), vol=1e-3 * (1 + rng.random(n)),
                          stress=rng.normal(size=(n, 6)) * 1e4, def_grad=np.eye(3) + 0.1 * rng.normal(size=(n, 3, 3)),
                          matid=rng.integers(0, 3, n))


def _reference_layout(p, scheme, cmat, dinv):
    """RestartOutput's text: ios::right | ios::scientific, default precision."""
    e = lambda v: "%15.6e" % v
    rows = lambda a, k: "".join("".join(e(v) for v in r) + "\n" for r in np.asarray(a).reshape(-1, k))
    n = p.n
    t = "%d     --- Number of solid material point --- \n" % n
    t += rows(p.coord, 3)
    for i in range(n):
        t += "%d%15d%15d%s%s%s\n" % (i, p.matid[i], 0, e(p.mass[i]), e(p.vol0[i]), e(p.vol[i]))
    t += rows(p.vel, 3) + rows(np.zeros((n, 3)), 3) + rows(p.stress, 6) + rows(p.def_grad, 9)
    if scheme == capi.SCHEME_TPIC:
        t += rows(cmat, 9) + rows(np.zeros((n, 9)), 9)
    if scheme == capi.SCHEME_APIC:
        t += rows(cmat, 9) + rows(np.zeros((n, 9)), 9) + rows(dinv, 9)
    return t


@pytest.mark.parametrize("scheme", [capi.SCHEME_FLIP, capi.SCHEME_TPIC, capi.SCHEME_APIC])
def test_restart_writer_is_the_reference_layout(tmp_path, scheme):
    p = _state()
    rng = np.random.default_rng(1)
    cm, di = rng.normal(size=(p.n, 3, 3)), rng.normal(size=(p.n, 3, 3))
    f = tmp_path / "sp0_re.txt"
    capi.restart_save(str(f), p, scheme, cm, di if scheme == capi.SCHEME_APIC else None)
    assert f.read_text() == _reference_layout(p, scheme, cm, di)


@pytest.mark.parametrize("scheme", [capi.SCHEME_PIC, capi.SCHEME_APIC])
def test_restart_round_trip(tmp_path, scheme):
    p = _state()
    rng = np.random.default_rng(2)
    cm, di = rng.normal(size=(p.n, 3, 3)), rng.normal(size=(p.n, 3, 3))
    ids, surf = np.arange(p.n)[::-1].copy(), rng.integers(0, 2, p.n)
    full, short = tmp_path / "full.txt", tmp_path / "short.txt"
    capi.restart_save(str(full), p, scheme, cm, di, ids, surf, digits=17)
    capi.restart_save(str(short), p, scheme, cm, di, ids, surf, digits=0)
    q, qcm, qdi, qid, qsurf = capi.restart_load(str(full), scheme)
    for fld in ("coord", "vel", "mass", "vol0", "vol", "stress", "def_grad", "matid"):
        assert np.array_equal(getattr(q, fld), getattr(p, fld)), fld
    assert np.array_equal(q.det_def_grad, p.vol / p.vol0)  # solid_data_io.cpp:79
    assert np.array_equal(qid, ids) and np.array_equal(qsurf, surf)
    if scheme == capi.SCHEME_APIC:
        assert np.array_equal(qcm, cm) and np.array_equal(qdi, di)
    else:
        assert qcm is None and qdi is None
    r = capi.restart_load(str(short), scheme)[0]
    np.testing.assert_allclose(r.coord, p.coord, rtol=5e-7)          # 7 significant digits
    assert not np.array_equal(r.coord, p.coord)


def test_restart_errors(tmp_path):
    with pytest.raises(capi.MPMError):
        capi.restart_load(str(tmp_path / "missing.txt"))
    f = tmp_path / "bad.txt"
    f.write_text("3     --- Number of solid material point --- \n1 2 3\n")
    with pytest.raises(capi.MPMError):
        capi.restart_load(str(f))
    p = _state(3)
    capi.restart_save(str(f), p, capi.SCHEME_PIC)
    with pytest.raises(capi.MPMError):   # an APIC reader expects the matrix blocks
        capi.restart_load(str(f), capi.SCHEME_APIC)


def _spdata_lines(path):
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    return lines, n


def test_spdata_with_the_solver_readers_five_columns(beam):
    """The solver's reader takes 'id matid surf_point mass vol0'
    (module/solid/solid_data_io.cpp:22); a five-column file gives the same
    mass and vol0 as the generator's four-column one."""
    case, c, b, tmp = beam
    lines, n = _spdata_lines(tmp / "spdata.txt")
    five = lines[: 1 + n]
    for l in lines[1 + n: 1 + 2 * n]:
        t = l.split()
        five.append("%10s%10s%10d%15s%15s" % (t[0], t[1], 1, t[2], t[3]))
    (tmp / "sp5.txt").write_text("\n".join(five) + "\n")
    again = capi.Case(beam_case.BEAM_INPUT, str(tmp / "griddata.txt"), str(tmp / "sp5.txt"))
    p, q = case.particles(), again.particles()
    assert np.array_equal(p.mass, q.mass) and np.array_equal(p.vol0, q.vol0)
    assert np.array_equal(p.matid, q.matid) and np.array_equal(p.coord, q.coord)


@pytest.mark.parametrize("cols", [3, 6])
def test_spdata_with_another_column_count_is_rejected(beam, cols):
    case, c, b, tmp = beam
    lines, n = _spdata_lines(tmp / "spdata.txt")
    bad = lines[: 1 + n] + [" ".join(["1"] * cols)] * n
    (tmp / "bad.txt").write_text("\n".join(bad) + "\n")
    with pytest.raises(capi.MPMError) as ei:
        capi.Case(beam_case.BEAM_INPUT, str(tmp / "griddata.txt"), str(tmp / "bad.txt"))
    assert ei.value.code == capi.MPM_ERR_INVALID_ARGUMENT


def test_declared_counts_beyond_the_file_size_are_rejected(beam, tmp_path):
    """A corrupt count line cannot make the reader allocate for 2^31 points."""
    case, c, b, tmp = beam
    (tmp / "huge.txt").write_text("2000000000\n0.1 0.1 0.1\n")
    with pytest.raises(capi.MPMError):
        capi.Case(beam_case.BEAM_INPUT, str(tmp / "griddata.txt"), str(tmp / "huge.txt"))
    (tmp / "huge_re.txt").write_text("2000000000\n")
    import ctypes as C
    lib = capi.load()
    h = C.c_void_p()
    rc = lib.mpm_restart_load(str(tmp / "huge_re.txt").encode(), capi.SCHEME_PIC, C.byref(h))
    assert rc == capi.MPM_ERR_INVALID_ARGUMENT and not h.value
