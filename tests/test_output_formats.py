"""The reference's on-disk formats (host writers of sph2d_output.hip, no GPU).

Text layout is pinned against the reference's own shipped output files
(tests/golden/output_formats/, verbatim excerpts made by
tests/golden/make_output_fixtures.py): rows re-written from their own parsed
values must come back byte for byte, so the Fw.d edit descriptor, the
list-directed header and ZONE lines and the line ends all match. The
Ew.dE3 descriptor (VTK file) is pinned by its definition (0.d1..d9 mantissa, the
correctly rounded 9 significant digits, a signed 3-digit exponent), the binary
#!TDV101 file by the record sequence of Output_for_tecplot.f90:187-291, and the
restart record by a write/read round trip including the NaN step-back of
Distribute_Latest_Particle.f90.
"""
import math
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import sph2d
from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "output_formats")


def gold_lines(name):
    with open(os.path.join(GOLD, name + ".head"), "rb") as f:
        return f.read().split(b"\n")[:-1]


def scales():
    return sph2d.Scales(1.0, 1.0, 1000.0, 1.0)        # TGV: Ly = 1, U = 1, rho0 U^2 = 1000, Lx/U = 1


def test_final_results_rewrite_reproduces_the_shipped_file(tmp_path):
    gold = gold_lines("Final_Doamin_Results.dat")
    rows = np.array([[float(v) for v in ln.split()] for ln in gold[3:]])
    n = len(rows)
    out = tmp_path / "Final_Doamin_Results.dat"
    sph2d.write_final_results(out, rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3], rows[:, 4], rows[:, 5])
    mine = out.read_bytes().split(b"\n")[:-1]
    assert mine[:2] == gold[:2]
    assert mine[2] == b" ZONE I=%12d  F=POINT" % n
    assert gold[2] == b" ZONE I=       10000  F=POINT"          # the shipped run's particle count
    assert mine[3:] == gold[3:]


def test_surface_zone_lines_match_the_shipped_file(tmp_path):
    gold = gold_lines("Free_Surface.dat")
    out = tmp_path / "Free_Surface.dat"
    for k in range(1, 5):                                 # save_time_step 500 x dt 1e-4
        sph2d.write_surface_zone(out, 500 * k * 1e-4, [], [], [], [], scales())
    assert out.read_bytes().split(b"\n")[:-1] == gold


@pytest.mark.parametrize("name,first", [("Kinetic_energy.dat", 0), ("Norm_Pressure.dat", 8)])
def test_sampling_rows_reproduce_the_shipped_files(tmp_path, name, first):
    gold = gold_lines(name)
    for ln in gold:
        vals = [float(v) for v in ln.split()]
        out14 = np.zeros(14)
        out14[first:first + len(vals) - 1] = vals[1:]
        sph2d.write_sampling(tmp_path, vals[0], out14)
    assert (tmp_path / name).read_bytes().split(b"\n")[:-1] == gold


def test_fixed_format_edge_cases():
    f = lambda v, w=20, d=8: sph2d.format_fortran(v, "F", w, d)
    assert f(0.5) == " " * 10 + "0.50000000"
    assert f(-0.5) == " " * 9 + "-0.50000000"
    assert f(-1e-12) == " " * 9 + "-0.00000000"            # Fortran keeps the sign of a negative value
    assert f(float("nan")) == " " * 17 + "NaN"
    assert f(float("-inf")) == " " * 11 + "-Infinity"     # log10(0) of an exact analytic match
    assert f(1e15, 12, 5) == "*" * 12                      # does not fit: asterisks
    assert sph2d.time_name(0.05, 12, 5) == "0.05000" and sph2d.time_name(0.05, 8, 4) == "0.0500"


@pytest.mark.parametrize("v", [1.0, 0.5, -0.123456789, 123456.789, 9.9999999995e5, 1e-300, -2.5e300, 0.0,
                               1.0 / 3.0, 5e-324])
def test_e_format_definition(v):
    s = sph2d.format_fortran(v, "E", 20, 9)
    assert len(s) == 20
    body = s.strip()
    neg = body.startswith("-")
    mant, ex = body.lstrip("-").split("E")
    assert mant.startswith("0.") and len(mant) == 11 and ex[0] in "+-" and len(ex) == 4
    if v == 0.0:
        assert mant == "0.000000000" and ex == "+000" and not neg
        return
    assert neg == (v < 0)
    digits, e = mant[2:], int(ex)
    assert digits[0] != "0"                                # normalised to 0.1 <= m < 1
    # the same 9 significant digits Python's correctly rounded formatting gives
    ref = "%.8e" % abs(v)
    assert digits == ref[0] + ref[2:10] and e == int(ref.split("e")[1]) + 1


def test_tecplot_zone_columns(tmp_path):
    x, y = np.array([0.25, 0.5, 0.75]), np.array([0.1, 0.2, 0.3])
    u, v = np.array([0.3, -0.4, 0.0]), np.array([0.4, 0.3, -0.2])
    p = np.array([100.0, -50.0, 0.0])
    vort = np.array([1.5, -2.0, 0.25])
    ana = np.array([[0.31, 0.41, 101.0, 0.5], [-0.39, 0.31, -50.0, 0.49], [0.0, -0.2, 1e-3, 0.2]])
    out = tmp_path / "video_result.dat"
    sc = sph2d.Scales(2.0, 0.5, 250.0, 4.0)
    sph2d.write_tecplot_zone(out, 0.05, x, y, u, v, p, vort, ana, sc)
    lines = out.read_text().split("\n")
    assert lines[0] == " TITLE='DISTRIBUTION'"
    assert lines[1] == " VARIABLES= 'X' 'Y' 'VX' 'VY' 'P' 'U' 'Vor' 'EVX' 'EVY' 'EP' 'EU' "
    assert lines[2] == ' ZONE T=" 0.05000 ", I=           3  F=POINT'
    for i in range(3):
        row = lines[3 + i]
        assert len(row) == 11 * 20
        cols = [row[20 * k:20 * k + 20] for k in range(11)]
        vm = math.sqrt(u[i] ** 2 + v[i] ** 2)
        with np.errstate(divide="ignore"):
            expect = [x[i] / 2, y[i] / 2, u[i] / 0.5, v[i] / 0.5, p[i] / 250, vm / 0.5, vort[i] * 4.0,
                      np.log10(abs((u[i] - ana[i, 0]) / 0.5)), np.log10(abs((v[i] - ana[i, 1]) / 0.5)),
                      np.log10(abs((p[i] - ana[i, 2]) / 250)), np.log10(abs((vm - ana[i, 3]) / 0.5))]
        assert cols == [sph2d.format_fortran(e, "F", 20, 8) for e in expect]
    assert lines[3 + 2][7 * 20:8 * 20].strip() == "-Infinity"   # u equals its analytic value exactly


def _read_i32(b, o):
    return struct.unpack_from("<i", b, o)[0], o + 4


def _read_str(b, o):
    s = []
    while True:
        v, o = _read_i32(b, o)
        if v == 0:
            return "".join(s), o
        s.append(chr(v))


def test_tecplot_binary_record_sequence(tmp_path):
    n = 5
    rng = np.random.default_rng(3)
    x, y, u, v, p = (rng.standard_normal(n) for _ in range(5))
    sc = sph2d.Scales(2.0, 0.5, 250.0, 4.0)
    out = tmp_path / "0.0500.dat"
    sph2d.write_tecplot_binary(out, 0.05, x, y, u, v, p, sc)
    b = out.read_bytes()
    assert b[:8] == b"#!TDV101"
    o = 8
    one, o = _read_i32(b, o)
    assert one == 1
    title, o = _read_str(b, o)
    assert title == "TITLE='0.0500'"
    nv, o = _read_i32(b, o)
    names = []
    for _ in range(nv):
        s, o = _read_str(b, o)
        names.append(s)
    assert names == ["X", "Y", "VX", "VY", "P", "U"]
    assert struct.unpack_from("<f", b, o)[0] == 299.0
    zone, o = _read_str(b, o + 4)
    assert zone == "ZONE 0.0500"
    ints = struct.unpack_from("<9i", b, o)
    assert ints == (-1, 0, 1, 0, 0, n, 1, 1, 0)            # colour, ordered, point, no location, no faces, IJK, no aux
    o += 36
    assert struct.unpack_from("<ff", b, o) == (357.0, 299.0)
    o += 8
    assert struct.unpack_from("<6i", b, o) == (2,) * 6 and struct.unpack_from("<2i", b, o + 24) == (0, -1)
    o += 32
    data = np.frombuffer(b[o:], dtype="<f8").reshape(n, 6)
    assert len(b) == o + 48 * n
    expect = np.stack([x / 2, y / 2, u / 0.5, v / 0.5, p / 250, np.sqrt(u ** 2 + v ** 2) / 0.5], axis=1)
    assert np.array_equal(data, expect)


def test_paraview_file_parses_as_vtk_xml(tmp_path):
    n = 7
    rng = np.random.default_rng(4)
    x, y, u, v, p = (rng.standard_normal(n) for _ in range(5))
    sc = sph2d.Scales(2.0, 0.5, 250.0, 4.0)
    out = tmp_path / "flow.vtu.0.05000"
    sph2d.write_paraview(out, x, y, u, v, p, sc)
    root = ET.parse(out).getroot()
    assert root.tag == "VTKFile" and root.get("type") == "UnstructuredGrid"
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == n and int(piece.get("NumberOfCells")) == 0
    arr = lambda path: np.array(piece.find(path).text.split(), dtype=float)
    pos = arr("Points/DataArray").reshape(n, 3)
    vel = arr("PointData/DataArray[@Name='Velocity']").reshape(n, 3)
    pr = arr("PointData/DataArray[@Name='Pressure']")
    np.testing.assert_allclose(pos[:, :2], np.stack([x / 2, y / 2], 1), rtol=6e-9, atol=0)
    np.testing.assert_allclose(vel[:, :2], np.stack([u / 0.5, v / 0.5], 1), rtol=6e-9, atol=0)
    np.testing.assert_allclose(pr, p / 250, rtol=6e-9, atol=0)
    assert np.all(pos[:, 2] == 0) and np.all(vel[:, 2] == 0)
    lines = out.read_text().split("\n")
    assert lines[3] == ' <Piece NumberOfPoints="           7 " NumberOfCells="           0 " >'


def test_restart_zone_round_trip_and_nan_step_back(tmp_path):
    n = 6
    rng = np.random.default_rng(5)
    out = tmp_path / "Latest_Time_Results.dat"
    zones = []
    for k, step in enumerate((5000, 10000, 15000)):
        st = {f: rng.standard_normal(n) for f in ("x", "y", "u", "v", "p", "rho", "mass", "energy", "c")}
        st["type"] = np.full(n, 2, dtype=np.int32)
        if k == 2:
            st["p"][3] = np.nan
        sph2d.write_latest_results_zone(out, step, st["x"], st["y"], st["u"], st["v"], st["p"], st["rho"],
                                        st["mass"], st["energy"], st["c"], st["type"])
        zones.append(st)
    lines = out.read_text().split("\n")
    assert lines[1] == " VARIABLES= 'X' 'Y' 'VX' 'VY' 'P' 'Rho' 'Mass' 'Ene' 'C' 'Type' 'Wmak_L' 'Step' "
    assert lines[2] == " ZONE I=           6  F=POINT" and len(lines[3]) == 9 * 20 + 3 * 12
    got = sph2d.read_latest_results(out)                  # the last zone holds a NaN: zone 2 is taken
    assert got["step"] == 10000 and len(got["x"]) == n
    for f in ("x", "y", "u", "v", "p", "rho", "mass", "energy", "c"):
        assert np.array_equal(got[f], np.round(zones[1][f], 8)), f
    assert np.array_equal(got["type"], zones[1]["type"])
    first = sph2d.read_latest_results(out, zone=1)
    assert first["step"] == 5000 and np.array_equal(first["rho"], np.round(zones[0]["rho"], 8))
    with pytest.raises(sph2d.SphError) as e:
        sph2d.read_latest_results(out, zone=1, capacity=3)
    assert e.value.code == 2


# --- post-processing grid files ----------------------------------------------------------
def test_initial_grid_files_equal_the_shipped_ones(tmp_path):
    """Distribute_Grid_For_Postproceeding.f90 at the shipped case (100 x 100 elements,
    element_dx = 1/100): both files byte for byte (SHA-256 of the reference's files),
    which pins the node order, the F20.10 node rows, the quadrilateral numbering
    (8I8) and the list-directed ZONE line broken at column 80."""
    import hashlib
    import json
    gold = json.load(open(os.path.join(GOLD, "initial_grid.json")))
    g, nd = tmp_path / "Initial_Post_Proceeding_Grid.dat", tmp_path / "Initial_Post_Proceeding_Node.dat"
    sph2d.write_initial_grid(g, nd, 100, 100, 1.0 / 100)
    for path, name in ((g, "Initial_Post_Proceeding_Grid.dat"), (nd, "Initial_Post_Proceeding_Node.dat")):
        data = path.read_bytes()
        assert [ln.decode() for ln in data.split(b"\n")[:8]] == gold[name]["head"], name
        assert data.count(b"\n") == gold[name]["lines"], name
        assert hashlib.sha256(data).hexdigest() == gold[name]["sha256"], name


def test_grid_zone_and_binary(tmp_path):
    nex, ney, edx = 3, 2, 0.25
    nn = (nex + 1) * (ney + 1)
    rng = np.random.default_rng(6)
    u, v, p, w = (rng.standard_normal(nn) for _ in range(4))
    fai = np.linspace(-1, 1, nn)
    ana = rng.standard_normal((nn, 4))
    sc = sph2d.Scales(1.0, 1.0, 1000.0, 1.0)
    out = tmp_path / "Post_Proceeding_Grid.dat"
    sph2d.write_grid_zone(out, 0.05, nex, ney, edx, u, v, p, w, fai, ana, sc)
    lines = out.read_text().split("\n")
    assert lines[0] == " TITLE='mesh'"
    assert lines[1] == " VARIABLES= 'X' 'Y' 'VX' 'VY' 'P' 'U' 'Vor' 'EVX' 'EVY' 'EP' 'EU' 'Fai'"
    assert lines[2] == ' ZONE T=" 0.05000 ",N=          12 ,E=           6 '
    assert lines[3] == " ,DATAPACKING=POINT,ZONETYPE=FEQUADRILATERAL "
    row = lines[4 + 5]                                       # node 6: (1*edx, 1*edx)
    cols = [float(row[20 * k:20 * k + 20]) for k in range(12)]
    assert cols[:2] == [0.25, 0.25] and abs(cols[11] - fai[5]) < 5e-9
    assert lines[4 + nn] == "%8d%8d%8d%8d" % (1, 2, 6, 5)    # first quadrilateral
    assert lines[4 + nn + 3] == "%8d%8d%8d%8d" % (5, 6, 10, 9)     # first of the second element row
    b = tmp_path / "g.dat"
    sph2d.write_grid_binary(b, 0.05, nex, ney, edx, u, v, p, fai, sc)
    data = b.read_bytes()
    o = 8 + 4
    title, o = _read_str(data, o)
    nv, o = _read_i32(data, o)
    names = []
    for _ in range(nv):
        s, o = _read_str(data, o)
        names.append(s)
    assert title == "TITLE='0.0500'" and names == ["X", "Y", "VX", "VY", "P", "U", "Fai"]
    zone, o = _read_str(data, o + 4)
    assert struct.unpack_from("<9i", data, o) == (-1, 0, 1, 0, 0, nex + 1, ney + 1, 1, 0)
    vals = np.frombuffer(data[-56 * nn:], dtype="<f8").reshape(nn, 7)
    assert np.array_equal(vals[:, 6], fai) and np.array_equal(vals[:, 2], u)


# --- Sampling_subroutine.f90 files -----------------------------------------------------------
def test_wave_front_file_is_reproduced_byte_for_byte(tmp_path):
    """Wave_front_results.dat (write(127,*), ifort list-directed REAL(8)): all 500
    shipped records, E form for values below 0.1 and F form with 15 significant
    digits above, rewritten from their own parsed values."""
    lines = gold_lines("Wave_front_results.dat")
    assert len(lines) == 500
    out = tmp_path / "Wave_front_results.dat"
    for ln in lines:
        sph2d.write_list_directed(out, [float(t) for t in ln.split()])
    assert out.read_bytes().split(b"\n")[:-1] == lines


@pytest.mark.parametrize("v,text", [
    (0.01, "  1.000000000000000E-002"),
    (0.995331222756290, "  0.995331222756290     "),
    (4.99, "   4.99000000000000     "),
    (-0.5, " -0.500000000000000     "),
    (0.09999999999999999, "  0.100000000000000     "),     # 0.1 after rounding to 15 digits: F form
    (0.0999999999999, "  9.999999999990000E-002"),
    (123456.789, "   123456.789000000     "),
    (999999999999999.0, "   999999999999999.     "),
    (1e15, "  1.000000000000000E+015"),
    (-2.5e-300, " -2.500000000000000E-300"),
])
def test_list_directed_real8_forms(v, text):
    """The G-editing choice on the value rounded to 15 significant digits, the F
    form's 5 trailing blanks and the 1P E form's 16 digits and 3-digit exponent
    (the two forms the shipped file shows, and their boundaries)."""
    item = sph2d.format_list_directed(v)
    assert len(item) == 23
    assert " " + item == text


def test_list_directed_records_wrap_at_column_80(tmp_path):
    """Three items fit one 80-column record (1 + 3 x 23 + 2 blanks = 72); a fourth
    starts a new record with its own leading blank."""
    out = tmp_path / "ld.dat"
    sph2d.write_list_directed(out, [0.5, 0.25, 0.125, 2.0])
    recs = out.read_text().split("\n")[:-1]
    assert len(recs) == 2 and len(recs[0]) == 72 and recs[1] == "   2.00000000000000     "


def test_pressure_probe_lines_use_format_310(tmp_path):
    """The pressure probe files take format 310 (20F20.8) as Kinetic_energy.dat does:
    a shipped row rewritten through the probe writer comes back unchanged."""
    row = gold_lines("Kinetic_energy.dat")[3]
    out = tmp_path / "Probe_1_Kalman_Pressure.dat"
    sph2d.write_f20_8(out, [float(t) for t in row.split()])
    assert out.read_bytes() == row + b"\n"


def test_kalman_filter_first_sample_and_half_gain():
    """Sampling_subroutine.f90:300-322: the first sampling passes the measurement and
    sets the state to (x0 = 1, P = measurement); with P equal to R (1.0e-5 as REAL(4))
    the next gain is exactly 1/2, so the output is the midpoint of x0 and the new
    measurement and P becomes R/2 + Q (Q = 1.0e-6 as REAL(4))."""
    R, Q = float(np.float32(1.0e-5)), float(np.float32(1.0e-6))
    lx, lp = np.zeros(2), np.zeros(2)
    first = sph2d.kalman_filter(1, [R, 0.3], lx, lp)
    assert list(first) == [R, 0.3] and list(lx) == [1.0, 1.0] and list(lp) == [R, 0.3]
    second = sph2d.kalman_filter(2, [3.0, 0.3], lx, lp)
    assert second[0] == 2.0 and lx[0] == 2.0 and lp[0] == R - 0.5 * R + Q
    assert 0.3 < second[1] < 1.0                       # between the state and the measurement
    for k in range(3, 400):                            # a constant measurement is approached
        last = sph2d.kalman_filter(k, [3.0, 0.3], lx, lp)
    assert abs(last[1] - 0.3) < 1e-6 and abs(last[0] - 3.0) < 1e-6


def test_sampling_line_points_and_position_file(tmp_path):
    """Initialize_SPH_Calculation.f90:362-436: 41 evenly spaced points per line from
    the first to the last point (both hit exactly), written under a two-line header
    as 6f12.6 rows."""
    x, y = sph2d.sampling_line_points([0.0, 0.25], [1.0, 0.25], 41)
    assert x[0] == 0.0 and x[-1] == 1.0 and np.all(y == 0.25)
    assert np.max(np.abs(np.diff(x) - 0.025)) < 1e-15
    sph2d.write_sampling_line_positions(tmp_path, 2, x, y)
    lines = (tmp_path / "2th_SamplingLine_Position.dat").read_text().split("\n")
    assert lines[0] == " TITLE='DISTRIBUTION'" and lines[1] == " VARIABLES= 'X' 'Y'"
    assert lines[2] == "    0.000000    0.250000" and lines[3] == "    0.025000    0.250000"
    assert lines[42] == "    1.000000    0.250000" and lines[43] == ""
