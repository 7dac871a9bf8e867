"""The reference's file formats on the host (CPU only, no device call).

Pinned against the reference's own code: oracle/_ref/ref_tools is the
reference's Mesher and vendored JSON library built from their sources; its
outputs on the committed inputs are the fixtures under tests/golden/io
(tests/golden/make_io_golden.py).  Where the reference tree and ref_tools
are present, the same comparisons also run live on the reference's own
mesh directories and JSON files.  Closed forms pin the Mesher further
(uniform node lists) and buildAP is checked against the oracle.
"""
import glob
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from falm_amd import falmio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
IO = os.path.join(GOLD, "io")
CASE = os.path.join(GOLD, "casedir")
REF_TOOLS = os.path.join(ROOT, "oracle", "_ref", "ref_tools")
REFERENCE = "/root/reference"
LIBDIR = os.path.join(ROOT, "falm_amd", "lib")


@pytest.fixture(scope="module")
def io_tool(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libfalm_amd.so")):
        pytest.fail("libfalm_amd.so is not built (python -m falm_amd.build)")
    exe = str(tmp_path_factory.mktemp("io_tool") / "io_tool")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "native", "io_tool.cpp"), "-L", LIBDIR, "-lfalm_amd",
                    "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read(path):
    with open(path) as f:
        return f.read()


def live_reference():
    return os.path.exists(REF_TOOLS) and os.path.isdir(REFERENCE)


# --------------------------------------------------------------------- JSON
@pytest.mark.parametrize("src,expected", [
    (os.path.join(IO, "json_doc.json"), os.path.join(IO, "json_doc.json.sorted")),
    (os.path.join(CASE, "setup.json"), os.path.join(IO, "casedir_setup.json.sorted")),
    (os.path.join(CASE, "turbineProperties.json"), os.path.join(IO, "casedir_turbineProperties.json.sorted")),
])
def test_json_dump_matches_reference_library(io_tool, src, expected):
    assert run(io_tool, "json", src) == read(expected)


def test_ordered_json_keeps_insertion_order(io_tool):
    assert run(io_tool, "ojson", os.path.join(IO, "json_doc.json")) == read(os.path.join(IO, "json_doc.json.ordered"))


@pytest.mark.parametrize("name", ["numbers", "numbers_random"])
def test_json_numbers_match_reference_library(io_tool, name):
    nums = read(os.path.join(IO, name + ".txt")).split()
    assert run(io_tool, "num", *nums) == read(os.path.join(IO, name + ".ref"))


def test_json_cached_powers_of_ten():
    """The Grisu2 table in json.h: f * 2^e is 10^k rounded to 64 bits."""
    import re
    from fractions import Fraction
    text = read(os.path.join(ROOT, "include", "falm", "json.h"))
    rows = re.findall(r"\{0x([0-9A-F]{16})ull, (-?\d+), (-?\d+)\}", text)
    assert [int(k) for _, _, k in rows] == list(range(-300, 325, 8))
    for f, e, k in rows:
        f, e, k = int(f, 16), int(e), int(k)
        assert 2 ** 63 <= f < 2 ** 64
        exact = Fraction(10) ** k / Fraction(2) ** e
        assert abs(exact - f) <= Fraction(1, 2)


def test_json_numbers_read_back_exactly(io_tool):
    rng = np.random.default_rng(7)
    vals = np.concatenate([rng.random(400) * 10.0 ** rng.integers(-30, 30, 400),
                           -rng.random(100) * 10.0 ** rng.integers(-300, 300, 100)])
    out = run(io_tool, "num", *[repr(float(v)) for v in vals]).split()
    assert [float(s) for s in out] == [float(v) for v in vals]


def test_json_reference_files_live(io_tool):
    if not live_reference():
        pytest.skip("reference tree or oracle/_ref/ref_tools absent")
    files = sorted(glob.glob(os.path.join(REFERENCE, "app", "**", "*.json"), recursive=True) +
                   glob.glob(os.path.join(REFERENCE, "test", "*.json")))
    assert files
    for f in files:
        for mode in ("json", "ojson"):
            ref = subprocess.run([REF_TOOLS, mode, f], capture_output=True, text=True)
            if ref.returncode != 0:
                continue  # not a JSON document the reference library accepts
            assert run(io_tool, mode, f) == ref.stdout, (mode, f)


# ------------------------------------------------------------------- Mesher
MESH_CASES = [("innerNode", os.path.join(CASE, "mesh"), 2, "casedir_innerNode_gc2.cv"),
              ("innerNode", os.path.join(CASE, "mesh"), 3, "casedir_innerNode_gc3.cv"),
              ("cell", os.path.join(CASE, "mesh"), 2, "casedir_cell_gc2.cv"),
              ("cell", os.path.join(IO, "mesh2"), 2, "mesh2_cell_gc2.cv")]


@pytest.mark.parametrize("center,mesh,gc,expected", MESH_CASES)
def test_mesher_matches_reference_mesher(tmp_path, io_tool, center, mesh, gc, expected):
    out = str(tmp_path / "cv.txt")
    falmio.build_mesh(center, mesh, out, gc)  # C ABI
    assert read(out) == read(os.path.join(IO, expected))
    out2 = str(tmp_path / "cv2.txt")
    run(io_tool, "mesh", center, mesh, out2, gc)  # C++ Mesher::build_mesh
    assert read(out2) == read(os.path.join(IO, expected))


def test_mesher_reference_mesh_dirs_live(tmp_path):
    if not live_reference():
        pytest.skip("reference tree or oracle/_ref/ref_tools absent")
    dirs = sorted(os.path.dirname(p) for p in glob.glob(os.path.join(REFERENCE, "app", "*", "mesh", "**", "x.txt"),
                                                        recursive=True))
    dirs = [d for d in dirs if os.path.exists(os.path.join(d, "mesh.json"))]
    assert dirs
    from falm_amd import runtime
    compared = 0
    for d in dirs:
        for center in ("innerNode", "cell"):
            nmin = min(int(read(os.path.join(d, ax + ".txt")).split()[0]) for ax in "xyz")
            a, b = str(tmp_path / "a.cv"), str(tmp_path / "b.cv")
            if center == "innerNode" and nmin < 4:
                # the reference reads outside its arrays here (2-D node lists)
                with pytest.raises(runtime.FalmError):
                    falmio.build_mesh(center, d, a, 2)
                continue
            subprocess.run([REF_TOOLS, "mesh", center, d, b, "2"], check=True)
            falmio.build_mesh(center, d, a, 2)
            assert read(a) == read(b), (center, d)
            compared += 1
    assert compared >= 4


def _uniform_mesh(tmp_path, n, h, x0):
    d = tmp_path / "uniform"
    d.mkdir()
    (d / "mesh.json").write_text('{"coordinatePoint": "node"}')
    for ax in "xyz":
        (d / f"{ax}.txt").write_text(f"{n}\n" + "".join(f"{x0 + h * i!r}\n" for i in range(n)))
    return str(d)


def test_mesher_closed_forms(tmp_path):
    """Uniform nodes: "cell" gives the face midpoints and the node spacing for
    every padded cell; "innerNode" gives the nodes themselves and the node
    spacing, except that the upper z ghosts repeat the last width (the
    reference's z loop, src/mesher/mesher.hpp:84), which for a uniform list
    is the same number."""
    n, h, x0 = 9, 0.25, -1.0
    d = _uniform_mesh(tmp_path, n, h, x0)
    for center, gc in (("cell", 2), ("cell", 3), ("innerNode", 2), ("innerNode", 3)):
        out = str(tmp_path / f"{center}{gc}.cv")
        falmio.build_mesh(center, d, out, gc)
        cv = falmio.read_cv_file(out)
        assert cv["gc"] == gc
        if center == "cell":
            assert cv["shape"] == (n - 1 + 2 * gc,) * 3
            first = x0 + 0.5 * h - gc * h
        else:
            assert cv["shape"] == (n + 2 * (gc - 1),) * 3
            first = x0 - (gc - 1) * h
        for ax in "xyz":
            L = len(cv[ax])
            np.testing.assert_allclose(cv["h" + ax], h, rtol=1e-14)
            np.testing.assert_allclose(cv[ax], first + h * np.arange(L), rtol=0, atol=1e-14)


def test_mesher_rejects_bad_input(tmp_path):
    from falm_amd import runtime
    d = _uniform_mesh(tmp_path, 9, 0.25, 0.0)
    with pytest.raises(runtime.FalmError):
        falmio.build_mesh("faces", d, str(tmp_path / "x.cv"), 2)
    with pytest.raises(runtime.FalmError):
        falmio.build_mesh("cell", str(tmp_path / "missing"), str(tmp_path / "x.cv"), 2)


# ------------------------------------------------------------- FalmIO files
def test_control_volume_file_round_trip(tmp_path, io_tool):
    src = os.path.join(IO, "casedir_innerNode_gc2.cv")
    out = str(tmp_path / "rt.cv")
    run(io_tool, "cv", src, out)
    assert read(out) == read(src)
    cv = falmio.read_cv_file(src)
    assert cv["shape"] == (44, 24, 24) and cv["gc"] == 2
    nodes = [float(v) for v in read(os.path.join(CASE, "mesh", "x.txt")).split()[1:]]
    np.testing.assert_array_equal(cv["x"][1:-1], nodes)  # innerNode: the nodes are the centres


def test_index_file(tmp_path, io_tool):
    out = str(tmp_path / "run.json")
    run(io_tool, "index", out, 40, 20, 20, 2, "4:0.04:0", "8:0.08:1")
    idx = json.loads(read(out))
    assert idx == {"global": [40, 20, 20], "guidePoint": 2,
                   "ranks": [{"rank": 0, "voxel": [40, 20, 20], "offset": [0, 0, 0]}],
                   "outputSteps": [{"time": 0.04, "step": 4}, {"time": 0.08, "step": 8, "timeAvg": 1}],
                   "dataType": "float64", "variables": ["u", "v", "w", "p"]}
    out0 = str(tmp_path / "empty.json")
    run(io_tool, "index", out0, 40, 20, 20, 2)
    assert "outputSteps" not in json.loads(read(out0))  # as before the first snapshot
    if live_reference():  # the text is what the reference's JSON library writes
        for f in (out, out0):
            assert read(f) == subprocess.run([REF_TOOLS, "json", f], capture_output=True, text=True,
                                             check=True).stdout


def test_vector_file_layout(tmp_path, io_tool):
    out = str(tmp_path / "run_000000_0000000008")
    run(io_tool, "vec", out, 7, 5, 4, 4, 2, 8, 0.08)
    hd, data = falmio.read_vector_file(out)
    assert hd == {"imax": 7, "jmax": 5, "kmax": 4, "n": 4, "gc": 2, "step": 8, "time": 0.08, "dsize": 8}
    # the value written at (i, j, k, n) is its flat index i + 7 (j + 5 (k + 4 n))
    n, k, j, i = np.meshgrid(range(4), range(4), range(5), range(7), indexing="ij")
    np.testing.assert_array_equal(data, i + 7 * (j + 5 * (k + 4 * n)))
    assert os.path.getsize(out) == 8 * 8 + 7 * 5 * 4 * 4 * 8


# ------------------------------------------------------------- setup files
def test_setup_parse(io_tool):
    got = json.loads(run(io_tool, "setup", os.path.join(CASE, "setup.json")))
    assert got["status"] == 0
    py = falmio.parse_settings(os.path.join(CASE, "setup.json"))
    # (0.29 - 0) / 0.01 = 28.999999999999996 truncates to 28, as in the reference
    assert got["maxIt"] == py["maxIt"] == 28
    assert (got["timeAvgStartIt"], got["timeAvgEndIt"]) == (py["timeAvgStartIt"], py["timeAvgEndIt"]) == (5, 28)
    assert (got["outputStartIt"], got["outputEndIt"], got["outputIntervalIt"]) == (0, 28, 10)
    assert (py["outputStartIt"], py["outputEndIt"], py["outputIntervalIt"]) == (0, 28, 10)
    assert got["outputPrefix"] == py["outputPrefix"] == "run"
    assert got["workdir"] == py["workdir"] == CASE and got["setupFile"] == "setup.json"
    assert (got["dt"], got["Re"], got["Cs"]) == (0.01, 1e4, 0.1)
    assert (got["advection"], got["sgs"], got["solver"], got["pc"]) == ("Upwind3", "Smagorinsky", "PBiCGStab", "SOR")
    assert (got["maxit"], got["tol"], got["pc_maxit"], got["pc_relax"]) == (500, 1e-12, 2, 1.5)
    assert (got["cvFile"], got["cvCenter"], got["convert"], got["convertSrc"]) == \
        ("controlVolume.txt", "innerNode", True, "mesh")


def test_setup_errors_are_reported(tmp_path, io_tool):
    bad = tmp_path / "setup.json"
    bad.write_text('{"runtime": {"time": {"start": 0}}}')
    assert json.loads(run(io_tool, "setup", str(bad)))["status"] != 0
    assert json.loads(run(io_tool, "setup", str(tmp_path / "none.json")))["status"] != 0


def test_turbine_json_and_ap_file(tmp_path, io_tool):
    from oracle import pyoracle as O
    case = tmp_path / "case"
    shutil.copytree(CASE, case)
    setup = json.loads(read(str(case / "setup.json")))
    tp = tmp_path / "turbine.json"
    tp.write_text(json.dumps(setup["turbine"]))
    times = [-1.0, 0.0, 0.025, 0.05, 0.1, 0.2, 0.3]
    got = json.loads(run(io_tool, "turbine", str(case), str(tp), *times))
    assert got["status"] == 0
    assert (got["radius"], got["hub_radius"], got["n_blade"], got["n_ap_per_blade"], got["eps"]) == \
        (0.4, 0.05, 3, 8, 0.2)
    assert (got["n_ap"], got["n_attack"]) == (48, 15)
    t0, t1 = got["turbines"]
    assert t0["hub"] == t1["hub"] == [-0.05, 0.0, 1.0]
    assert t0["base"] == [0.0, -0.45, 0.0] and t0["tip_rate"] == 4.0
    # pitch motion {amplitude 3 deg, period 0.5, phase 10 deg} (src/alm/turbineHandler.h:171-176)
    assert t0["angle_type"] == 2
    assert t0["motion"] == pytest.approx([math.radians(3.0), 2 * math.pi / 0.5, math.radians(10.0)], rel=1e-15)
    # fixed roll 2 deg: amplitude 2 deg at phase pi/2 (src/alm/turbineHandler.h:164-170)
    assert t1["angle_type"] == 1 and t1["tip_rate"] == -3.0
    assert t1["motion"] == pytest.approx([math.radians(2.0), 0.0, 0.5 * math.pi], rel=1e-15)
    # base x of turbine 1 from base_x.txt: (0, 0) (0.05, 0.02) (0.2, -0.03)
    want = {-1.0: (0.0, 0.0), 0.0: (0.0, 0.4), 0.025: (0.01, 0.4), 0.05: (0.02, -1 / 3.0),
            0.1: (0.02 - 0.05 / 3.0, -1 / 3.0), 0.2: (-0.03, 0.0), 0.3: (-0.03, 0.0)}
    for t, b, v in t1["track"]:
        assert b[0] == pytest.approx(want[t][0], abs=1e-15) and v[0] == pytest.approx(want[t][1], rel=1e-14)
        assert b[1:] == [0.45, 0.0] and v[1:] == [0.0, 0.0]
    for t, b, v in t0["track"]:
        assert b == [0.0, -0.45, 0.0] and v == [0.0, 0.0, 0.0]
    # the Python mirror of the same parse
    alm, tracks = falmio.alm_from_turbine_json(str(case), setup["turbine"])
    assert set(tracks) == {(1, 0)}
    for t in times:
        assert tracks[(1, 0)].eval(t) == pytest.approx(tuple(want[t]), abs=1e-15)
    # ap.json: the blade table interpolated by the oracle's buildAP
    ap = json.loads(read(str(case / "ap.json")))
    prop = json.loads(read(str(case / "turbineProperties.json")))
    bp = prop["bladeProperties"]
    af = bp["airfoils"]
    rr, ch, tw = (np.array([a[k] for a in af]) for k in ("r/R", "chord/R", "twist[deg]"))
    cl, cd = np.array([a["Cl"] for a in af]), np.array([a["Cd"] for a in af])
    r, c, w, pcl, pcd = O.build_ap(rr, ch, tw, cl, cd, 2, 3, 8, 0.4, 0.05)
    nat = len(bp["attacks"])
    assert [a["id"] for a in ap["aps"]] == list(range(48)) and ap["attacks"] == bp["attacks"]
    assert list(ap["aps"][0].keys()) == ["id", "r", "chord", "twist[deg]", "Cl", "Cd"]
    tol = dict(rtol=1e-15, atol=1e-16)
    np.testing.assert_allclose([a["r"] for a in ap["aps"]], r, **tol)
    np.testing.assert_allclose([a["chord"] for a in ap["aps"]], c, **tol)
    np.testing.assert_allclose([a["twist[deg]"] for a in ap["aps"]], w, **tol)
    np.testing.assert_allclose(np.array([a["Cl"] for a in ap["aps"]]).T, np.reshape(pcl, (nat, 48)), **tol)
    np.testing.assert_allclose(np.array([a["Cd"] for a in ap["aps"]]).T, np.reshape(pcd, (nat, 48)), **tol)
    np.testing.assert_allclose(alm.ap_r, r, **tol)
    np.testing.assert_allclose(alm.cl, pcl, **tol)
    if live_reference():  # the file is laid out as the reference's ordered_json writes it
        text = read(str(case / "ap.json"))
        assert text == subprocess.run([REF_TOOLS, "ojson", str(case / "ap.json")], capture_output=True, text=True,
                                      check=True).stdout


def test_turbine_json_errors(tmp_path, io_tool):
    case = tmp_path / "case"
    shutil.copytree(CASE, case)
    tp = tmp_path / "turbine.json"
    tp.write_text(json.dumps({"turbineProperties": "missing.json", "bladePointNumber": 8, "projectionWidth": 0.2,
                              "turbineArray": [], "apFile": "ap.json"}))
    assert json.loads(run(io_tool, "turbine", str(case), str(tp)))["status"] != 0


def test_case_directory_on_the_oracle(tmp_path):
    """case_from_setup: the control volumes of the converted mesh reach the
    step driver's configuration (here the oracle's): cell centres, widths
    and metrics follow the control-volume file; the rotors load the flow."""
    from falm_amd import _abi as abi
    from oracle import pyoracle as O
    case = tmp_path / "case"
    shutil.copytree(CASE, case)
    setup = falmio.case_from_setup(str(case / "setup.json"))
    assert read(str(case / "controlVolume.txt")) == read(os.path.join(IO, "casedir_innerNode_gc2.cv"))
    assert tuple(setup.struct.n) == (40, 20, 20) and setup.settings["maxIt"] == 28
    ref = O.Case(setup)
    cv = setup.cv
    nx, ny, nz = cv["shape"]
    xyz = ref.get(abi.FIELD_XYZ).reshape(3, nz, ny, nx)
    kx = ref.get(abi.FIELD_KX).reshape(3, nz, ny, nx)
    ja = ref.get(abi.FIELD_JA).reshape(nz, ny, nx)
    np.testing.assert_array_equal(xyz[0, 0, 0, :], cv["x"])
    np.testing.assert_array_equal(xyz[2, :, 0, 0], cv["z"])
    np.testing.assert_allclose(kx[1, 0, :, 0], 1.0 / cv["hy"], rtol=1e-15)
    np.testing.assert_allclose(ja, cv["hz"][:, None, None] * cv["hy"][None, :, None] * cv["hx"][None, None, :],
                               rtol=1e-15)
    for it in (1, 2):
        for tid, (b, v) in falmio.base_motion(setup, it * setup.struct.dt).items():
            ref.set_base(tid, b, v)
        info = ref.step(1)
    assert info.div < 1e-8
    _, _, tq, th = ref.alm_get()
    # turbine 1 turns backwards (tipRate -3) against the same blade twist
    assert th[0] > 0 and th[1] != 0 and np.all(np.abs(tq) > 0)
