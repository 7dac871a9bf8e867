"""The reference's case-directory and output formats, host side (Python).

Input (what a reference case directory holds):
  setup.json             runtime / inflow / solver / mesh / turbine blocks
                         (FalmCore::parse_settings, src/falm.h:222-340)
  mesh/ node lists       converted by the library's Mesher (falm_mesher_build,
                         src/mesher/mesher.hpp) into the control-volume file
  control-volume file    "imax jmax kmax gc" + centre/width lines
                         (src/FalmIO.h:99-133)
  turbine JSON           turbine array + property file with the blade table
                         (src/alm/alm.h:20-57, turbineHandler.h:121-205,
                         bladeHandler.h:16-104)

Output (what the C++ FalmCore writes, src/falm.h:446-479, src/FalmIO.h:25-97):
  PREFIX_RRRRRR_IIIIIIIIII       u, v, w, p of the padded domain (vector file)
  PREFIX_tavg_RRRRRR_IIIIIIIIII  their running time averages
  PREFIX.json                    index file; PREFIX.cv, PREFIX_setup.json copies
  PREFIX_PT.csv                  per-turbine power and thrust (src/alm/alm.h:126-138)

case_from_setup() turns a case directory into the falm_case configuration the
step driver runs (runtime.Case), with the mesh of its control-volume file and
the rotors of its turbine block, so reference case directories run on the
device in FP64 or mixed precision.  The boundary conditions are the wind
tunnel's (app/nedo/bcdevcall.cu, app/windtunnel/bcdevcall.cu): inlet at x-,
convective outflow at x+, slip y+-, no-slip z-, slip z+.
"""
import ctypes as C
import json
import math
import os

import numpy as np

from . import _abi as abi
from . import cases

GUIDE_CELL = 2


# ------------------------------------------------------------ file formats
def glue_path(workdir, name):
    """src/cpputil.hpp:6-14."""
    if name.startswith("/"):
        return name
    return workdir + name if workdir.endswith("/") else workdir + "/" + name


def build_mesh(cvcenter, meshpath, outputfile, gc=GUIDE_CELL):
    """Mesher::build_mesh through the library (host code, no device)."""
    from . import runtime
    runtime.call("falm_mesher_build", cvcenter.encode(), meshpath.encode(), outputfile.encode(), int(gc))


def read_cv_file(path):
    """Control-volume file -> dict(shape, gc, x, hx, y, hy, z, hz)."""
    with open(path) as f:
        tok = f.read().split()
    imax, jmax, kmax, gc = (int(t) for t in tok[:4])
    vals = np.array([float(t) for t in tok[4:4 + 2 * (imax + jmax + kmax)]], dtype=np.float64)
    if vals.size != 2 * (imax + jmax + kmax):
        raise ValueError(f"{path}: short control-volume file")
    pairs = vals.reshape(-1, 2)
    out = {"shape": (imax, jmax, kmax), "gc": gc}
    o = 0
    for ax, n in zip("xyz", (imax, jmax, kmax)):
        out[ax], out["h" + ax] = pairs[o:o + n, 0].copy(), pairs[o:o + n, 1].copy()
        o += n
    return out


_HEADER = np.dtype([("imax", "<u8"), ("jmax", "<u8"), ("kmax", "<u8"), ("n", "<u8"), ("gc", "<u8"),
                    ("step", "<u8"), ("time", "<f8"), ("dsize", "<u8")])


def read_vector_file(path):
    """Vector file -> (header dict, array [n, kmax, jmax, imax]); the layout
    the reference's reconstructor and file-info script read
    (src/reconstructor/reconstructor.cpp:141-176, script/file-info.py)."""
    with open(path, "rb") as f:
        h = np.fromfile(f, dtype=_HEADER, count=1)
        if h.size != 1:
            raise ValueError(f"{path}: short header")
        hd = {k: (float(h[k][0]) if k == "time" else int(h[k][0])) for k in _HEADER.names}
        dt = {4: np.float32, 8: np.float64}[hd["dsize"]]
        count = hd["n"] * hd["kmax"] * hd["jmax"] * hd["imax"]
        data = np.fromfile(f, dtype=dt, count=count)
    if data.size != count:
        raise ValueError(f"{path}: short data")
    return hd, data.reshape(hd["n"], hd["kmax"], hd["jmax"], hd["imax"])


def snapshot_path(workdir, prefix, step, rank=0, tavg=False):
    """FalmCore::outputUVWP file name (src/falm.h:455-473)."""
    p = prefix + ("_tavg" if tavg else "")
    return glue_path(workdir, "%s_%06d_%010d" % (p, rank, step))


def read_index_file(path):
    with open(path) as f:
        return json.load(f)


# ------------------------------------------------------------ setup parsing
def parse_settings(setup_path):
    """FalmCore::parse_settings (src/falm.h:222-340): the parsed JSON plus
    the derived iteration counts, with the reference's truncations."""
    cut = setup_path.rfind("/")
    if cut < 0:
        workdir, setup_file = ".", setup_path
    elif cut == 0:
        workdir, setup_file = "/", setup_path[1:]
    else:
        workdir, setup_file = setup_path[:cut], setup_path[cut + 1:]
    with open(glue_path(workdir, setup_file)) as f:
        params = json.load(f)
    rt = params["runtime"]
    start, dt = float(rt["time"]["start"]), float(rt["time"]["dt"])
    s = {"params": params, "workdir": workdir, "setupFile": setup_file, "startTime": start, "dt": dt,
         "maxIt": int((float(rt["time"]["end"]) - start) / dt)}
    if "timeAvg" in rt:
        ta = rt["timeAvg"]
        s["timeAvgStartIt"] = int((max(float(ta["start"]), start) - start) / dt) if "start" in ta else 0
        s["timeAvgEndIt"] = int((float(ta["end"]) - start) / dt) if "end" in ta else s["maxIt"]
    else:
        s["timeAvgStartIt"], s["timeAvgEndIt"] = 0, -1
    if "output" in rt:
        o = rt["output"]
        s["outputStartIt"] = int((max(float(o["start"]), start) - start) / dt) if "start" in o else 0
        s["outputEndIt"] = int((float(o["end"]) - start) / dt) if "end" in o else s["maxIt"]
        s["outputIntervalIt"] = int(float(o["interval"]) / dt)
        s["outputPrefix"] = o["prefix"]
    else:
        s["outputStartIt"], s["outputEndIt"], s["outputIntervalIt"], s["outputPrefix"] = 0, -1, 1, ""
    return s


_LS = {"PBiCGStab": abi.LS_PBICGSTAB, "SOR": abi.LS_SOR, "Jacobi": abi.LS_JACOBI}
_ADV = {"Upwind1": abi.ADV_UPWIND1, "Upwind3": abi.ADV_UPWIND3, "QUICK": abi.ADV_QUICK,
        "UTOPIA": abi.ADV_UTOPIA, "KK": abi.ADV_KK}


class BaseTrack:
    """A base coordinate over time from a "t value" text file ('#' comments):
    linear on t[q] <= t < t[q+1], constant before the first and from the
    last sample on; the segment's slope is the base velocity.  The
    reference parses such file names but never reads them
    (src/alm/turbineHandler.h:144-150)."""

    def __init__(self, path):
        self.path = path
        t, v = [], []
        with open(path) as f:
            for line in f:
                line = line.split("#", 1)[0].replace(",", " ").split()
                if len(line) >= 2:
                    t.append(float(line[0]))
                    v.append(float(line[1]))
        if not t or any(b <= a for a, b in zip(t, t[1:])):
            raise ValueError(f"{path}: need increasing sample times")
        self.t, self.v = t, v

    def eval(self, x):
        t, v = self.t, self.v
        if x < t[0]:
            return v[0], 0.0
        if x >= t[-1]:
            return v[-1], 0.0
        q = 0
        while t[q + 1] <= x:
            q += 1
        slope = (v[q + 1] - v[q]) / (t[q + 1] - t[q])
        return v[q] + slope * (x - t[q]), slope


def build_ap(blade, n_turbine, n_blade, n_ap_per_blade, radius, hub_radius):
    """BladeHandler::buildAP values (falm_build_ap, host code): per-AP r,
    chord, twist and Cl/Cd [attack][ap] from the blade table JSON."""
    from . import runtime
    af, attack = blade["airfoils"], np.array(blade["attacks"], dtype=np.float64)
    nst, nat = len(af), attack.size
    rr = np.array([a["r/R"] for a in af], dtype=np.float64)
    ch = np.array([a["chord/R"] for a in af], dtype=np.float64)
    tw = np.array([a["twist[deg]"] for a in af], dtype=np.float64)
    cl = np.ascontiguousarray([a["Cl"][:nat] for a in af], dtype=np.float64)
    cd = np.ascontiguousarray([a["Cd"][:nat] for a in af], dtype=np.float64)
    nap = n_turbine * n_blade * n_ap_per_blade
    out = [np.zeros(nap), np.zeros(nap), np.zeros(nap), np.zeros(nap * nat), np.zeros(nap * nat)]
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    runtime.call("falm_build_ap", nst, nat, vp(rr), vp(ch), vp(tw), vp(cl), vp(cd), n_turbine, n_blade,
                 n_ap_per_blade, float(radius), float(hub_radius), *[vp(o) for o in out])
    return attack, out


def alm_from_turbine_json(workdir, tp, cutoff=0.0):
    """AlmHandler::init(workdir, turbine JSON) (src/alm/alm.h:20-57):
    AlmSetup plus {(turbine, axis): BaseTrack} for file-driven bases."""
    with open(glue_path(workdir, tp["turbineProperties"])) as f:
        prop = json.load(f)
    radius = float(prop["radius"])
    hub_radius = float(prop.get("hubRadius", 0.0))
    n_blade = int(prop["bladeNumber"])
    napb = int(tp["bladePointNumber"])
    turbines, tracks = [], {}
    for tid, tj in enumerate(tp["turbineArray"]):
        base = [0.0, 0.0, 0.0]
        for i, b in enumerate(tj["baseLocation"][:3]):
            if isinstance(b, str):
                tracks[(tid, i)] = BaseTrack(glue_path(workdir, b))
                base[i] = tracks[(tid, i)].eval(0.0)[0]
            else:
                base[i] = float(b)
        motion, sw = [0.0, 0.0, 0.0], [0, 0, 0]
        for i, a in enumerate(tj["angle"][:3]):
            if isinstance(a, (int, float)):
                if a != 0:
                    motion = [float(a) * math.pi / 180.0, 0.0, 0.5 * math.pi]
                    sw[i] = 1
            else:
                motion = [float(a["amplitude"]) * math.pi / 180.0, 2.0 * math.pi / float(a["period"]),
                          (math.pi / 180.0) * float(a["phase"])]
                sw[i] = 1
        angle_type = 0 if sum(sw) != 1 else sw.index(1) + 1
        hub = (-float(prop["overhang"]), 0.0, float(prop["tower"]))
        turbines.append(cases.make_turbine(base, hub=hub, tip_rate=float(tj["tipRate"]), angle_type=angle_type,
                                           motion=motion))
    attack, (r, chord, twist, cl, cd) = build_ap(prop["bladeProperties"], len(turbines), n_blade, napb, radius,
                                                 hub_radius)
    alm = cases.AlmSetup(turbines, n_blade, napb, radius, hub_radius, float(tp["projectionWidth"]), r, chord,
                         twist, attack, cl, cd, cutoff=cutoff)
    return alm, tracks


def case_from_setup(setup_path, precision=abi.CASE_F64, build=True, solver_kw=None):
    """A reference case directory as a falm_case configuration (CaseSetup).

    The mesh comes from the control-volume file, converted from the node
    list first when the setup asks for it (and build is true).  Extra
    attributes: .settings (parse_settings), .cv (read_cv_file), .tracks
    (file-driven turbine bases).  The inlet velocity is the setup's inflow
    velocity (the reference's nedo app writes 1.0 there whatever the setup
    says, app/nedo/bcdevcall.cu:16)."""
    st = parse_settings(setup_path)
    p, wd = st["params"], st["workdir"]
    mesh = p["mesh"]
    cv_path = glue_path(wd, mesh["controlVolumeFile"])
    if "convert" in mesh and build:
        build_mesh(mesh["controlVolumeCenter"], glue_path(wd, mesh["convert"]), cv_path, GUIDE_CELL)
    cv = read_cv_file(cv_path)
    if cv["gc"] != GUIDE_CELL:
        raise ValueError(f"{cv_path}: {cv['gc']} guide cells, the step driver uses {GUIDE_CELL}")
    s = abi.CaseConfig()
    s.kind = abi.CASE_WINDTUNNEL
    keep = []
    for m, ax in enumerate("xyz"):
        s.n[m] = cv["shape"][m] - 2 * GUIDE_CELL
        for q, arr in enumerate((cv[ax], cv["h" + ax])):
            a = np.ascontiguousarray(arr, dtype=np.float64)
            keep.append(a)
            s.cv[2 * m + q] = a.ctypes.data_as(C.POINTER(C.c_double))
    cfd = p["solver"]["cfd"]
    s.Re, s.dt = float(cfd["Re"]), st["dt"]
    s.adv_scheme = _ADV.get(cfd["advectionScheme"], abi.ADV_UPWIND1)
    sgs = cfd.get("SGS", "Empty")
    s.sgs_model = abi.SGS_SMAGORINSKY if sgs == "Smagorinsky" else abi.SGS_EMPTY
    s.Cs = float(cfd["Cs"]) if sgs == "Smagorinsky" else 0.1
    s.u_inflow = float(p["inflow"]["velocity"]) if "inflow" in p else 1.0
    s.zminus_slip = 0
    s.precision = precision
    ls = p["solver"]["linearSolver"]
    kw = dict(type=_LS.get(ls["type"], abi.LS_EMPTY), maxit=int(ls["iteration"]), tol=float(ls["tolerance"]),
              relax=float(ls.get("relaxationFactor", 1.0)), pc_type=abi.LS_EMPTY, pc_maxit=5, pc_relax=1.0,
              precision=abi.SOLVE_MIXED if precision == abi.CASE_MIXED else abi.SOLVE_F64)
    if kw["type"] == abi.LS_PBICGSTAB:
        pc = ls["preconditioner"]
        kw.update(pc_type=_LS.get(pc["type"], abi.LS_EMPTY), pc_maxit=int(pc["iteration"]),
                  pc_relax=float(pc.get("relaxationFactor", 1.0)))
    kw.update(solver_kw or {})
    s.solver = abi.solver_config(**kw)
    alm, tracks = (None, {})
    if "turbine" in p:
        alm, tracks = alm_from_turbine_json(wd, p["turbine"])
    setup = cases.CaseSetup(s, alm, name=f"case directory {wd}")
    setup.settings, setup.cv, setup.tracks, setup._keep = st, cv, tracks, keep
    return setup


def base_motion(setup, t):
    """{turbine: (base[3], velocity[3])} of the file-driven bases at time t."""
    out = {}
    for (tid, axis), tr in setup.tracks.items():
        if tid not in out:
            T = setup.alm.turbines[tid]
            out[tid] = ([T.base[0], T.base[1], T.base[2]], [0.0, 0.0, 0.0])
        out[tid][0][axis], out[tid][1][axis] = tr.eval(t)
    return out
