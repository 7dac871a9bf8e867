"""Writes the file-format fixtures under tests/golden/casedir and tests/golden/io.

Inputs (written by this script, synthetic, in the reference's formats):
  casedir/                 a small wind-turbine case directory: setup.json,
                           mesh/ node lists, turbineProperties.json, a base
                           motion file
  io/mesh2/                node lists with a single z cell
  io/json_doc.json         a JSON document with the value kinds the setup,
                           index and AP files use
  io/numbers*.txt          doubles to print as JSON numbers

Expected outputs (written by oracle/_ref/ref_tools, i.e. by the reference's
own Mesher and its vendored JSON library, built by oracle/build.py):
  io/*.cv                  control-volume files from Mesher::build_mesh
  io/*.sorted, *.ordered   json / ordered_json parse + dump(2)
  io/numbers*.ref          json(double).dump() per number

    python oracle/build.py && python tests/golden/make_io_golden.py
"""
import json
import math
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_tools")
CASE = os.path.join(HERE, "casedir")
IO = os.path.join(HERE, "io")


def write_nodes(path, nodes):
    with open(path, "w") as f:
        f.write(f"{len(nodes)}\n")
        for v in nodes:
            f.write(f"{v:.15e}\n")


def case_inputs():
    os.makedirs(os.path.join(CASE, "mesh"), exist_ok=True)
    with open(os.path.join(CASE, "mesh", "mesh.json"), "w") as f:
        f.write('{\n    "coordinatePoint":"node"\n}\n')
    # x: 0.1 spacing, the last six gaps stretched by 1.15; y: uniform;
    # z: gaps growing by 3 % from the ground
    x = [-1.0 + 0.1 * i for i in range(36)]
    for k in range(6):
        x.append(x[-1] + 0.1 * 1.15 ** (k + 1))
    y = [-1.05 + 0.1 * j for j in range(22)]
    z = [-0.04]
    for k in range(21):
        z.append(z[-1] + 0.08 * 1.03 ** k)
    for ax, v in zip("xyz", (x, y, z)):
        write_nodes(os.path.join(CASE, "mesh", ax + ".txt"), v)

    attacks = [float(a) for a in range(-8, 21, 2)]
    stations = [0.0, 0.08, 0.16, 0.24, 0.32, 0.4]  # absolute radius (read as r/R by the reference)
    chord = [0.1, 0.09, 0.07, 0.05, 0.04, 0.03]
    twist = [25.0, 18.0, 12.0, 8.0, 5.0, 3.0]
    airfoils = []
    for s, (r, c, t) in enumerate(zip(stations, chord, twist)):
        scale = 1.0 - 0.05 * s
        cl = [round(scale * 2 * math.pi * math.radians(max(-12.0, min(12.0, a))), 6) for a in attacks]
        cd = [round(0.01 + 0.05 * math.radians(a) ** 2, 6) for a in attacks]
        airfoils.append({"r/R": r, "chord/R": c, "twist[deg]": t, "Cl": cl, "Cd": cd})
    prop = {"radius": 0.4, "hubRadius": 0.05, "bladeNumber": 3, "tower": 1.0, "overhang": 0.05,
            "bladeProperties": {"airfoils": airfoils, "attacks": attacks}}
    with open(os.path.join(CASE, "turbineProperties.json"), "w") as f:
        json.dump(prop, f, indent=4)
    with open(os.path.join(CASE, "base_x.txt"), "w") as f:
        f.write("# t  x of the base of turbine 1\n0.0 0.0\n0.05 0.02\n0.2 -0.03\n")
    setup = {
        "runtime": {"time": {"start": 0, "end": 0.29, "dt": 0.01}, "timeAvg": {"start": 0.05},
                    "output": {"start": 0.0, "interval": 0.1, "prefix": "run"}},
        "inflow": {"type": "uniform", "velocity": 1.0},
        "solver": {"linearSolver": {"type": "PBiCGStab", "iteration": 500, "tolerance": 1e-12,
                                    "preconditioner": {"type": "SOR", "iteration": 2, "relaxationFactor": 1.5}},
                   "cfd": {"Re": 1e4, "advectionScheme": "Upwind3", "SGS": "Smagorinsky", "Cs": 0.1}},
        "mesh": {"convert": "mesh", "controlVolumeFile": "controlVolume.txt", "controlVolumeCenter": "innerNode"},
        "turbine": {"turbineProperties": "turbineProperties.json", "bladePointNumber": 8, "projectionWidth": 0.2,
                    "turbineArray": [
                        {"baseLocation": [0, -0.45, 0],
                         "angle": [0, {"amplitude": 3, "period": 0.5, "phase": 10}, 0], "tipRate": 4},
                        {"baseLocation": ["base_x.txt", 0.45, 0], "angle": [2, 0, 0], "tipRate": -3}],
                    "apFile": "ap.json", "writePowerThrust": True},
    }
    with open(os.path.join(CASE, "setup.json"), "w") as f:
        json.dump(setup, f, indent=4)


def io_inputs():
    os.makedirs(os.path.join(IO, "mesh2"), exist_ok=True)
    with open(os.path.join(IO, "mesh2", "mesh.json"), "w") as f:
        f.write('{"coordinatePoint": "node"}\n')
    write_nodes(os.path.join(IO, "mesh2", "x.txt"), [0.0, 0.25, 0.5, 0.8, 1.2])
    write_nodes(os.path.join(IO, "mesh2", "y.txt"), [0.0, 0.5, 1.0])
    write_nodes(os.path.join(IO, "mesh2", "z.txt"), [0.0, 0.3])
    doc = ('{"zeta": [1, -2, 3.5, 1e-6, 2.5e5, 0.1, 1E3, -0.0, 12345678901234, 0.30000000000000004],\n'
           ' "alpha": {"b": true, "a": false, "n": null, "s": "tab\\t quote\\" slash\\\\ \\u00e9 \\/ nl\\n"},\n'
           ' "empty": {"arr": [], "obj": {}}, "dup": 1, "dup": 2,\n'
           ' "Mixed": [{"t": 0.04, "step": 4}, {"t": 0.08, "step": 8, "timeAvg": 1}], "u": "\\u0001"}\n')
    with open(os.path.join(IO, "json_doc.json"), "w") as f:
        f.write(doc)
    nums = ["0.1", "2.5e-4", "1e-6", "250000", "1e16", "1e15", "0.3", "-0.0", "1e-300", "5e-324",
            "1.7976931348623157e308", "0.0001", "0.00001", "12.5", "100", "3.14159", "0.30000000000000004",
            "0.04", "0.08", "0.12", "1e4", "2.5e5", "6.283185307179586", "0.05235987755982988", "-123.456",
            "1e21", "1e-5", "9.999999999999999e-05", "123456789.125", "0.0012345"]
    with open(os.path.join(IO, "numbers.txt"), "w") as f:
        f.write(" ".join(nums) + "\n")
    # random bit patterns and decimals of assorted sizes
    import random
    import struct
    rnd = random.Random(20241202)
    vals = []
    while len(vals) < 2000:
        r = rnd.random()
        if r < 0.5:
            v = struct.unpack("<d", struct.pack("<Q", rnd.getrandbits(64)))[0]
        elif r < 0.8:
            v = rnd.random() * 10.0 ** rnd.randint(-20, 20)
        else:
            v = round(rnd.random() * 10.0 ** rnd.randint(0, 6), rnd.randint(0, 8))
        if math.isfinite(v):
            vals.append(repr(v))
    with open(os.path.join(IO, "numbers_random.txt"), "w") as f:
        f.write(" ".join(vals) + "\n")


def run(args, out=None):
    r = subprocess.run([REF] + args, check=True, capture_output=True, text=True)
    if out:
        with open(out, "w") as f:
            f.write(r.stdout)


def expected():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/ref_tools is not built (python oracle/build.py, needs the reference tree)")
    mesh = os.path.join(CASE, "mesh")
    run(["mesh", "innerNode", mesh, os.path.join(IO, "casedir_innerNode_gc2.cv"), "2"])
    run(["mesh", "innerNode", mesh, os.path.join(IO, "casedir_innerNode_gc3.cv"), "3"])
    run(["mesh", "cell", mesh, os.path.join(IO, "casedir_cell_gc2.cv"), "2"])
    run(["mesh", "cell", os.path.join(IO, "mesh2"), os.path.join(IO, "mesh2_cell_gc2.cv"), "2"])
    for name in ("json_doc.json",):
        src = os.path.join(IO, name)
        run(["json", src], os.path.join(IO, name + ".sorted"))
        run(["ojson", src], os.path.join(IO, name + ".ordered"))
    for name in ("setup.json", "turbineProperties.json"):
        run(["json", os.path.join(CASE, name)], os.path.join(IO, "casedir_" + name + ".sorted"))
    for name in ("numbers", "numbers_random"):
        with open(os.path.join(IO, name + ".txt")) as f:
            nums = f.read().split()
        run(["num"] + nums, os.path.join(IO, name + ".ref"))


if __name__ == "__main__":
    case_inputs()
    io_inputs()
    expected()
