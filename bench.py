#!/usr/bin/env python3
"""bench.py — throughput of the fractional-step LES step on one MI355X.

Workload: wind-tunnel channel of 512x128x128 cells (dx = D/32) with one
actuator-line rotor (3 blades x 32 points, TSR 7), Upwind3 + Smagorinsky,
BiCGStab with a red-black SOR preconditioner.  The headline figure is MLUPS
of the mixed-precision step (float stencils and Krylov passes, double defect
correction) with the Poisson solve stopped at |r|/N <= 1.5e-8 of the scaled
system, the loosest of 1e-8, 1.5e-8, 2e-8, 3e-8, 5e-8 at which the divergence
monitor sqrt(sum div^2)/N (app/windtunnel/main.cpp:185) stays below 1e-6 in
both precisions (7.9e-7 after the timed steps, 5 BiCGStab iterations per
step; 2e-8 gives 1.06e-6); the FP64 step runs beside it on the same
workload with the same tolerance.  A second pair of legs stops the
solve at |r|/N <= 1e-6, the reference's criterion (src/FalmEq.cpp:330),
where one BiCGStab iteration per step suffices and the monitor reads ~1e-5.
A step is one full time step (ALM, pseudo velocity, face flux, Poisson RHS,
Poisson solve, projection, boundary fill, SGS, divergence monitor) with all
fields resident in HBM.

The accuracy leg runs both precisions for 1000 steps from the same start
and reports the mixed step's velocity and pressure error against the FP64
step, beside the FP64 step's own difference from an FP64 run at a ten times
tighter Poisson tolerance (the error the tolerance itself allows; the
channel's pressure has a smooth near-null mode that |r|/N barely
constrains).  It uses CFL 0.1: the reference's pseudo-velocity step is explicit
Euler with Upwind3 (ua = un + dt(-adv + vis + ff), src/FalmCFD.cpp:5-67),
which amplifies resolved modes at CFL 0.3 (|u|max grows about 4x by step
400 and overflows near step 450 at 256x96x96, in the GPU path and in the
CPU restatement alike); the timed steps (the first ~30) are unaffected.

Per-kernel HBM rooflines come from HIP events around each launch of one
kernel family at a time (extra steps after the timed ones, so the timed
steps carry no event overhead), with algorithmic bytes per launch from the
operand-list model (every array a kernel touches counted once per cell).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--only mixed|fp64]
                    [--tol T] [--solver-tol-leg T2] [--no-cpu] [--calibrate]
                    [--accuracy-steps N]
    python bench.py --adjoint     (rotor-power gradient report, 256x96x96)
    python bench.py --farm        (floating tandem-rotor segment, 1024x256x256)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + steps/s @512x128x128 ALM: mixed vs FP64 GPU at div<1e-6; HBM GB/s"
PEAK_HBM_GBS = 8000.0
DOMINANT = "pseudo_u"  # the largest kernel of the step


def families():
    from falm_amd import _abi as abi
    return {
        "pseudo_u": (abi.TIME_PSEUDO_U, "k_pseudo_u (Upwind3 + diffusion + force)"),
        "project": (abi.TIME_PROJECT, "k_project (pressure-gradient projection, cell + face)"),
        "face_flux": (abi.TIME_FACE_FLUX, "k_u_to_uu_rhs (face flux + Poisson right-hand side)"),
        "rhs": (abi.TIME_RHS, "k_divergence (Poisson right-hand side)"),
        "sgs": (abi.TIME_SGS, "k_sgs_div (Smagorinsky + divergence + |div|^2)"),
        "divergence": (abi.TIME_DIVERGENCE, "k_div_norm (divergence + |div|^2)"),
        "solver_init": (abi.TIME_SOLVER_INIT, "k_bicg_init (b - A x in double + dots)"),
        "sor": (abi.TIME_SOR, "k_sor_op_km / k_sor3_km (separate preconditioner sweeps, pc_maxit >= 2)"),
        "mv": (abi.TIME_MV, "k_bicg_mv_km<., ., 1> (q = A p^ + (rr, q); preconditioners other than one SOR pass)"),
        "mv2": (abi.TIME_MV2, "k_bicg_mv_km<., ., 2> (t = A s^ + (t, s), (t, t); the same)"),
        "pcmv": (abi.TIME_PCMV, "k_bicg_pcmv_km<., ., 0> (r = s - omega t, p = r + beta(p - omega q), "
                                "q = A M(p), (rr, q), y += alpha p + omega s; one-pass RB-SOR M on a two-cell ring)"),
        "pcmv_s": (abi.TIME_PCMV_S, "k_bicg_pcmv_km<., ., 1> (s = r - alpha q, t = A M(s), (t, s), (t, t), (s, s), "
                                    "(s, rr), (t, rr))"),
        "pc_apply": (abi.TIME_PC_APPLY, "k_bicg_close_km (end of a pass: x += M(y + alpha p + omega s))"),
        "solver_update": (abi.TIME_SOLVER_UPDATE, "k_bicg_update (x += alpha p^ + omega s^, r = s - omega t + dots; "
                                                  "preconditioners other than one SOR pass)"),
        "alm": (abi.TIME_ALM, "ALM stage (points, box, factors, projection, totals): a latency chain, "
                              "bytes = point tables + force-box cells written"),
    }


# stages whose time is launch latency rather than streaming: reported in us,
# with no bandwidth fraction
LATENCY_FAMILIES = ("alm",)


def run_gpu(precision, steps, warmup, n, fam_steps, tol):
    """Every falm_case_step call returns after synchronising the case's
    stream, so the wall clock brackets completed device work."""
    from falm_amd import _abi as abi
    from falm_amd import cases, runtime

    setup = cases.windtunnel(n=n, precision=precision, tol=tol)
    c = runtime.Case(setup)
    for _ in range(warmup):
        c.step(1)
    its, outer = [], []
    t0 = time.perf_counter()
    info = None
    for _ in range(steps):
        info = c.step(1)
        its.append(info.solver_it)
        outer.append(info.outer_it)
    dt = time.perf_counter() - t0
    u = c.get(abi.FIELD_U)
    kern = {}
    for name, (fam, desc) in families().items():
        runtime.timing_enable(fam)
        for _ in range(fam_steps):
            c.step(1)
        launches, kms, kbytes = runtime.timing_read()
        runtime.timing_enable(abi.TIME_NONE)
        if launches == 0 or kms <= 0:
            continue
        gbs = kbytes / (kms * 1e-3) / 1e9
        kern[name] = {"kernel": desc, "launches_per_step": launches / fam_steps,
                      "avg_launch_us": 1e3 * kms / launches, "bytes_per_launch": kbytes / launches,
                      "gbs": gbs, "frac": gbs / PEAK_HBM_GBS}
        if name in LATENCY_FAMILIES:
            kern[name].update(bound="latency", frac=None)
    c.close()
    cells = setup.n_interior
    return {
        "tol": tol,
        "ms_per_step": 1e3 * dt / steps,
        "steps_per_s": steps / dt,
        "mlups": cells * steps / dt / 1e6,
        "solver_it_per_step": sum(its) / len(its),
        "defect_passes_per_step": sum(outer) / len(outer),
        "div": info.div,
        "solver_err": info.solver_err,
        "status": info.status,
        "kernels": kern,
    }, u, setup


def calibrate_copy(nbytes=1 << 30, reps=3):
    """A plain streaming kernel of known traffic (falm_scale over a double
    array: read + write of nbytes each) for scaling HBM counter readings."""
    import ctypes as C

    from falm_amd import _abi as abi
    from falm_amd import runtime
    p = C.c_void_p()
    runtime.call("falm_malloc", C.byref(p), nbytes)
    for _ in range(reps):
        runtime.call("falm_scale", abi.F64, p, nbytes // 8, 1.0, None)
    runtime.call("falm_stream_sync", None)
    runtime.call("falm_free", p)
    return {"kernel": "k_scale", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps}


def stream_reference(shape, reps=20):
    """What plain streaming reaches on this box at the benchmark's array
    size: falm_vecaxby (r = a x + b y, one element per thread; two arrays
    read, one written) over the padded 512x128x128 field in float and double,
    and falm_scale over 1 GiB of doubles, timed with HIP events on the null
    stream the calls run on."""
    import ctypes as C

    from falm_amd import _abi as abi
    from falm_amd import runtime

    n = shape[0] * shape[1] * shape[2]
    out = {}
    hip = hip_runtime()
    ev = [C.c_void_p(), C.c_void_p()]
    for e in ev:
        if hip.hipEventCreate(C.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")

    def timed(fn):
        fn()
        runtime.call("falm_stream_sync", None)
        hip.hipEventRecord(ev[0], None)
        for _ in range(reps):
            fn()
        hip.hipEventRecord(ev[1], None)
        hip.hipEventSynchronize(ev[1])
        ms = C.c_float()
        if hip.hipEventElapsedTime(C.byref(ms), ev[0], ev[1]) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value / reps * 1e-3  # s

    for name, prec, tb in (("float", abi.F32, 4), ("double", abi.F64, 8)):
        bufs = [C.c_void_p() for _ in range(3)]
        for b in bufs:
            runtime.call("falm_malloc", C.byref(b), n * tb)
        t = timed(lambda: runtime.call("falm_vecaxby", prec, 1.0, bufs[0], 0.5, bufs[1], bufs[2], n, None))
        gbs = 3 * n * tb / t / 1e9
        out[f"vecaxby_{name}"] = {"bytes_per_launch": 3 * n * tb, "avg_launch_us": t * 1e6, "gbs": gbs,
                                  "frac": gbs / PEAK_HBM_GBS}
        for b in bufs:
            runtime.call("falm_free", b)
    nb = 1 << 30
    p = C.c_void_p()
    runtime.call("falm_malloc", C.byref(p), nb)
    t = timed(lambda: runtime.call("falm_scale", abi.F64, p, nb // 8, 1.0, None))
    runtime.call("falm_free", p)
    gbs = 2 * nb / t / 1e9
    out["scale_double_1GiB"] = {"bytes_per_launch": 2 * nb, "avg_launch_us": t * 1e6, "gbs": gbs,
                                "frac": gbs / PEAK_HBM_GBS}
    for e in ev:
        hip.hipEventDestroy(e)
    return out


def cpu_baseline(n, steps=2, warmup=1):
    """The oracle (FP64 C++ restatement of the reference scheme, OpenMP) on
    the host cores: a bounded sample of the same workload."""
    from falm_amd import cases
    from oracle import pyoracle as O

    setup = cases.windtunnel(n=n)
    c = O.Case(setup)
    for _ in range(warmup):
        c.step(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.step(1)
    dt = time.perf_counter() - t0
    return {
        "value": setup.n_interior * steps / dt / 1e6,
        "unit": "MLUPS",
        "cores": O.num_threads(),
        "cpu_model": cpu_model(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "kind": "port",
        "sample": f"{steps} steps after {warmup} warm-up step(s) of the same 512x128x128 + rotor workload, FP64, "
                  f"{dt / steps:.2f} s/step",
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_accuracy(n, steps, cfl, tol):
    """Mixed against FP64 on the GPU after `steps` steps of the same
    workload from the same initial state: relative L2 errors of velocity
    and pressure over the interior, and each run's divergence monitor."""
    import numpy as np

    from falm_amd import _abi as abi
    from falm_amd import cases, runtime

    out = {}
    for key, prec, t in (("fp64", abi.CASE_F64, tol), ("mixed", abi.CASE_MIXED, tol),
                         ("fp64_tight", abi.CASE_F64, 0.1 * tol)):
        setup = cases.windtunnel(n=n, precision=prec, cfl=cfl, tol=t)
        c = runtime.Case(setup)
        t0 = time.perf_counter()
        info = c.step(steps)
        dt = time.perf_counter() - t0
        out[key] = {"u": c.get(abi.FIELD_U), "p": c.get(abi.FIELD_P), "div": info.div,
                    "solver_err": info.solver_err, "status": info.status, "s": dt}
        c.close()
    nx, ny, nz = setup.shape
    mask = np.zeros((nz, ny, nx), dtype=bool)
    mask[2:-2, 2:-2, 2:-2] = True
    m1 = mask.ravel()
    m3 = np.tile(m1, 3)

    def rel(a, b, m):
        return float(np.linalg.norm(a[m] - b[m]) / np.linalg.norm(b[m]))

    f, x, ft = out["fp64"], out["mixed"], out["fp64_tight"]
    return {
        "workload": f"windtunnel {n[0]}x{n[1]}x{n[2]} + rotor, CFL {cfl}, Poisson tol {tol:g}, {steps} steps "
                    "from the same start",
        "velocity_rel_l2": rel(x["u"], f["u"], m3),
        "pressure_rel_l2": rel(x["p"], f["p"], m1),
        # the solver tolerance's own effect: FP64 at tol against FP64 at tol/10
        "fp64_tol_vs_tol_over_10": {"velocity_rel_l2": rel(f["u"], ft["u"], m3),
                                    "pressure_rel_l2": rel(f["p"], ft["p"], m1)},
        "div_fp64": f["div"], "div_mixed": x["div"],
        "u_max_fp64": float(np.abs(f["u"][m3]).max()), "u_max_mixed": float(np.abs(x["u"][m3]).max()),
        "status": [f["status"], x["status"]],
        "wall_s": [f["s"], x["s"]],
    }


def run_lid3d(n, steps, tol):
    """The 3-D lid-driven cavity LES (128^3, Re 1e4, Cs 0.1, dt 5e-4,
    BiCGStab + RB-SOR at |r|/N <= tol, app/lid3d/main.cpp:122-175) in FP64 and
    in mixed precision for `steps` steps from rest: the mixed run's velocity,
    pressure and divergence against the FP64 run, and each run's time."""
    import numpy as np

    from falm_amd import _abi as abi
    from falm_amd import cases, runtime

    out = {}
    for key, prec in (("fp64", abi.CASE_F64), ("mixed", abi.CASE_MIXED)):
        setup = cases.lid3d(n=n, precision=prec, tol=tol)
        c = runtime.Case(setup)
        t0 = time.perf_counter()
        for q in range(0, steps, 100):
            info = c.step(min(100, steps - q))
        dt = time.perf_counter() - t0
        out[key] = {"u": c.get(abi.FIELD_U), "p": c.get(abi.FIELD_P), "div_field": c.get(abi.FIELD_DIV),
                    "div": info.div, "solver_err": info.solver_err, "solver_it_last": info.solver_it,
                    "status": info.status, "s": dt}
        c.close()
    nx, ny, nz = setup.shape
    mask = np.zeros((nz, ny, nx), dtype=bool)
    mask[2:-2, 2:-2, 2:-2] = True
    m1 = mask.ravel()
    m3 = np.tile(m1, 3)
    f, x = out["fp64"], out["mixed"]
    # Neumann walls: the pressure is defined up to a constant
    pf, px = f["p"][m1] - f["p"][m1].mean(), x["p"][m1] - x["p"][m1].mean()
    return {
        "workload": f"lid-driven cavity LES {n}^3, Re 1e4, Cs 0.1, dt 5e-4, BiCGStab+RB-SOR(2 passes, w=1.5) "
                    f"tol {tol:g} on |r|/N, {steps} steps from rest",
        "velocity_rel_l2": float(np.linalg.norm(x["u"][m3] - f["u"][m3]) / np.linalg.norm(f["u"][m3])),
        "pressure_rel_l2_mean_free": float(np.linalg.norm(px - pf) / np.linalg.norm(pf)),
        "divergence_field_max_abs_diff": float(np.abs(x["div_field"][m1] - f["div_field"][m1]).max()),
        "div_fp64": f["div"], "div_mixed": x["div"],
        "status": [f["status"], x["status"]],
        "solver_it_last_step": [f["solver_it_last"], x["solver_it_last"]],
        "wall_s": [f["s"], x["s"]],
        "ms_per_step": [1e3 * f["s"] / steps, 1e3 * x["s"] / steps],
    }


def run_farm(n, steps, window, ckpt, fam_steps=2, adj_tol=1e-9, settle=5):
    """The floating tandem-rotor segment (two rotors 7 D apart, the upstream
    one in prescribed surge and pitch, 8 % inlet turbulence, CFL 0.1) in
    mixed precision and in FP64 for `steps` steps from the same start: MLUPS,
    per-kernel HBM rooflines of the mixed run, the mixed run's error against
    the FP64 run; then, on the FP64 run's state, the adjoint gradient of the
    summed power of both rotors over `window` steps with respect to the
    upstream rotor's yaw and its platform's pitch (the static part of the
    pitch angle), each against a central difference (h = 1e-4 rad)."""
    import numpy as np

    from falm_amd import _abi as abi
    from falm_amd import cases, runtime

    def note(msg):
        print(f"[farm] {msg}", file=sys.stderr, flush=True)

    out = {}
    fields = {}
    for key, prec in (("mixed", abi.CASE_MIXED), ("fp64", abi.CASE_F64)):
        setup = cases.farm_segment(n=n, precision=prec)
        c = runtime.Case(setup)
        its = []
        t0 = time.perf_counter()
        for q in range(0, steps, 100):
            info = c.step(min(100, steps - q))
            its.append(info.solver_it)
            note(f"{key}: {min(q + 100, steps)}/{steps} steps, {time.perf_counter() - t0:.1f} s, div {info.div:.3e}")
        dt = time.perf_counter() - t0
        rec = {"ms_per_step": 1e3 * dt / steps, "mlups": setup.n_interior * steps / dt / 1e6,
               "steps_per_s": steps / dt, "div": info.div, "solver_err": info.solver_err,
               "solver_it_last_step": info.solver_it, "status": info.status}
        _, _, tq, th = c.alm_get()
        rec["torque"], rec["thrust"] = [float(x) for x in tq], [float(x) for x in th]
        fields[key] = (c.get(abi.FIELD_U), c.get(abi.FIELD_P))
        if key == "mixed":
            kern = {}
            for name, (fam, desc) in families().items():
                runtime.timing_enable(fam)
                c.step(fam_steps)
                launches, kms, kbytes = runtime.timing_read()
                runtime.timing_enable(abi.TIME_NONE)
                if launches == 0 or kms <= 0:
                    continue
                gbs = kbytes / (kms * 1e-3) / 1e9
                kern[name] = {"launches_per_step": launches / fam_steps, "avg_launch_us": 1e3 * kms / launches,
                              "gbs": gbs, "frac": None if name in LATENCY_FAMILIES else gbs / PEAK_HBM_GBS}
            rec["kernels"] = kern
            c.close()
        out[key] = rec
    nx, ny, nz = setup.shape
    mask = np.zeros((nz, ny, nx), dtype=bool)
    mask[2:-2, 2:-2, 2:-2] = True
    m1 = mask.ravel()
    m3 = np.tile(m1, 3)
    (um, pm), (uf, pf) = fields["mixed"], fields["fp64"]
    out["velocity_rel_l2_mixed_vs_fp64"] = float(np.linalg.norm(um[m3] - uf[m3]) / np.linalg.norm(uf[m3]))
    out["pressure_rel_l2_mixed_vs_fp64"] = float(np.linalg.norm(pm[m1] - pf[m1]) / np.linalg.norm(pf[m1]))
    del fields
    # The adjoint and its central-difference check run on the FP64 run's case
    # with a tighter Poisson tolerance from here on (a few steps to settle):
    # the adjoint treats each Poisson solve as exact, and at tol 1e-6 the
    # solver's stopping noise in J is of the size of the central
    # differences' signal at h = 1e-4 rad.
    c.set_tolerance(adj_tol, 20000)
    info = c.step(settle)
    # In the window every Poisson solve runs a fixed number of BiCGStab
    # iterations, the number the last solve needed for adj_tol, so that J is
    # a smooth function of the parameters: a tolerance test lets the
    # stopping iteration jump between the central differences' two runs, and
    # iterating far past convergence makes BiCGStab's iterates (and J) vary
    # erratically from one parameter value to the next.
    fd_iters = max(1, info.solver_it)
    c.set_tolerance(1e-30, fd_iters)
    info = c.step(1)
    fixed_err = info.solver_err
    note(f"fixed {fd_iters} iterations per solve: |r|/N {fixed_err:.3e}")
    t0 = time.perf_counter()
    J0 = c.window(window)
    t_fwd = time.perf_counter() - t0
    J0b = c.window(window)  # the window restores the state: the same J again
    note(f"forward window J = {J0:.8g} (again: {J0b:.8g}), {t_fwd:.2f} s")
    t0 = time.perf_counter()
    J, gp, gt, nbytes = c.adjoint(window, ckpt, 1e-10)
    t_adj = time.perf_counter() - t0
    gy, go = c.adjoint_attitude()
    note(f"adjoint {t_adj:.2f} s, {nbytes / 2**30:.2f} GiB: dJ/dyaw {gy}, dJ/doffset {go}")
    checks = []
    for which, ad in (("yaw", gy[0]), ("pitch_offset", go[0])):
        def Jat(e):
            c.set_attitude(0, e if which == "yaw" else 0.0, e if which != "yaw" else 0.0)
            return c.window(window)
        for hr in (1e-4, 1e-3, 1e-2):
            h = float(np.degrees(hr))
            fd = (Jat(h) - Jat(-h)) / (2 * h)
            checks.append({"parameter": which, "fd_step_rad": hr, "adjoint_per_deg": float(ad), "fd_per_deg": fd,
                           "rel_err": abs(fd - ad) / abs(fd) if fd != 0 else None})
            note(f"{which}: adjoint {ad:.8e} central difference (h {hr:g} rad) {fd:.8e}")
        Jat(0.0)
    c.close()
    out["adjoint"] = {
        "objective": f"mean summed power of both rotors over {window} FP64 steps after the FP64 run's {steps} "
                     f"steps, {settle} more at Poisson tol {adj_tol:g} and one more; every Poisson solve in the "
                     f"window runs {fd_iters} BiCGStab iterations (the count the last solve at tol {adj_tol:g} "
                     f"needed)",
        "poisson_err_after_fixed_iterations": fixed_err,
        "J": J, "J_forward_window": J0, "J_forward_window_repeat": J0b,
        "grad_yaw_per_deg": [float(x) for x in gy], "grad_pitch_offset_per_deg": [float(x) for x in go],
        "grad_blade_pitch_per_deg": [float(x) for x in gp[0, :3]],
        "checks": checks,
        "time_forward_window_s": t_fwd, "time_adjoint_s": t_adj, "adjoint_device_bytes": nbytes,
    }
    out["workload"] = (f"farm segment {n[0]}x{n[1]}x{n[2]} cells (dx=D/32), rotors at 4D and 11D, upstream surge "
                       f"0.1D sin(2 pi t/T) and pitch 3 deg sin(2 pi t/T + pi/2), T = 20 revolutions, 8% inlet "
                       f"turbulence, CFL 0.1, Poisson tol 1e-6, {steps} steps")
    return out


def hip_runtime():
    """The HIP runtime the library already loaded (the same instance, so its
    device and streams are the library's)."""
    import ctypes as C

    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    if path is None:
        raise RuntimeError("libamdhip64 is not loaded")
    return C.CDLL(path)


def device_free_bytes():
    """hipMemGetInfo through the HIP runtime the library already loaded."""
    import ctypes as C

    hip = hip_runtime()
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return free.value


def run_adjoint(n, spinup, window, ckpt, ndir, tol, adj_rtol, cfl, seed=7):
    """The rotor-power gradient over a window of FP64 steps after a spin-up:
    J = mean power, dJ/d(3 blade pitches, n_ap_per_blade twist offsets) by the
    checkpointed device adjoint, checked in ndir seeded random unit
    directions against central differences of the device forward with
    h = 1e-4 rad; time and device memory against the forward window."""
    import numpy as np

    from falm_amd import cases, runtime

    # two SOR passes: the tight forward solves of the central differences
    # then stop at the same iteration
    setup = cases.windtunnel(n=n, cells_per_d=24, tol=tol, cfl=cfl, pc_maxit=2, pc_relax=1.5)
    free0 = device_free_bytes()
    c = runtime.Case(setup)
    case_bytes = free0 - device_free_bytes()  # the forward case's device footprint
    def note(msg):
        print(f"[adjoint] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    for q in range(0, spinup, 50):
        c.step(min(50, spinup - q))
        note(f"spin-up {min(q + 50, spinup)}/{spinup} steps, {time.perf_counter() - t0:.1f} s")
    t_spin = time.perf_counter() - t0
    t0 = time.perf_counter()
    J0 = c.window(window)
    t_fwd = time.perf_counter() - t0
    note(f"forward window J = {J0:.6g}, {t_fwd:.2f} s")
    t0 = time.perf_counter()
    J, gp, gt, nbytes = c.adjoint(window, ckpt, adj_rtol)
    t_adj = time.perf_counter() - t0
    note(f"adjoint {t_adj:.2f} s, {nbytes / 2**30:.2f} GiB")
    napb = setup.alm.n_ap_per_blade
    g = np.concatenate([gp[0, :3], gt])
    h = float(np.degrees(1e-4))
    rng = np.random.default_rng(seed)
    dirs = []
    t0 = time.perf_counter()
    for _ in range(ndir):
        v = rng.standard_normal(3 + napb)
        v /= np.linalg.norm(v)

        def Jat(e):
            c.set_pitch(0, e * v[:3])
            c.set_twist_offset(e * v[3:])
            return c.window(window)

        fd = (Jat(h) - Jat(-h)) / (2 * h)
        ad = float(g @ v)
        dirs.append({"fd_per_deg": fd, "adjoint_per_deg": ad, "rel_err": abs(fd - ad) / abs(fd)})
        note(f"direction {len(dirs)}: fd {fd:.6g} adjoint {ad:.6g}")
    Jat(0.0)
    t_fd = time.perf_counter() - t0
    c.close()
    return {
        "workload": f"windtunnel {n[0]}x{n[1]}x{n[2]} cells + 1 ALM rotor (3x{napb} points), dx=D/24, FP64, "
                    f"CFL {cfl}, Poisson tol {tol:g}, {spinup}-step spin-up, {window}-step window, checkpoints every {ckpt}",
        "J_mean_power": J,
        "J_forward_window": J0,
        "grad_pitch_per_deg": [float(x) for x in gp[0, :3]],
        "grad_twist_per_deg": [float(x) for x in gt],
        "directions": dirs,
        "max_rel_err_vs_fd": max(d["rel_err"] for d in dirs),
        "fd_step_rad": 1e-4,
        "time_spinup_s": t_spin,
        "time_forward_window_s": t_fwd,
        "time_adjoint_s": t_adj,
        "adjoint_time_vs_forward": t_adj / t_fwd,
        "time_fd_s": t_fd,
        "adjoint_device_bytes": nbytes,
        "forward_case_device_bytes": case_bytes,
        "adjoint_bytes_vs_forward_case": nbytes / case_bytes if case_bytes > 0 else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, nargs=3, default=[512, 128, 128])
    ap.add_argument("--only", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--fam-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=15)
    ap.add_argument("--calibrate", action="store_true")
    ap.add_argument("--adjoint", action="store_true",
                    help="rotor-power gradient report (256x96x96, D/24) instead of the step benchmark")
    ap.add_argument("--farm", action="store_true",
                    help="floating tandem-rotor segment report (1024x256x256, two rotors) instead of the step benchmark")
    ap.add_argument("--farm-n", type=int, nargs=3, default=[1024, 256, 256])
    ap.add_argument("--farm-steps", type=int, default=2000)
    ap.add_argument("--farm-window", type=int, default=10)
    ap.add_argument("--farm-ckpt", type=int, default=5)
    ap.add_argument("--adj-spinup", type=int, default=500)
    ap.add_argument("--adj-window", type=int, default=200)
    ap.add_argument("--adj-ckpt", type=int, default=20)
    ap.add_argument("--adj-dirs", type=int, default=4)
    ap.add_argument("--adj-tol", type=float, default=1e-12)
    ap.add_argument("--adj-cfl", type=float, default=0.1)
    ap.add_argument("--tol", type=float, default=1.5e-8,
                    help="Poisson tolerance on |r|/N of the headline legs (1.5e-8 keeps sqrt(sum div^2)/N < 1e-6)")
    ap.add_argument("--solver-tol-leg", type=float, default=1e-6,
                    help="second pair of legs stopped at this |r|/N (0: skip)")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="mixed-vs-FP64 error after this many steps (0: skip)")
    ap.add_argument("--accuracy-cfl", type=float, default=0.1)
    ap.add_argument("--lid3d-steps", type=int, default=1000,
                    help="cavity LES mixed-vs-FP64 leg: steps (0: skip)")
    ap.add_argument("--lid3d-n", type=int, default=128)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from falm_amd import _abi as abi
    from falm_amd import runtime

    runtime.init(0)
    if args.farm:
        rep = run_farm(tuple(args.farm_n), args.farm_steps, args.farm_window, args.farm_ckpt)
        print(json.dumps({"metric": "MLUPS + mixed vs FP64 error + attitude gradient: floating tandem-rotor segment",
                          "unit": "MLUPS", "value": rep["mixed"]["mlups"], "n_gpus": 1, "farm": rep}))
        return
    if args.adjoint:
        rep = run_adjoint((256, 96, 96), args.adj_spinup, args.adj_window, args.adj_ckpt, args.adj_dirs,
                          args.adj_tol, 1e-10, args.adj_cfl)
        print(json.dumps({"metric": "rotor-power adjoint gradient vs central differences",
                          "unit": "relative error", "n_gpus": 1, "adjoint": rep}))
        return
    n = tuple(args.n)
    calib = calibrate_copy() if args.calibrate else None

    def leg(tol):
        """Both precisions at one Poisson tolerance; velocity difference of
        the two after the timed and timing steps."""
        r64 = rmx = u64 = umx = setup = None
        if args.only in ("both", "fp64"):
            r64, u64, setup = run_gpu(abi.CASE_F64, args.steps, args.warmup, n, args.fam_steps, tol)
        if args.only in ("both", "mixed"):
            rmx, umx, setup = run_gpu(abi.CASE_MIXED, args.steps, args.warmup, n, args.fam_steps, tol)
        rel_l2 = None
        if r64 is not None and rmx is not None:
            nx, ny, nz = setup.shape
            mask = np.zeros((nz, ny, nx), dtype=bool)
            mask[2:-2, 2:-2, 2:-2] = True
            m3 = np.tile(mask.ravel(), 3)
            rel_l2 = float(np.linalg.norm(umx[m3] - u64[m3]) / np.linalg.norm(u64[m3]))
        return r64, rmx, rel_l2, setup

    # headline: the tolerance at which the divergence monitor
    # sqrt(sum div^2)/N stays below 1e-6 in both precisions
    r64, rmx, rel_l2, setup = leg(args.tol)
    head = rmx if rmx is not None else r64
    divs = [r["div"] for r in (r64, rmx) if r is not None]
    tol_leg = None
    if args.solver_tol_leg > 0:
        s64, smx, srel, _ = leg(args.solver_tol_leg)
        tol_leg = {
            "what": f"same workload with the solver stopped at |r|/N <= {args.solver_tol_leg:g} (the reference's "
                    "criterion, src/FalmEq.cpp:330) instead of at the divergence condition",
            "mixed": smx, "fp64": s64,
            "speedup_mixed_vs_fp64": (s64["ms_per_step"] / smx["ms_per_step"]) if (s64 and smx) else None,
            "velocity_rel_l2_mixed_vs_fp64": srel,
        }

    # dominant kernel: the streaming family (one kernel each) with the most
    # time per step
    stream_fams = {k_: v for k_, v in head["kernels"].items() if k_ not in LATENCY_FAMILIES}
    dom_name = max(stream_fams, key=lambda k_: stream_fams[k_]["avg_launch_us"] * stream_fams[k_]["launches_per_step"]) \
        if stream_fams else None
    dom = stream_fams.get(dom_name)
    roofline = None
    if dom:
        roofline = {
            "bound": "hbm",
            "kernel": dom["kernel"],
            "achieved": dom["gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": dom["frac"],
            "traffic": None,
            "avg_launch_us": dom["avg_launch_us"],
            "algorithmic_bytes_per_launch": dom["bytes_per_launch"],
        }
        tr_path = os.path.join(ROOT, "profiles", f"traffic_{dom_name}.json")
        if os.path.exists(tr_path):
            with open(tr_path) as f:
                tr = json.load(f)
            key = "mixed" if rmx is not None else "fp64"
            if key in tr:
                roofline["traffic"] = tr[key].get("hbm_bytes_per_launch")

    acc = run_accuracy(n, args.accuracy_steps, args.accuracy_cfl, args.tol) if args.accuracy_steps > 0 else None
    lid = run_lid3d(args.lid3d_n, args.lid3d_steps, 1e-6) if args.lid3d_steps > 0 else None
    stream = stream_reference(setup.shape)
    cpu = None if args.no_cpu else cpu_baseline(n, steps=args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils+Krylov / f64 defect correction" if rmx is not None else "f64",
        "data": "synthetic (uniform inflow + seeded inlet noise, synthetic blade polar)",
        "config": {
            "workload": f"windtunnel {n[0]}x{n[1]}x{n[2]} cells + 1 ALM rotor (3x32 points), dx=D/32, "
                        f"Re 2.5e5, Cs 0.1, dt 0.3dx/U, BiCGStab+RB-SOR({setup.struct.solver.pc_maxit} pass, "
                        f"w={setup.struct.solver.pc_relax:g}) tol {args.tol:g} on |r|/N",
            "cells": setup.n_interior,
            "precision": "mixed" if rmx is not None else "fp64",
        },
        "steps_per_s": head["steps_per_s"],
        "div_condition": {"bound": 1e-6, "monitor": "sqrt(sum div^2)/N after the last timed step",
                          "met": bool(divs) and max(divs) < 1e-6, "poisson_tol": args.tol},
        "mixed": rmx,
        "fp64": r64,
        "speedup_mixed_vs_fp64": (r64["ms_per_step"] / rmx["ms_per_step"]) if (r64 and rmx) else None,
        "velocity_rel_l2_mixed_vs_fp64": rel_l2,
        "solver_tol_leg": tol_leg,
        "roofline": roofline,
        "stream_reference": stream,
        "accuracy": acc,
        "lid3d_mixed_vs_fp64": lid,
        "cpu_baseline": cpu,
    }
    if calib:
        out["calibration"] = calib
    print(json.dumps(out))


if __name__ == "__main__":
    main()
