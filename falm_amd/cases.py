"""Case set-ups (host side): the configurations the step driver runs.

Each builder returns a CaseSetup whose `.struct` is the falm_case_config
passed to falm_case_create (and to the CPU oracle's orc_case_create).  The
numpy arrays referenced by the ALM config are kept alive by the setup.

Workloads (interior cells):
  lid2d      128x128x1  2-D lid-driven cavity, Re 1000, RB-SOR (omega 1.2, tol 1e-8)
  lid3d      128^3      3-D lid-driven cavity LES, Re 1e4, Cs 0.1, BiCGStab+RB-SOR
  windtunnel 512x128x128 channel with one actuator-line rotor (D = 1, dx = D/32)
  farm_segment 1024x256x256 channel with two rotors 7D apart, the upstream one
             floating (surge + pitch), 8 % inlet turbulence
"""
import ctypes as C
import math

import numpy as np

from . import _abi as abi


class AlmSetup:
    """Actuator-line configuration plus the arrays it points to."""

    def __init__(self, turbines, n_blade, n_ap_per_blade, radius, hub_radius, eps, ap_r,
                 ap_chord, ap_twist, attack, cl, cd, twist_offset=None, cutoff=0.0, flags=0):
        self.n_turbine = len(turbines)
        self.n_blade = n_blade
        self.n_ap_per_blade = n_ap_per_blade
        self.n_ap = self.n_turbine * n_blade * n_ap_per_blade
        self.turbines = (abi.Turbine * self.n_turbine)(*turbines)
        f64 = lambda v: np.ascontiguousarray(v, dtype=np.float64)  # noqa: E731
        self.ap_r, self.ap_chord, self.ap_twist = f64(ap_r), f64(ap_chord), f64(ap_twist)
        self.attack, self.cl, self.cd = f64(attack), f64(cl), f64(cd)
        self.twist_offset = None if twist_offset is None else f64(twist_offset)
        assert self.ap_r.size == self.n_ap and self.cl.size == self.n_ap * self.attack.size
        s = abi.AlmConfig()
        s.n_turbine, s.n_blade, s.n_ap_per_blade = self.n_turbine, n_blade, n_ap_per_blade
        s.n_attack = self.attack.size
        s.radius, s.hub_radius, s.eps = radius, hub_radius, eps
        s.turbines = C.cast(self.turbines, C.POINTER(abi.Turbine))
        dp = lambda a: None if a is None else a.ctypes.data_as(abi.DP)  # noqa: E731
        s.ap_r, s.ap_chord, s.ap_twist = dp(self.ap_r), dp(self.ap_chord), dp(self.ap_twist)
        s.attack, s.cl, s.cd = dp(self.attack), dp(self.cl), dp(self.cd)
        s.twist_offset = dp(self.twist_offset)
        s.cutoff, s.flags = cutoff, flags
        self.struct = s


class CaseSetup:
    def __init__(self, struct, alm=None, name=""):
        self.struct = struct
        self.alm = alm
        self.name = name
        if alm is not None:
            self.struct.alm = C.pointer(alm.struct)

    @property
    def n(self):
        return tuple(self.struct.n)

    @property
    def n_interior(self):
        n = self.struct.n
        return n[0] * n[1] * n[2]

    @property
    def shape(self):
        n = self.struct.n
        return (n[0] + 4, n[1] + 4, n[2] + 4)


def make_turbine(base, hub=(0.0, 0.0, 0.0), tip_rate=0.0, angle_type=0, motion=(0.0, 0.0, 0.0),
                 surge=(0.0, 0.0, 0.0), pitch_deg=(), yaw_deg=0.0, angle_offset_deg=0.0):
    t = abi.Turbine()
    for m in range(3):
        t.base[m], t.hub[m] = base[m], hub[m]
        t.motion[m], t.surge[m] = motion[m], surge[m]
    t.tip_rate = tip_rate
    t.angle_type = angle_type
    t.yaw_deg, t.angle_offset_deg = yaw_deg, angle_offset_deg
    for b, v in enumerate(pitch_deg):
        t.pitch_deg[b] = v
    return t


def synthetic_polar(alpha_deg):
    """Cl = 2*pi*alpha clipped at |alpha| <= 12 deg, Cd = 0.01 + 0.05*alpha^2
    (alpha in radians), tabulated at the given attack angles."""
    a = np.radians(np.asarray(alpha_deg, dtype=np.float64))
    cl = 2.0 * math.pi * np.clip(a, -math.radians(12.0), math.radians(12.0))
    cd = 0.01 + 0.05 * a * a
    return cl, cd


def blade_table(radius, chord_root, chord_tip, twist_root, twist_tip, n_station=17,
                attack=np.arange(-45.0, 46.0, 1.0)):
    """Blade stations in absolute radius (the reference's buildAP reads the
    r/R column as an absolute radius, src/alm/bladeHandler.h:31,57-83)."""
    rr = np.linspace(0.0, radius, n_station)
    s = rr / radius
    chord = chord_root + (chord_tip - chord_root) * s
    twist = twist_root + (twist_tip - twist_root) * s
    cl1, cd1 = synthetic_polar(attack)
    cl = np.tile(cl1, (n_station, 1))
    cd = np.tile(cd1, (n_station, 1))
    return rr, chord, twist, np.asarray(attack, dtype=np.float64), cl, cd


def build_ap_host(rr, chord, twist, cl, cd, n_turbine, n_blade, n_ap_per_blade, radius, hub_radius):
    """BladeHandler::buildAP (src/alm/bladeHandler.h:16-104) in numpy:
    per-AP radius, chord, twist and the Cl/Cd columns [attack][ap]."""
    nst, nat = cl.shape
    nap = n_turbine * n_blade * n_ap_per_blade
    dr = (radius - hub_radius) / n_ap_per_blade
    ap_r = np.zeros(nap)
    ap_chord = np.zeros(nap)
    ap_twist = np.zeros(nap)
    ap_cl = np.zeros((nat, nap))
    ap_cd = np.zeros((nat, nap))
    for ap in range(nap):
        r = (ap % n_ap_per_blade + 0.5) * dr + hub_radius
        ap_r[ap] = r
        if r < rr[0]:
            i, p = 0, 0.0
        elif r >= rr[-1]:
            i, p = nst - 1, 0.0
        else:
            i = int(np.nonzero((rr[:-1] <= r) & (rr[1:] > r))[0][0])
            p = (r - rr[i]) / (rr[i + 1] - rr[i])
        j = min(i + 1, nst - 1)
        ap_chord[ap] = (1.0 - p) * chord[i] + p * chord[j]
        ap_twist[ap] = (1.0 - p) * twist[i] + p * twist[j]
        ap_cl[:, ap] = (1.0 - p) * cl[i] + p * cl[j]
        ap_cd[:, ap] = (1.0 - p) * cd[i] + p * cd[j]
    return ap_r, ap_chord, ap_twist, ap_cl.ravel(), ap_cd.ravel()


def rotor(h, base, radius=0.5, tsr=7.0, u_inf=1.0, n_blade=3, n_ap_per_blade=32, eps_cells=2.0,
          chord=(0.08, 0.02), twist=(20.0, 0.0), hub_radius=0.0, cutoff=abi.ALM_CUTOFF_ROUNDING, flags=0,
          angle_type=0, motion=(0.0, 0.0, 0.0), surge=(0.0, 0.0, 0.0), pitch_deg=(),
          twist_offset=None, extra_bases=(), yaw_deg=0.0, angle_offset_deg=0.0):
    """One (or more, via extra_bases) three-bladed rotor of diameter 2*radius:
    chord and twist linear in r, synthetic polar, Gaussian width eps_cells*h,
    tip rate Omega = tsr*u_inf/radius.  cutoff: culling radius in units of
    eps (default: terms below 2^-53 of a point's peak are dropped; 0 keeps
    every term, the reference's all-pairs sum).  The platform motion
    (angle_type, motion, angle_offset_deg, surge) and the yaw belong to the
    first rotor; the rotors at extra_bases stand still."""
    rr, ch, tw, attack, cl, cd = blade_table(radius, chord[0], chord[1], twist[0], twist[1])
    bases = [tuple(base)] + [tuple(b) for b in extra_bases]
    zero = (0.0, 0.0, 0.0)
    turbines = [make_turbine(b, tip_rate=tsr * u_inf / radius, angle_type=angle_type if i == 0 else 0,
                             motion=motion if i == 0 else zero, surge=surge if i == 0 else zero,
                             pitch_deg=pitch_deg, yaw_deg=yaw_deg if i == 0 else 0.0,
                             angle_offset_deg=angle_offset_deg if i == 0 else 0.0)
                for i, b in enumerate(bases)]
    ap = build_ap_host(rr, ch, tw, cl, cd, len(turbines), n_blade, n_ap_per_blade, radius, hub_radius)
    return AlmSetup(turbines, n_blade, n_ap_per_blade, radius, hub_radius, eps_cells * h,
                    ap[0], ap[1], ap[2], attack, ap[3], ap[4], twist_offset=twist_offset,
                    cutoff=cutoff, flags=flags)


def _base_struct(kind, n, h, origin=(0.0, 0.0, 0.0)):
    s = abi.CaseConfig()
    s.kind = kind
    for m in range(3):
        s.n[m], s.h[m], s.origin[m] = int(n[m]), float(h[m]), float(origin[m])
    return s


def windtunnel(n=(512, 128, 128), cells_per_d=32, re=2.5e5, cfl=0.3, u_inf=1.0, cs=0.1,
               inlet_intensity=0.05, seed=20241202, precision=abi.CASE_F64, with_rotor=True,
               hub_x_d=4.0, pc_maxit=1, pc_relax=1.0, tol=1e-6, maxit=1000, rotor_kw=None,
               solver_kw=None):
    """Wind-tunnel channel (app/windtunnel, app/nedo) with one actuator-line
    rotor: dx = D/cells_per_d, inflow u_inf + Gaussian inlet noise, convective
    outlet, slip side walls (z- slip too), fixed dt = cfl*dx/u_inf.
    Preconditioner: one red-black Gauss-Seidel pass (pc_maxit 1, omega 1);
    with the pressure warm-started, one BiCGStab iteration per step meets
    tol 1e-6 whether the preconditioner makes 1, 2 or 3 passes, so the
    lightest one is the fastest."""
    h = 1.0 / cells_per_d
    s = _base_struct(abi.CASE_WINDTUNNEL, n, (h, h, h))
    s.Re, s.dt, s.Cs = re, cfl * h / u_inf, cs
    s.adv_scheme, s.sgs_model = abi.ADV_UPWIND3, abi.SGS_SMAGORINSKY
    s.u_inflow, s.lid_velocity = u_inf, 0.0
    s.inlet_intensity, s.seed, s.zminus_slip = inlet_intensity * u_inf, seed, 1
    s.precision = precision
    skw = dict(type=abi.LS_PBICGSTAB, maxit=maxit, tol=tol, pc_type=abi.LS_SOR, pc_maxit=pc_maxit,
               pc_relax=pc_relax,
               precision=abi.SOLVE_MIXED if precision == abi.CASE_MIXED else abi.SOLVE_F64)
    skw.update(solver_kw or {})
    s.solver = abi.solver_config(**skw)
    alm = None
    if with_rotor:
        ly, lz = n[1] * h, n[2] * h
        kw = dict(rotor_kw or {})
        alm = rotor(h, (hub_x_d, 0.5 * ly, 0.5 * lz), u_inf=u_inf, **kw)
    return CaseSetup(s, alm, name=f"windtunnel {n[0]}x{n[1]}x{n[2]}" + (" + ALM" if alm else ""))


def farm_segment(n=(1024, 256, 256), cells_per_d=32, spacing_d=7.0, hub_x_d=4.0, re=2.5e5, cfl=0.1, u_inf=1.0,
                 cs=0.1, inlet_intensity=0.08, seed=20241203, precision=abi.CASE_MIXED, tol=1e-6, maxit=1000,
                 surge_d=0.1, pitch_amp_deg=3.0, period_revs=20.0, yaw_deg=0.0, angle_offset_deg=0.0,
                 pc_maxit=1, pc_relax=1.0, rotor_kw=None):
    """Floating tandem-rotor segment: two rotors spacing_d diameters apart in
    line on the channel's centre line, the upstream one on a platform in
    prescribed surge surge_d*D*sin(2 pi t/T) and pitch
    pitch_amp_deg*sin(2 pi t/T + pi/2), T = period_revs rotor revolutions,
    with 8 % inlet turbulence.  CFL 0.1: the explicit Upwind3 step amplifies
    the longest resolved waves by 2.4 % per step at CFL 0.3, which a
    2000-step run does not survive."""
    h = 1.0 / cells_per_d
    radius, tsr = 0.5, 7.0
    omega = tsr * u_inf / radius
    w = omega / period_revs  # 2 pi / T with T = period_revs * 2 pi / Omega
    ly, lz = n[1] * h, n[2] * h
    kw = dict(rotor_kw or {})
    kw.setdefault("angle_type", 2)
    kw.setdefault("motion", (math.radians(pitch_amp_deg), w, 0.5 * math.pi))
    kw.setdefault("surge", (surge_d * 2 * radius, w, 0.0))
    setup = windtunnel(n=n, cells_per_d=cells_per_d, re=re, cfl=cfl, u_inf=u_inf, cs=cs,
                       inlet_intensity=inlet_intensity, seed=seed, precision=precision, with_rotor=False,
                       tol=tol, maxit=maxit, pc_maxit=pc_maxit, pc_relax=pc_relax)
    alm = rotor(h, (hub_x_d, 0.5 * ly, 0.5 * lz), radius=radius, tsr=tsr, u_inf=u_inf,
                extra_bases=[(hub_x_d + spacing_d * 2 * radius, 0.5 * ly, 0.5 * lz)], yaw_deg=yaw_deg,
                angle_offset_deg=angle_offset_deg, **kw)
    return CaseSetup(setup.struct, alm, name=f"farm segment {n[0]}x{n[1]}x{n[2]} + 2 rotors {spacing_d:g}D apart")


def lid3d(n=128, re=1e4, dt=5e-4, cs=0.1, lid=1.0, precision=abi.CASE_F64, tol=1e-6, maxit=1000,
          pc_maxit=2, pc_relax=1.5, solver_kw=None):
    """3-D lid-driven cavity LES on the unit cube (app/lid3d): no-slip walls,
    lid on y+ moving in +x, Smagorinsky, BiCGStab + RB-SOR.  n may also be a
    triple of interior cell counts: a box of cubic cells of width 1/n[0]."""
    nn = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
    h = 1.0 / nn[0]
    s = _base_struct(abi.CASE_LID3D, nn, (h, h, h))
    s.Re, s.dt, s.Cs = re, dt, cs
    s.adv_scheme, s.sgs_model = abi.ADV_UPWIND3, abi.SGS_SMAGORINSKY
    s.u_inflow, s.lid_velocity = 0.0, lid
    s.precision = precision
    skw = dict(type=abi.LS_PBICGSTAB, maxit=maxit, tol=tol, pc_type=abi.LS_SOR, pc_maxit=pc_maxit,
               pc_relax=pc_relax,
               precision=abi.SOLVE_MIXED if precision == abi.CASE_MIXED else abi.SOLVE_F64)
    skw.update(solver_kw or {})
    s.solver = abi.solver_config(**skw)
    return CaseSetup(s, None, name=f"lid3d {nn[0]}x{nn[1]}x{nn[2]}")


def lid2d(n=128, re=1000.0, dt=1e-3, lid=1.0, omega=1.2, tol=1e-8, maxit=100000):
    """2-D lid-driven cavity (app/ld2d) as a z-thin domain: plain method,
    no SGS, red-black SOR pressure solver."""
    h = 1.0 / n
    s = _base_struct(abi.CASE_LID2D, (n, n, 1), (h, h, h))
    s.Re, s.dt, s.Cs = re, dt, 0.0
    s.adv_scheme, s.sgs_model = abi.ADV_UPWIND3, abi.SGS_EMPTY
    s.u_inflow, s.lid_velocity = 0.0, lid
    s.precision = abi.CASE_F64
    s.solver = abi.solver_config(type=abi.LS_SOR, maxit=maxit, tol=tol, relax=omega,
                                 pc_type=abi.LS_EMPTY)
    return CaseSetup(s, None, name=f"lid2d {n}x{n}")
