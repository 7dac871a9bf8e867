"""ctypes mirror of the structs and constants in include/falm_amd.h.

The layouts here must match the C header field for field; tests check the
sizes against the library's own view (falm_abi_sizes).
"""
import ctypes as C

FALM_OK = 0
FALM_ERR_MALLOC = 1
FALM_ERR_SETUP = 2
FALM_ERR_INVALID = 10
FALM_ERR_ALIAS = 11
FALM_ERR_UNSUPPORTED = 12
FALM_ERR_NONFINITE = 13
FALM_ERR_NOT_CONVERGED = 14
FALM_ERR_BREAKDOWN = 15
FALM_ERR_NO_DEVICE = 16
FALM_ERR_HIP = 100

F64, F32 = 0, 1

ADV_UPWIND1, ADV_UPWIND3, ADV_QUICK, ADV_UTOPIA, ADV_KK = 0, 1, 2, 3, 4
SGS_EMPTY, SGS_SMAGORINSKY, SGS_CSM = 0, 1, 2
LS_EMPTY, LS_JACOBI, LS_SOR, LS_PBICGSTAB = 0, 1, 2, 4
BLACK, RED = 0, 1

RED_COL_MAX, RED_COL_MIN, RED_COL_ABSMAX, RED_COL_ABSMIN, RED_VEC_MAX, RED_VEC_MIN = range(6)

SOLVE_F64, SOLVE_F32, SOLVE_MIXED = 0, 1, 2

BC_NONE, BC_INLET, BC_CONVECTIVE, BC_SLIP, BC_WALL0, BC_REFLECT = range(6)
PBC_NONE, PBC_NEUMANN, PBC_ZERO = 0, 1, 2

POISSON_NEUMANN, POISSON_OUTLET_XPLUS, POISSON_2D = 0, 1, 2

ALM_FIX_DRAG_TANGENT = 1
ALM_CUTOFF_ROUNDING = 6.1  # FALM_ALM_CUTOFF_ROUNDING

CASE_WINDTUNNEL, CASE_LID3D, CASE_LID2D = 0, 1, 2
CASE_F64, CASE_MIXED = 0, 1

FIELD_U, FIELD_UU, FIELD_P, FIELD_NUT, FIELD_FF, FIELD_DIV, FIELD_RHS, FIELD_POI_A, \
    FIELD_XYZ, FIELD_KX, FIELD_G, FIELD_JA = range(12)
FIELD_NCOL = {FIELD_U: 3, FIELD_UU: 3, FIELD_P: 1, FIELD_NUT: 1, FIELD_FF: 3, FIELD_DIV: 1,
              FIELD_RHS: 1, FIELD_POI_A: 7, FIELD_XYZ: 3, FIELD_KX: 3, FIELD_G: 3, FIELD_JA: 1}

TIME_NONE, TIME_SOR, TIME_MV, TIME_PSEUDO_U, TIME_ALM, TIME_SOLVER_UPDATE, TIME_PROJECT, TIME_FACE_FLUX, \
    TIME_SOLVER_INIT, TIME_SGS, TIME_RHS, TIME_DIVERGENCE, TIME_SOR_S, TIME_MV2, TIME_PCMV, TIME_PCMV_S, \
    TIME_PC_APPLY = range(17)


class Domain(C.Structure):
    _fields_ = [("shape", C.c_int * 3), ("offset", C.c_int * 3), ("gshape", C.c_int * 3),
                ("gc", C.c_int)]

    @classmethod
    def make(cls, shape, gc=2, offset=(0, 0, 0), gshape=None):
        d = cls()
        for m in range(3):
            d.shape[m] = int(shape[m])
            d.offset[m] = int(offset[m])
            d.gshape[m] = int((gshape or shape)[m])
        d.gc = int(gc)
        return d

    @property
    def ncell(self):
        return self.shape[0] * self.shape[1] * self.shape[2]


class FaceBC(C.Structure):
    _fields_ = [("type", C.c_int), ("value", C.c_double * 3)]


class InletNoise(C.Structure):
    _fields_ = [("intensity", C.c_double), ("seed", C.c_uint64), ("step", C.c_int64)]


class SolverConfig(C.Structure):
    _fields_ = [("type", C.c_int), ("maxit", C.c_int), ("tol", C.c_double), ("relax", C.c_double),
                ("pc_type", C.c_int), ("pc_maxit", C.c_int), ("pc_relax", C.c_double),
                ("precision", C.c_int), ("inner_maxit", C.c_int), ("inner_rtol", C.c_double),
                ("check_every", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("it", C.c_int), ("err", C.c_double), ("outer_it", C.c_int), ("status", C.c_int)]


class Turbine(C.Structure):
    _fields_ = [("base", C.c_double * 3), ("hub", C.c_double * 3), ("tip_rate", C.c_double),
                ("angle_type", C.c_int), ("motion", C.c_double * 3), ("surge", C.c_double * 3),
                ("pitch_deg", C.c_double * 8), ("yaw_deg", C.c_double), ("angle_offset_deg", C.c_double)]


DP = C.POINTER(C.c_double)


class AlmConfig(C.Structure):
    _fields_ = [("n_turbine", C.c_int), ("n_blade", C.c_int), ("n_ap_per_blade", C.c_int),
                ("n_attack", C.c_int), ("radius", C.c_double), ("hub_radius", C.c_double),
                ("eps", C.c_double), ("turbines", C.POINTER(Turbine)), ("ap_r", DP),
                ("ap_chord", DP), ("ap_twist", DP), ("attack", DP), ("cl", DP), ("cd", DP),
                ("twist_offset", DP), ("cutoff", C.c_double), ("flags", C.c_int)]


class CaseConfig(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int * 3), ("h", C.c_double * 3),
                ("origin", C.c_double * 3), ("Re", C.c_double), ("dt", C.c_double),
                ("Cs", C.c_double), ("adv_scheme", C.c_int), ("sgs_model", C.c_int),
                ("u_inflow", C.c_double), ("lid_velocity", C.c_double),
                ("inlet_intensity", C.c_double), ("seed", C.c_uint64), ("zminus_slip", C.c_int),
                ("precision", C.c_int), ("solver", SolverConfig),
                ("alm", C.POINTER(AlmConfig)), ("cv", C.POINTER(C.c_double) * 6)]


class StepInfo(C.Structure):
    _fields_ = [("step", C.c_int64), ("t", C.c_double), ("div", C.c_double),
                ("solver_it", C.c_int), ("solver_err", C.c_double), ("outer_it", C.c_int),
                ("status", C.c_int)]


ABI_STRUCTS = {"falm_domain": Domain, "falm_face_bc": FaceBC, "falm_inlet_noise": InletNoise,
               "falm_solver_config": SolverConfig, "falm_solve_info": SolveInfo,
               "falm_turbine": Turbine, "falm_alm_config": AlmConfig,
               "falm_case_config": CaseConfig, "falm_step_info": StepInfo}


def solver_config(type=LS_PBICGSTAB, maxit=1000, tol=1e-6, relax=1.0, pc_type=LS_SOR,
                  pc_maxit=2, pc_relax=1.5, precision=SOLVE_F64, inner_maxit=0,
                  inner_rtol=0.0, check_every=0):
    s = SolverConfig()
    s.type, s.maxit, s.tol, s.relax = type, maxit, tol, relax
    s.pc_type, s.pc_maxit, s.pc_relax = pc_type, pc_maxit, pc_relax
    s.precision, s.inner_maxit, s.inner_rtol, s.check_every = precision, inner_maxit, inner_rtol, check_every
    return s
