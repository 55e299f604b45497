"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker of the HIP path.

Arrays are exchanged by the reference's names (parameters.f90:222-292) in the
reference's layout: Fortran column-major inner dimensions, one row per cell,
face or boundary face, 0-based indices with -1 for "no neighbour".
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

NATURAL = 0
COLORED = 1


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "nu", "Re", "alpha", "rho", "g_beta", "DT", "w_u", "w_v", "w_w", "w_p", "w_p2", "w_T",
        "v_limit", "R_i", "R_o")]


def reference_params(Ra=1.0e5, Pr=0.7, DT=0.01, R_i=None, R_o=2.0, rho=1.0, g_beta=1.0, th_tc=1.0,
                     relax=0.4, v_limit=1e-2):
    """Physical constants derived as parameters.f90:152-173 derives them."""
    if R_i is None:
        R_i = 1.0 * (3.14159265359 / 3.0)
    nu = 1. / np.sqrt(Ra / (g_beta * th_tc * Pr))
    p = Params()
    p.nu = nu
    p.Re = 1. / nu
    p.alpha = nu / Pr
    p.rho = rho
    p.g_beta = g_beta
    p.DT = DT
    for n in ("w_u", "w_v", "w_w", "w_p", "w_p2", "w_T"):
        setattr(p, n, relax)
    p.v_limit = v_limit
    p.R_i = R_i
    p.R_o = R_o
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.ora_create.restype = vp
        L.ora_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Params)]
        L.ora_destroy.argtypes = [vp]
        L.ora_field_info.restype = ctypes.c_long
        L.ora_field_info.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.ora_field_ptr.restype = vp
        L.ora_field_ptr.argtypes = [vp, ctypes.c_char_p]
        for n in ("ora_velocity_constants", "ora_temperature_constants", "ora_pressure_constants",
                  "ora_velocity_sources", "ora_pressure_sources", "ora_temperature_sources", "ora_pressure_bc",
                  "ora_zero_flux", "ora_pressure_sweep", "ora_check_steady_state", "ora_evaluate_D", "ora_update_time_step"):
            getattr(L, n).argtypes = [vp]
        for n in ("ora_velocity_sweep", "ora_temperature_sweep", "ora_pressure_stage"):
            getattr(L, n).argtypes = [vp, ctypes.c_int]
        L.ora_find_error.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.ora_set_p_measure.restype = None
        L.ora_set_p_measure.argtypes = [vp, ctypes.c_int]
        L.ora_iteration.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.ora_inner_solve.restype = ctypes.c_int
        L.ora_inner_solve.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
        L.ora_nusselt.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.ora_get_scalars.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.ora_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_int]
        L.ora_time_step.restype = ctypes.c_int
        L.ora_time_step.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_ic_noise.restype = ctypes.c_double
        L.oracle_ic_noise.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int64]
        L.ora_threads.restype = ctypes.c_int
        L.ora_threads.argtypes = []
        L.ora_set_threads.restype = None
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_nim_direction.restype = None
        L.ora_nim_direction.argtypes = [ctypes.c_double] * 5 + [ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One oracle context on an nt x nr x nz O-grid."""

    def __init__(self, nt, nr, nz, params):
        self.L = lib()
        self.params = params
        self.h = self.L.ora_create(nt, nr, nz, ctypes.byref(params))
        if not self.h:
            raise MemoryError("ora_create failed")
        self.nt, self.nr, self.nz = nt, nr, nz

    def close(self):
        if self.h:
            self.L.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _view(self, name):
        isint = ctypes.c_int()
        n = self.L.ora_field_info(self.h, name.encode(), ctypes.byref(isint))
        if n < 0:
            raise KeyError(name)
        ptr = self.L.ora_field_ptr(self.h, name.encode())
        ct = ctypes.c_int32 if isint.value else ctypes.c_double
        buf = (ct * n).from_address(ptr)
        return np.ctypeslib.as_array(buf)

    def set(self, name, values):
        v = self._view(name)
        v[:] = np.asarray(values).reshape(-1)

    def get(self, name):
        return self._view(name).copy()

    # stages
    def velocity_constants(self): self.L.ora_velocity_constants(self.h)
    def temperature_constants(self): self.L.ora_temperature_constants(self.h)
    def pressure_constants(self): self.L.ora_pressure_constants(self.h)
    def velocity_sources(self): self.L.ora_velocity_sources(self.h)
    def pressure_sources(self): self.L.ora_pressure_sources(self.h)
    def temperature_sources(self): self.L.ora_temperature_sources(self.h)
    def pressure_bc(self): self.L.ora_pressure_bc(self.h)
    def zero_flux(self): self.L.ora_zero_flux(self.h)
    def pressure_sweep(self): self.L.ora_pressure_sweep(self.h)
    def velocity_sweep(self, order=NATURAL): self.L.ora_velocity_sweep(self.h, order)
    def temperature_sweep(self, order=NATURAL): self.L.ora_temperature_sweep(self.h, order)
    def pressure_stage(self, order=NATURAL): self.L.ora_pressure_stage(self.h, order)
    def check_steady_state(self): self.L.ora_check_steady_state(self.h)
    def evaluate_D(self): self.L.ora_evaluate_D(self.h)
    def update_time_step(self): self.L.ora_update_time_step(self.h)

    def set_p_measure(self, mode):
        """0: max |dp| as find_error; 1: (max dp - min dp) / 2, blind to a
        change of p's constant level."""
        self.L.ora_set_p_measure(self.h, int(mode))

    def find_error(self):
        e = np.zeros(5)
        self.L.ora_find_error(self.h, _dp(e))
        return e

    def iteration(self, order=NATURAL):
        e = np.zeros(5)
        self.L.ora_iteration(self.h, order, _dp(e))
        return e

    def inner_solve(self, tol, order=NATURAL, max_iter=100000, min_iter=0):
        t = np.ascontiguousarray(tol, dtype=np.float64)
        e = np.zeros(5)
        it = ctypes.c_int()
        rc = self.L.ora_inner_solve(self.h, order, _dp(t), max_iter, min_iter, ctypes.byref(it), _dp(e))
        return rc, it.value, e

    def time_step(self, tol, order=NATURAL, max_iter=100000, min_iter=0):
        t = np.ascontiguousarray(tol, dtype=np.float64)
        scal = np.zeros(9)
        nu = np.zeros(5)
        rc = self.L.ora_time_step(self.h, order, _dp(t), max_iter, min_iter, _dp(scal), _dp(nu))
        return rc, scal, nu

    def nusselt(self):
        o = np.zeros(5)
        self.L.ora_nusselt(self.h, _dp(o))
        return o

    def scalars(self):
        o = np.zeros(6)
        self.L.ora_get_scalars(self.h, _dp(o))
        return dict(Dilatation=o[0], dudt_max=o[1], dTdt_max=o[2], TI=o[3], STEP=int(o[4]))

    def set_time(self, TI, STEP):
        self.L.ora_set_time(self.h, TI, STEP)


def ic_noise(seed, field, idx):
    """Vectorised oracle_ic_noise over an index array."""
    L = lib()
    return np.array([L.oracle_ic_noise(seed, field, int(i)) for i in np.asarray(idx).reshape(-1)])


# reference arrays (golden fixture names) -> oracle field names
GEOMETRY_MAP = {
    "meshes_faces": "mfaces", "meshes_neighbors": "nbr", "faces_BC": "faces_bc", "bc_faces": "bc_faces",
    "bc_position": "bc_pos", "bc_mesh": "bc_mesh", "bc_points": "bc_points", "volume": "volume",
    "xi_x": "xi_x", "gij": "gij", "gij_faces": "gij_faces", "G1": "G1", "G2": "G2", "G3": "G3",
    "faces_area": "faces_area", "faces_normal": "faces_normal", "faces_mid": "faces_mid", "bc_a": "bc_a",
    "bc_ap": "bc_ap", "AA_px": "AA_px", "AA_py": "AA_py", "AA_pz": "AA_pz",
}
ONE_BASED = {"meshes_faces", "meshes_neighbors", "bc_faces", "bc_mesh", "bc_points"}


def load_reference_geometry(orc, geo):
    """Copy a reference-layout geometry (1-based integer tables as dumped
    from the compiled reference) into an oracle context."""
    for src, dst in GEOMETRY_MAP.items():
        if src not in geo:
            continue
        v = np.asarray(geo[src])
        if src in ONE_BASED:
            v = v.astype(np.int64) - 1
        orc.set(dst, v)


def threads():
    """Threads the oracle's OpenMP loops use."""
    return int(lib().ora_threads())


def set_threads(n):
    lib().ora_set_threads(int(n))


def nim_direction(u, k, gp, gm, v_limit=1e-2):
    """The six exponential-fit constants of one direction (nim.f90:184-287)."""
    import numpy as np
    out = np.zeros(6)
    lib().ora_nim_direction(u, k, gp, gm, v_limit, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out
