"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it.  Arrays use the reference's [site][19] layout, site = x + nx*y.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

OPS = {"P1": 1, "P3": 3, "F1": 11, "F2": 12, "F3": 13}
Q = 19


class Case(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("op", C.c_int), ("chi", C.c_double),
                ("xi", C.c_double), ("tau", C.c_double), ("A", C.c_double), ("beta", C.c_double)]


class DropDiag(C.Structure):
    _fields_ = ([(n, C.c_double) for n in
                 "raio raio_grad raio_linha p_in p_out delta_p sigma u_max u_rms massa_R massa_B qx qy qz "
                 "sd_in sd_out sigma_linha K_perfil r_perfil esp_perfil res_perfil".split()]
                + [(n, C.c_int) for n in "n_perfil n_in n_out".split()]
                + [(n, C.c_double) for n in
                   "x0 y0 area p_centro p_canto mean_x mean_y mean_z drop_x drop_y max_ref rms_ref max_mean "
                   "rms_mean max_lab rms_lab min_f min_rho min_phi max_phi axis_ratio center_dx center_dy "
                   "x_unwrapped".split()]
                + [("negative_total", C.c_longlong), ("negative_color", C.c_longlong)]
                + [(n, C.c_double) for n in "R0 a2 a4 a8 rmin rmax".split()]
                + [("ok", C.c_int), ("n_ruim", C.c_int)])


class InclDiag(C.Structure):
    _fields_ = ([(n, C.c_double) for n in
                 "sigma_pop int_grad K res d0 u_max u_rms un_max ut_max massa_R massa_B qx qy SFx SFy min_f".split()]
                + [("neg_total", C.c_longlong), ("neg_cor", C.c_longlong)]
                + [(n, C.c_int) for n in "n_ruim n_ajuste ok".split()])


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_step.argtypes = [C.POINTER(Case), dp, dp, dp, dp, C.c_void_p, C.c_void_p]
        L.orc_step_geo.argtypes = [C.POINTER(Case), np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS"),
                                   C.c_double, dp, dp, dp, dp, C.c_void_p, C.c_void_p]
        L.orc_init_drop.argtypes = [C.POINTER(Case), C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_init_band.argtypes = [C.POINTER(Case), C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_force_field.argtypes = [C.POINTER(Case), dp, dp, dp]
        L.orc_perturbation.argtypes = [dp, C.c_double, dp, C.c_int, C.c_double]
        L.orc_source.argtypes = [C.c_double] * 8 + [dp]
        L.orc_drop_diag.argtypes = [C.POINTER(Case), dp, dp, C.c_void_p, C.c_double, C.c_int, C.c_double,
                                    C.c_double, C.POINTER(DropDiag)]
        L.orc_incl_diag.argtypes = [C.POINTER(Case), C.c_int, C.c_int, dp, dp, C.c_void_p, C.c_void_p,
                                    C.POINTER(InclDiag)]
        L.orc_e0.argtypes = [dp, dp, dp]
        L.orc_e0.restype = C.c_int
        _lib = L
    return _lib


def parse_op(text):
    """operador.txt line -> (op id, chi, xi) with the reference defaults (Operadores_tensao.cpp:51-86)."""
    t = text.split()
    op, chi, xi = OPS[t[0]], 0.5, 1.569988
    if t[0] == "P3" and len(t) > 1:
        chi = float(t[1])
    if t[0] == "F3" and len(t) > 1:
        xi = float(t[1])
    return op, chi, xi


def make_case(nx, ny, op="P3 0.5", tau=1.0, A=0.225, beta=0.8):
    o, chi, xi = parse_op(op)
    return Case(nx, ny, o, chi, xi, tau, A, beta)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Sim:
    """Oracle simulation holding the reference's state arrays (pre-step, post-stream)."""

    def __init__(self, case, solid=None, wett=0.0):
        """solid: optional nx*ny mask (1 = solid site, site = x + nx*y) and the wall's wettability."""
        self.cs = case
        self.solid = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8)
        self.wett = float(wett)
        n = case.nx * case.ny
        self.n = n
        self.fR = np.zeros(n * Q)
        self.fB = np.zeros(n * Q)
        self.fR_new = np.zeros(n * Q)
        self.fB_new = np.zeros(n * Q)
        self.F = np.zeros(3 * n) if case.op >= 11 else None
        self.gm = np.zeros(3 * n)
        self.post_R = None
        self.post_B = None

    def init_drop(self, R, width=3.45, v0=0.0, x0=None, y0=None):
        k = 2.0 * np.log(9.0) / width if width > 0 else -1.0
        x0 = self.cs.nx // 2 if x0 is None else x0
        y0 = self.cs.ny // 2 if y0 is None else y0
        lib().orc_init_drop(C.byref(self.cs), R, k, x0, y0, v0, self.fR, self.fB)

    def init_band(self, p, q, width=3.45):
        lib().orc_init_band(C.byref(self.cs), p, q, width, self.fR, self.fB)

    def step(self):
        """One step; afterwards post_R/post_B hold post-collision populations, fR/fB the next state."""
        if self.solid is None:
            lib().orc_step(C.byref(self.cs), self.fR, self.fB, self.fR_new, self.fB_new, _ptr(self.F), _ptr(self.gm))
        else:
            lib().orc_step_geo(C.byref(self.cs), self.solid, self.wett, self.fR, self.fB, self.fR_new, self.fB_new,
                               _ptr(self.F), _ptr(self.gm))
        self.post_R, self.post_B = self.fR, self.fB
        self.fR, self.fR_new = self.fR_new, self.fR
        self.fB, self.fB_new = self.fB_new, self.fB

    def drop_diag(self, v0=0.0, step=0):
        d = DropDiag()
        lib().orc_drop_diag(C.byref(self.cs), self.post_R, self.post_B, _ptr(self.F), v0, step,
                            float(self.cs.nx // 2), float(self.cs.ny // 2), C.byref(d))
        return d

    def incl_diag(self, p, q):
        d = InclDiag()
        lib().orc_incl_diag(C.byref(self.cs), p, q, self.post_R, self.post_B, _ptr(self.F), _ptr(self.gm),
                            C.byref(d))
        return d


def e0():
    err = np.zeros(64)
    tol = np.zeros(64)
    m4 = np.zeros(20)
    fails = lib().orc_e0(err, tol, m4)
    return fails, err[:37], tol[:37], m4
