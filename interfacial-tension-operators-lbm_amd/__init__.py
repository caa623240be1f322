"""Python binding of the itolbm C ABI (include/itolbm.h) over ctypes.

The HIP library libitolbm.so must have been built (make -C interfacial-tension-operators-lbm_amd);
loading fails loudly otherwise: there is no CPU fallback.

Arrays exchanged with the library use the reference layout [site][19] (site = x + nx*y) for
populations and [site][3] for vectors, as float64.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libitolbm.so")

Q = 19
OPS = {"P1": 1, "P3": 3, "F1": 11, "F2": 12, "F3": 13}
PREC_F64 = 0
PREC_F32 = 1
PREC_F16 = 2
PREC_F32_ZSYM = 3   # FP32 storage of the 14 distinct populations of a z-mirror-symmetric state
PREC_F16_ZSYM = 4
STATE_POST_STREAM = 0
STATE_POST_COLLISION = 1
STEP_RECORD_AUX = 1
VTK_DENSITY = 0
VTK_VELOCITY = 1
VTK_BINARY = 1
VTK_DOUBLE = 2

# every symbol include/itolbm.h declares
SYMBOLS = ["lbm_create", "lbm_destroy", "lbm_last_error", "lbm_device_bytes", "lbm_upload_state",
           "lbm_download_state", "lbm_init_drop", "lbm_init_band", "lbm_step", "lbm_step_case",
           "lbm_step_case_noalloc", "lbm_tile_shape", "lbm_stream_copy", "lbm_sync", "lbm_set_stream", "lbm_set_reference", "lbm_set_active", "lbm_set_graph",
           "lbm_download_aux", "lbm_diag_drop", "lbm_diag_incl", "lbm_read_geometry", "lbm_set_geometry",
           "lbm_write_vtk", "lbm_step_mode", "lbm_stored_bytes"]


class CaseDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("op", C.c_int), ("chi", C.c_double), ("xi", C.c_double),
                ("tau", C.c_double), ("A", C.c_double), ("beta", C.c_double)]


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

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class InclDiag(C.Structure):
    _fields_ = ([(n, C.c_double) for n in
                 "sigma_pop int_grad K res d0 u_max u_rms un_max ut_max massa_R massa_B qx qy SFx SFy min_f".split()]
                + [("neg_total", C.c_longlong), ("neg_cor", C.c_longlong)]
                + [(n, C.c_int) for n in "n_ruim n_ajuste ok".split()])

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libitolbm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.lbm_create.argtypes = [C.POINTER(CaseDesc), C.c_int, C.c_int, C.POINTER(P)]
        L.lbm_destroy.argtypes = [P]
        L.lbm_last_error.argtypes = [P]
        L.lbm_last_error.restype = C.c_char_p
        L.lbm_device_bytes.argtypes = [P]
        L.lbm_device_bytes.restype = C.c_longlong
        L.lbm_upload_state.argtypes = [P, C.c_int, dp, dp]
        L.lbm_download_state.argtypes = [P, C.c_int, C.c_int, dp, dp]
        L.lbm_init_drop.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double]
        L.lbm_init_band.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double]
        L.lbm_step.argtypes = [P, C.c_int, C.c_int]
        L.lbm_step_case.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.lbm_step_case_noalloc.argtypes = [P, C.c_int, C.c_int, C.c_int]
        ip = C.POINTER(C.c_int)
        L.lbm_tile_shape.argtypes = [P, C.c_int, ip, ip, ip]
        L.lbm_stream_copy.argtypes = [P, P, C.c_ulonglong, P]
        L.lbm_sync.argtypes = [P]
        L.lbm_set_stream.argtypes = [P, C.c_void_p]
        L.lbm_set_reference.argtypes = [P, C.c_int, C.c_double, C.c_double]
        L.lbm_set_active.argtypes = [P, C.c_int, C.c_int]
        L.lbm_set_graph.argtypes = [P, C.c_int]
        L.lbm_download_aux.argtypes = [P, C.c_int, C.c_void_p, C.c_void_p]
        L.lbm_diag_drop.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.POINTER(DropDiag)]
        L.lbm_diag_incl.argtypes = [P, C.c_int, C.c_int, C.c_int, C.POINTER(InclDiag), C.c_void_p]
        L.lbm_step_mode.argtypes = [P, ip]
        L.lbm_stored_bytes.argtypes = [P, C.c_int, ip]
        L.lbm_read_geometry.argtypes = [C.c_char_p, C.c_int, C.c_int, ip, ip, ip, C.c_void_p, C.c_longlong]
        L.lbm_read_geometry.restype = C.c_longlong
        L.lbm_set_geometry.argtypes = [P, C.c_int, C.c_void_p, C.c_double]
        L.lbm_write_vtk.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_int]
        for s in SYMBOLS:
            if s not in ("lbm_last_error", "lbm_device_bytes", "lbm_read_geometry"):
                getattr(L, s).restype = C.c_int
        _lib = L
    return _lib


def parse_op(text):
    """operador.txt line -> (op id, chi, xi) with the reference defaults (Operadores_tensao.cpp:51-86)."""
    t = text.split()
    if not t or t[0] not in OPS:
        raise ValueError(f"bad operator line {text!r}: P1 | P3 chi | F1 | F2 | F3 xi")
    chi, xi = 0.5, 1.569988
    if t[0] == "P3" and len(t) > 1:
        chi = float(t[1])
    if t[0] == "F3" and len(t) > 1:
        xi = float(t[1])
    return OPS[t[0]], chi, xi


def case_desc(nx, ny, op="P3 0.5", tau=1.0, A=0.225, beta=0.8):
    o, chi, xi = parse_op(op)
    return CaseDesc(int(nx), int(ny), o, chi, xi, float(tau), float(A), float(beta))


class LBMError(RuntimeError):
    pass


def stream_copy(dst_ptr, src_ptr, nbytes, hip_stream=None):
    """lbm_stream_copy: the library's streaming copy kernel between two device addresses."""
    rc = lib().lbm_stream_copy(C.c_void_p(int(dst_ptr)), C.c_void_p(int(src_ptr)), int(nbytes),
                               None if hip_stream is None else C.c_void_p(int(hip_stream)))
    if rc != 0:
        raise LBMError(f"lbm_stream_copy failed ({rc})")


def read_geometry(path, pts_in=0, pts_out=0):
    """lbm_read_geometry (read_geo): returns (nx, ny, nz, meio), meio an int32 array of nx*ny*nz entries,
    0 for a solid site and 1, 2, ... for the fluid sites in site order."""
    nx, ny, nz = C.c_int(), C.c_int(), C.c_int()
    n = lib().lbm_read_geometry(os.fsencode(path), pts_in, pts_out, C.byref(nx), C.byref(ny), C.byref(nz), None, 0)
    if n < 0:
        raise LBMError(f"lbm_read_geometry({path!r}) failed ({n})")
    meio = np.zeros(nx.value * ny.value * nz.value, dtype=np.int32)
    n = lib().lbm_read_geometry(os.fsencode(path), pts_in, pts_out, None, None, None,
                                meio.ctypes.data_as(C.c_void_p), meio.size)
    if n < 0:
        raise LBMError(f"lbm_read_geometry({path!r}) failed ({n})")
    return nx.value, ny.value, nz.value, meio


def write_vtk(path, field, nx, ny, fR, fB=None, force=None, meio=None, binary=False, double=False):
    """lbm_write_vtk: rho (field VTK_DENSITY, of fR) or velocity (VTK_VELOCITY) file of host populations."""
    fR = np.ascontiguousarray(fR, dtype=np.float64).ravel()
    keep = [np.ascontiguousarray(a, dtype=dt).ravel() if a is not None else None
            for a, dt in ((fB, np.float64), (force, np.float64), (meio, np.int32))]
    ptr = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in keep]
    flags = (VTK_BINARY if binary else 0) | (VTK_DOUBLE if double else 0)
    rc = lib().lbm_write_vtk(os.fsencode(path), field, nx, ny, fR, ptr[0], ptr[1], ptr[2], flags)
    if rc != 0:
        raise LBMError(f"lbm_write_vtk({path!r}) failed ({rc})")


class Context:
    """A set of cases resident on the GPU (lbm_create)."""

    def __init__(self, cases, precision=PREC_F64):
        L = lib()
        if isinstance(cases, CaseDesc):
            cases = [cases]
        self.descs = list(cases)
        arr = (CaseDesc * len(self.descs))(*self.descs)
        h = C.c_void_p()
        rc = L.lbm_create(arr, len(self.descs), precision, C.byref(h))
        if rc != 0:
            raise LBMError(f"lbm_create failed ({rc})")
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            raise LBMError(f"itolbm error {rc}: {lib().lbm_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_sites(self, case=0):
        d = self.descs[case]
        return d.nx * d.ny

    def upload(self, fR, fB, case=0):
        n = self.n_sites(case) * Q
        fR = np.ascontiguousarray(fR, dtype=np.float64).reshape(-1)
        fB = np.ascontiguousarray(fB, dtype=np.float64).reshape(-1)
        if fR.size != n or fB.size != n:
            raise ValueError("population arrays must hold nx*ny*19 values")
        self._chk(lib().lbm_upload_state(self.h, case, fR, fB))

    def init_drop(self, R, width=3.45, x0=None, y0=None, v0=0.0, case=0):
        """initial_conditions_drop on the host (library code), centre (nx/2, ny/2) by default."""
        d = self.descs[case]
        x0 = d.nx // 2 if x0 is None else int(x0)
        y0 = d.ny // 2 if y0 is None else int(y0)
        self._chk(lib().lbm_init_drop(self.h, case, float(R), float(width), x0, y0, float(v0)))

    def init_band(self, p, q, width=3.45, case=0):
        self._chk(lib().lbm_init_band(self.h, case, int(p), int(q), float(width)))

    def download(self, which=STATE_POST_COLLISION, case=0):
        n = self.n_sites(case) * Q
        fR = np.empty(n)
        fB = np.empty(n)
        self._chk(lib().lbm_download_state(self.h, case, which, fR, fB))
        return fR, fB

    def step(self, n=1, record=False, case=None):
        flags = STEP_RECORD_AUX if record else 0
        if case is None:
            self._chk(lib().lbm_step(self.h, n, flags))
        else:
            self._chk(lib().lbm_step_case(self.h, case, n, flags))

    def step_noalloc(self, n=1, record=False, case=0):
        """lbm_step_case_noalloc: one case, nothing allocated (static scratch on the device)."""
        self._chk(lib().lbm_step_case_noalloc(self.h, case, n, STEP_RECORD_AUX if record else 0))

    def step_mode(self):
        """lbm_step_mode: -1 undecided, 0 one launch per family, 1 one launch, 2 one persistent launch."""
        v = C.c_int()
        self._chk(lib().lbm_step_mode(self.h, C.byref(v)))
        return v.value

    def stored_bytes(self, case=0):
        """lbm_stored_bytes: bytes a step reads and writes per lattice update of the case's stored form."""
        v = C.c_int()
        self._chk(lib().lbm_stored_bytes(self.h, case, C.byref(v)))
        return v.value

    def tile_shape(self, case=0):
        """(rows per segment, pipeline depth, tuned) lbm_step uses for the case's kernel family."""
        r, d, t = C.c_int(), C.c_int(), C.c_int()
        self._chk(lib().lbm_tile_shape(self.h, case, C.byref(r), C.byref(d), C.byref(t)))
        return r.value, d.value, bool(t.value)

    def set_stream(self, hip_stream):
        """Queues later work on a caller's hipStream_t (an int handle, e.g. torch's cuda_stream); None = own."""
        self._chk(lib().lbm_set_stream(self.h, None if hip_stream is None else C.c_void_p(int(hip_stream))))

    def sync(self):
        self._chk(lib().lbm_sync(self.h))

    def set_graph(self, steps_per_graph):
        """Steps per captured HIP graph in step() (even; 0 = launch by launch)."""
        self._chk(lib().lbm_set_graph(self.h, int(steps_per_graph)))

    def set_active(self, active, case=0):
        """Includes or leaves out a case from step() over all cases."""
        self._chk(lib().lbm_set_active(self.h, case, 1 if active else 0))

    def set_geometry(self, meio, wett=0.0, case=0):
        """lbm_set_geometry: meio (nx*ny, 0 = solid) as read_geometry gives it, or None for all fluid."""
        if meio is None:
            self._chk(lib().lbm_set_geometry(self.h, case, None, float(wett)))
            return
        m = np.ascontiguousarray(meio, dtype=np.int32).ravel()
        if m.size != self.n_sites(case):
            raise ValueError(f"meio has {m.size} entries, the case {self.n_sites(case)} sites")
        self._chk(lib().lbm_set_geometry(self.h, case, m.ctypes.data_as(C.c_void_p), float(wett)))

    def set_reference(self, rho_red, rho_blue, case=0):
        """Reference phase densities of the FP32/FP16 deviation storage (no effect on FP64)."""
        self._chk(lib().lbm_set_reference(self.h, case, float(rho_red), float(rho_blue)))

    def aux(self, case=0):
        n = self.n_sites(case)
        gm = np.empty(3 * n)
        F = np.empty(3 * n)
        self._chk(lib().lbm_download_aux(self.h, case, gm.ctypes.data_as(C.c_void_p), F.ctypes.data_as(C.c_void_p)))
        return gm, F

    def diag_drop(self, v0=0.0, step=0, x_ini=None, y_ini=None, case=0):
        d = self.descs[case]
        x_ini = float(d.nx // 2) if x_ini is None else float(x_ini)
        y_ini = float(d.ny // 2) if y_ini is None else float(y_ini)
        out = DropDiag()
        self._chk(lib().lbm_diag_drop(self.h, case, float(v0), int(step), x_ini, y_ini, C.byref(out)))
        return out

    def diag_incl(self, p, q, profile=False, case=0):
        out = InclDiag()
        prof = np.empty(5 * self.n_sites(case)) if profile else None
        self._chk(lib().lbm_diag_incl(self.h, case, int(p), int(q), C.byref(out),
                                      None if prof is None else prof.ctypes.data_as(C.c_void_p)))
        return (out, prof.reshape(-1, 5)) if profile else out

    @property
    def device_bytes(self):
        return lib().lbm_device_bytes(self.h)
