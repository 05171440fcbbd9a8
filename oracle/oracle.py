"""ctypes binding of the CPU oracle (libzdx_oracle.so) and, where it was built,
of the reference harness (oracle/_ref/libzdx_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libzdx_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libzdx_ref.so")
REF_BIN = os.path.join(HERE, "_ref", "zdem_cpu")
REF_DIR = "/root/reference"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class PairCounts(C.Structure):
    _fields_ = [("tri_checks", C.c_int64), ("segments", C.c_int32), ("comps", C.c_int32),
                ("loops_raw", C.c_int32), ("loops", C.c_int32), ("loop_verts", C.c_int32),
                ("overflow", C.c_int32)]


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir(REF_DIR):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF_DIR={REF_DIR}"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.orc_load_stl.argtypes = [C.c_char_p, C.c_void_p, C.c_int]
        L.orc_build_mesh.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp]
        L.orc_pair_contact.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                       C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                       C.c_void_p, C.c_int, C.POINTER(PairCounts)]
        L.orc_sim_from_config.restype = C.c_void_p
        L.orc_sim_from_config.argtypes = [C.c_char_p, C.c_char_p]
        L.orc_sim_destroy.argtypes = [C.c_void_p]
        L.orc_sim_num_particles.argtypes = [C.c_void_p]
        L.orc_sim_get_state.argtypes = [C.c_void_p, _dp]
        L.orc_sim_set_state.argtypes = [C.c_void_p, _dp]
        L.orc_sim_get_forces.argtypes = [C.c_void_p, _dp, _dp, _ip]
        L.orc_sim_step.argtypes = [C.c_void_p, C.c_int, _lp]
        L.orc_sim_dt.argtypes = [C.c_void_p]
        L.orc_sim_dt.restype = C.c_double
        L.orc_sim_set_force_model.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_tri_tri.argtypes = [_dp, _dp, _dp]
        L.orc_sn_gn.argtypes = [_dp, C.c_int, _dp, _dp]
        L.orc_xc0.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.orc_loop_start_order.argtypes = [_ip, C.c_int, _ip]
        L.orc_sh_clip.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.orc_wall_contact.argtypes = [_dp, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_integrate.argtypes = [_dp, _dp, C.c_double, _dp, _dp, _dp, C.c_double]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference harness library (only where it was built)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.ref_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        R.ref_load_stl.argtypes = [C.c_char_p, C.c_void_p, C.c_int]
        R.ref_build_mesh.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp]
        R.ref_transform.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        R.ref_tri_tri.argtypes = [_dp, _dp, _dp]
        R.ref_pair_segments.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                        C.POINTER(C.c_int64)]
        R.ref_components.argtypes = [_dp, C.c_int, C.c_double, _ip, _ip]
        R.ref_loops.argtypes = [_dp, C.c_int, C.c_double, _dp, C.POINTER(C.c_int), _ip, _ip, C.c_int]
        R.ref_sn_xc0.argtypes = [_dp, C.c_int, _dp, _dp, _dp, _dp, _dp]
        R.ref_sh_clip.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int]
        R.ref_wall_contact.argtypes = [_dp, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp]
        R.ref_integrate.argtypes = [_dp, _dp, C.c_double, _dp, _dp, _dp, C.c_double]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def load_stl(path: str) -> np.ndarray:
    L = lib()
    T = L.orc_load_stl(path.encode(), None, 0)
    if T < 0:
        raise IOError(path)
    out = np.zeros((T, 3, 3))
    L.orc_load_stl(path.encode(), out.ctypes.data, T)
    return out


def build_mesh(tris: np.ndarray, center: bool = True):
    tris = _f64(tris)
    T = tris.shape[0]
    out = np.zeros_like(tris)
    props = np.zeros(16)
    nv = lib().orc_build_mesh(tris.reshape(-1), T, int(center), out.reshape(-1), props)
    return out, props, nv


def pair_contact(trisA, trisB, posA, quatA, posB, quatB, mean_edge_a, bbox_diag_a=0.0,
                 split=True, k=1.0, seg_cap=200000, loop_cap=4096, vtx_cap=200000):
    L = lib()
    tA, tB = _f64(trisA), _f64(trisB)
    seg = np.zeros((seg_cap, 6))
    rec = np.zeros((loop_cap, 28))
    vtx = np.zeros((vtx_cap, 3))
    c = PairCounts()
    rc = L.orc_pair_contact(tA.reshape(-1), tA.shape[0], tB.reshape(-1), tB.shape[0], _f64(posA), _f64(quatA),
                            _f64(posB), _f64(quatB), float(mean_edge_a), float(bbox_diag_a), int(split), float(k),
                            seg.ctypes.data, seg_cap, rec.ctypes.data, loop_cap, vtx.ctypes.data, vtx_cap,
                            C.byref(c))
    if rc != 0:
        raise RuntimeError("oracle pair capacity exceeded")
    return dict(tri_checks=c.tri_checks, segments=seg[:c.segments].copy(), comps=c.comps,
                loops_raw=c.loops_raw, loops=rec[:c.loops].copy(), loop_vtx=vtx[:c.loop_verts].copy())


class Sim:
    """The oracle time-stepper on a zdem_cpu config file."""

    def __init__(self, config_path: str, base_dir: str | None = None):
        L = lib()
        self._L = L
        self.h = L.orc_sim_from_config(config_path.encode(), base_dir.encode() if base_dir else None)
        if not self.h:
            raise IOError(f"oracle could not load config {config_path}")
        self.n = L.orc_sim_num_particles(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_sim_destroy(self.h)
            self.h = None

    def step(self, n: int = 1) -> np.ndarray:
        ctr = np.zeros(7, dtype=np.int64)
        self._L.orc_sim_step(self.h, n, ctr)
        return ctr

    def state(self) -> np.ndarray:
        s = np.zeros((self.n, 16))
        self._L.orc_sim_get_state(self.h, s.reshape(-1))
        return s

    def set_state(self, s: np.ndarray) -> None:
        self._L.orc_sim_set_state(self.h, _f64(s).reshape(-1))

    def set_force_model(self, model: int, k: float = 0.0) -> None:
        """0: the reference's particle-particle force; 1: minimal-energy force with stiffness k."""
        self._L.orc_sim_set_force_model(self.h, int(model), float(k))

    def forces(self):
        F = np.zeros((self.n, 3))
        T = np.zeros((self.n, 3))
        c = np.zeros(self.n, dtype=np.int32)
        self._L.orc_sim_get_forces(self.h, F.reshape(-1), T.reshape(-1), c)
        return F, T, c


def parse_state_txt(path: str) -> np.ndarray:
    """Reads a write_particle_state_txt dump (src/main.cpp:1842-1870):
    returns (N, 13) = pos quat vel omega."""
    rows, cur = [], {}
    for line in open(path):
        s = line.replace("=", " ").split()
        if not s:
            continue
        if s[0] == "particle":
            cur = {}
        elif s[0] in ("pos", "vel", "omega", "quat"):
            cur[s[0]] = [float(x) for x in s[1:]]
        elif s[0] == "end_particle":
            rows.append(cur["pos"] + cur["quat"] + cur["vel"] + cur["omega"])
    return np.asarray(rows, dtype=np.float64)
