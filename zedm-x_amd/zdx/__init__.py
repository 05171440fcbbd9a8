"""zdx — Python binding of libzdx.so, the MI355X Route-B contact step.

The library is the product: every call here goes through its C ABI
(include/zdx.h).  There is no CPU fallback; importing works anywhere, but
creating a Context needs the built library and a HIP device and raises
otherwise.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libzdx.so")
CLI_PATH = os.path.join(PKG, "bin", "zdx")

OK, BAD_ARG, IO, CAPACITY, HIP_ERROR, INTERNAL = range(6)
MODE_PLAIN, MODE_BATCHED = 0, 1
FORCE_ZDEM, FORCE_MINIMAL, FORCE_ZDEM_DAMPED = 0, 1, 2
COUNTERS = ["contacts", "tri_checks", "segments", "comps", "loops", "pairs", "wall_contacts", "tri_tests", "tri_hits",
            "tri_scanned", "topo_pairs", "seg_max"]
EXPORTS = ["zdx_create", "zdx_destroy", "zdx_last_error", "zdx_load_stl", "zdx_mesh_register", "zdx_contact_pairs",
           "zdx_contact_times", "zdx_contact_pairs_noalloc", "zdx_wall_contacts",
           "zdx_sim_load_config", "zdx_sim_create", "zdx_sim_destroy", "zdx_sim_num_particles", "zdx_sim_step",
           "zdx_sim_get_state", "zdx_sim_set_state", "zdx_sim_get_forces", "zdx_sim_get_pairs", "zdx_sim_write_vtk",
           "zdx_sim_write_state_txt", "zdx_sim_checkpoint", "zdx_sim_restore",
           "zdx_sim_stage_times", "zdx_sim_set_force_model", "zdx_sim_set_tiles", "zdx_sim_autotune"]
LOOP_REC = 28


class ZdxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"zdx status {status}: {msg}")
        self.status = status


class PairIn(C.Structure):
    _fields_ = [("mesh_a", C.c_int32), ("mesh_b", C.c_int32), ("pos_a", C.c_double * 3), ("quat_a", C.c_double * 4),
                ("pos_b", C.c_double * 3), ("quat_b", C.c_double * 4)]


# zdx_pair_in as a numpy record (same offsets and size as the C struct), so a
# batch of pairs is filled with array assignments
PAIR_IN_DTYPE = np.dtype({"names": ["mesh_a", "mesh_b", "pos_a", "quat_a", "pos_b", "quat_b"],
                          "formats": [np.int32, np.int32, (np.float64, 3), (np.float64, 4), (np.float64, 3),
                                      (np.float64, 4)],
                          "offsets": [PairIn.mesh_a.offset, PairIn.mesh_b.offset, PairIn.pos_a.offset,
                                      PairIn.quat_a.offset, PairIn.pos_b.offset, PairIn.quat_b.offset],
                          "itemsize": C.sizeof(PairIn)})


# zdx_body_pose (mesh, reserved, pos[3], quat[4]: 64 bytes)
BODY_POSE_DTYPE = np.dtype({"names": ["mesh", "pos", "quat"], "formats": [np.int32, (np.float64, 3), (np.float64, 4)],
                            "offsets": [0, 8, 32], "itemsize": 64})


class ContactOpts(C.Structure):
    _fields_ = [("split_contacts", C.c_int32), ("mode", C.c_int32), ("k", C.c_double), ("sensitivities", C.c_int32),
                ("reserved", C.c_int32), ("tol", C.c_double)]


class PairOut(C.Structure):
    _fields_ = [("tri_checks", C.c_void_p), ("n_segments", C.c_void_p), ("n_comps", C.c_void_p),
                ("n_loops_raw", C.c_void_p), ("loop_offset", C.c_void_p), ("seg_offset", C.c_void_p),
                ("loop_rec", C.c_void_p), ("loop_cap", C.c_int64), ("vtx_offset", C.c_void_p),
                ("loop_vtx", C.c_void_p), ("vtx_cap", C.c_int64), ("segments", C.c_void_p), ("seg_cap", C.c_int64),
                ("n_loops_total", C.c_int64), ("n_vtx_total", C.c_int64), ("n_seg_total", C.c_int64),
                ("loop_jac", C.c_void_p)]


class SimDesc(C.Structure):
    _fields_ = [("n_particles", C.c_int32), ("mesh_id", C.c_void_p), ("init", C.c_void_p), ("material", C.c_void_p),
                ("n_walls", C.c_int32), ("wall_mesh_id", C.c_void_p), ("wall_pose", C.c_void_p),
                ("wall_material", C.c_void_p), ("gravity", C.c_double * 3), ("dt", C.c_double),
                ("split_contacts", C.c_int32)]


_lib = None


def lib():
    """Loads libzdx.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libzdx.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        L.zdx_create.argtypes = [C.POINTER(vp), i32]
        L.zdx_destroy.argtypes = [vp]
        L.zdx_destroy.restype = None
        L.zdx_last_error.argtypes = [vp]
        L.zdx_last_error.restype = C.c_char_p
        L.zdx_load_stl.argtypes = [C.c_char_p, vp, i32, C.POINTER(i32)]
        L.zdx_mesh_register.argtypes = [vp, vp, i32, i32, C.POINTER(i32), vp]
        L.zdx_contact_pairs.argtypes = [vp, vp, i32, C.POINTER(ContactOpts), C.POINTER(PairOut)]
        L.zdx_contact_times.argtypes = [vp, vp, i32]
        L.zdx_contact_pairs_noalloc.argtypes = [vp, vp, i32, C.POINTER(ContactOpts), C.POINTER(PairOut)]
        L.zdx_wall_contacts.argtypes = [vp, vp, i32, vp, vp, i32, i32, vp]
        L.zdx_sim_load_config.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(vp)]
        L.zdx_sim_create.argtypes = [vp, C.POINTER(SimDesc), C.POINTER(vp)]
        L.zdx_sim_destroy.argtypes = [vp]
        L.zdx_sim_destroy.restype = None
        L.zdx_sim_num_particles.argtypes = [vp]
        L.zdx_sim_step.argtypes = [vp, i32, i32, vp]
        L.zdx_sim_get_state.argtypes = [vp, vp]
        L.zdx_sim_set_state.argtypes = [vp, vp]
        L.zdx_sim_get_forces.argtypes = [vp, vp, vp, vp]
        L.zdx_sim_get_pairs.argtypes = [vp, vp, vp, i64, C.POINTER(i64)]
        L.zdx_sim_write_vtk.argtypes = [vp, C.c_char_p]
        L.zdx_sim_write_state_txt.argtypes = [vp, C.c_char_p]
        L.zdx_sim_checkpoint.argtypes = [vp, C.c_char_p]
        L.zdx_sim_restore.argtypes = [vp, C.c_char_p]
        L.zdx_sim_stage_times.argtypes = [vp, vp, i32]
        L.zdx_sim_set_force_model.argtypes = [vp, i32, C.c_double]
        L.zdx_sim_set_tiles.argtypes = [vp, i32, i32]
        L.zdx_sim_autotune.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
        for f in EXPORTS:
            if f not in ("zdx_destroy", "zdx_sim_destroy", "zdx_last_error", "zdx_sim_num_particles"):
                getattr(L, f).restype = i32
        L.zdx_sim_num_particles.restype = i32
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def load_stl(path: str) -> np.ndarray:
    L = lib()
    T = C.c_int32(0)
    st = L.zdx_load_stl(path.encode(), None, 0, C.byref(T))
    if st != OK:
        raise ZdxError(st, f"cannot load {path}")
    out = np.zeros((T.value, 3, 3))
    st = L.zdx_load_stl(path.encode(), out.ctypes.data, T.value, C.byref(T))
    if st != OK:
        raise ZdxError(st, f"cannot load {path}")
    return out


class Context:
    """A device context (one HIP stream) owning registered meshes."""

    def __init__(self, device: int = 0):
        L = lib()
        self._L = L
        h = C.c_void_p()
        st = L.zdx_create(C.byref(h), int(device))
        if st != OK:
            raise ZdxError(st, "zdx_create failed (no HIP device?)")
        self.h = h
        self.mesh_props: dict[int, np.ndarray] = {}
        self._sims = weakref.WeakSet()

    def close(self):
        """Destroys the context, after every simulation still open on it (a
        simulation must not outlive its context)."""
        if getattr(self, "h", None):
            for sim in list(self._sims):
                sim.close()
            self._L.zdx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        if st != OK:
            raise ZdxError(st, self._L.zdx_last_error(self.h).decode())

    def register_mesh(self, tris: np.ndarray, center: bool = True) -> int:
        t = np.ascontiguousarray(tris, dtype=np.float64)
        mid = C.c_int32(0)
        props = np.zeros(16)
        self._check(self._L.zdx_mesh_register(self.h, t.ctypes.data, t.shape[0], int(center), C.byref(mid),
                                              props.ctypes.data))
        self.mesh_props[mid.value] = props
        return mid.value

    def contact_pairs(self, mesh_a, mesh_b, pos_a, quat_a, pos_b, quat_b, split=True, mode=MODE_BATCHED, k=1.0,
                      segments=False, vertices=True, sensitivities=False, tol=0.0, noalloc=False):
        """Batched per-pair contact evaluation.  Arrays are per pair; returns a
        dict with per-pair counters, per-loop records (ZDX_LOOP_REC layout) and,
        optionally, loop vertices and narrow-phase segments.  tol > 0 replaces
        mesh A's snap / adjacency tolerance for every pair.  noalloc: through
        zdx_contact_pairs_noalloc (no device allocation; the library's
        per-GPU scratch array bounds the batch)."""
        n = len(mesh_a)
        arr = np.zeros(max(n, 1), dtype=PAIR_IN_DTYPE)
        if n:
            arr["mesh_a"][:n] = np.asarray(mesh_a, dtype=np.int32)
            arr["mesh_b"][:n] = np.asarray(mesh_b, dtype=np.int32)
            arr["pos_a"][:n] = np.asarray(pos_a, dtype=np.float64).reshape(n, 3)
            arr["quat_a"][:n] = np.asarray(quat_a, dtype=np.float64).reshape(n, 4)
            arr["pos_b"][:n] = np.asarray(pos_b, dtype=np.float64).reshape(n, 3)
            arr["quat_b"][:n] = np.asarray(quat_b, dtype=np.float64).reshape(n, 4)
        opts = ContactOpts(int(split), int(mode), float(k), int(bool(sensitivities)), 0, float(tol))
        loop_cap, vtx_cap, seg_cap = max(64, 4 * n), max(1024, 64 * n), max(1024, 64 * n)
        for _ in range(4):
            tc = np.zeros(n, dtype=np.int64)
            ns = np.zeros(n, dtype=np.int32)
            nc = np.zeros(n, dtype=np.int32)
            nr = np.zeros(n, dtype=np.int32)
            lo = np.zeros(n + 1, dtype=np.int32)
            so = np.zeros(n + 1, dtype=np.int32) if segments else None
            rec = np.zeros((loop_cap, LOOP_REC))
            vo = np.zeros(loop_cap + 1, dtype=np.int32) if vertices else None
            vtx = np.zeros((vtx_cap, 3)) if vertices else None
            seg = np.zeros((seg_cap, 6)) if segments else None
            jac = np.zeros((loop_cap, 72)) if sensitivities else None
            out = PairOut(tc.ctypes.data, ns.ctypes.data, nc.ctypes.data, nr.ctypes.data, lo.ctypes.data, _ptr(so),
                          rec.ctypes.data, loop_cap, _ptr(vo), _ptr(vtx), vtx_cap, _ptr(seg), seg_cap, 0, 0, 0,
                          _ptr(jac))
            fn = self._L.zdx_contact_pairs_noalloc if noalloc else self._L.zdx_contact_pairs
            st = fn(self.h, arr.ctypes.data, n, C.byref(opts), C.byref(out))
            if st == CAPACITY:
                loop_cap = max(loop_cap, int(out.n_loops_total) + 1)
                vtx_cap = max(vtx_cap, int(out.n_vtx_total) + 1)
                seg_cap = max(seg_cap, int(out.n_seg_total) + 1)
                continue
            self._check(st)
            nl = int(out.n_loops_total)
            res = dict(tri_checks=tc, n_segments=ns, n_comps=nc, n_loops_raw=nr, loop_offset=lo, loops=rec[:nl])
            if sensitivities:
                res["loop_jac"] = jac[:nl]
            if vertices:
                res["vtx_offset"] = vo[:nl + 1]
                res["loop_vtx"] = vtx[:int(out.n_vtx_total)]
            if segments:
                res["seg_offset"] = so
                res["segments"] = seg[:int(out.n_seg_total)]
            return res
        raise ZdxError(CAPACITY, "contact_pairs output did not fit")

    def wall_contacts(self, mesh, pos, quat, wall_tris, wall_normals, jac=False) -> np.ndarray:
        """Contact of each body (mesh id, pos, quat per body) with each wall
        triangle (world coordinates, (W, 3, 3), normals (W, 3)): records
        (n, W, 32) -- flag, area, point, normal, then with jac d area/dxi and
        d point/dxi for rigid motions xi = (d, w) of the body."""
        n = len(mesh)
        wt = np.ascontiguousarray(wall_tris, dtype=np.float64).reshape(-1, 9)
        wn = np.ascontiguousarray(wall_normals, dtype=np.float64).reshape(-1, 3)
        arr = np.zeros(max(n, 1), dtype=BODY_POSE_DTYPE)
        if n:
            arr["mesh"][:n] = np.asarray(mesh, dtype=np.int32)
            arr["pos"][:n] = np.asarray(pos, dtype=np.float64).reshape(n, 3)
            arr["quat"][:n] = np.asarray(quat, dtype=np.float64).reshape(n, 4)
        out = np.zeros((max(n, 1), max(len(wt), 1), 32))
        self._check(self._L.zdx_wall_contacts(self.h, arr.ctypes.data, n, wt.ctypes.data, wn.ctypes.data, len(wt),
                                              int(bool(jac)), out.ctypes.data))
        return out[:n, :len(wt)]

    def contact_times(self) -> dict:
        """Device ms of the last contact_pairs call (transform, narrow phase,
        topology) and its narrow-phase attempts."""
        v = np.zeros(4)
        self._check(self._L.zdx_contact_times(self.h, v.ctypes.data, 4))
        return dict(transform=float(v[0]), narrow=float(v[1]), topology=float(v[2]), attempts=int(v[3]))


class Sim:
    """A simulation (zdem_cpu step) on a context."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.h = handle
        self.n = ctx._L.zdx_sim_num_particles(handle)
        ctx._sims.add(self)

    @classmethod
    def from_config(cls, ctx: Context, path: str, base_dir: str | None = None) -> "Sim":
        h = C.c_void_p()
        ctx._check(ctx._L.zdx_sim_load_config(ctx.h, path.encode(), base_dir.encode() if base_dir else None,
                                              C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def create(cls, ctx: Context, mesh_id, init, material, gravity, dt, split=True, walls=()):
        """init: (N, 13) pos quat vel omega; material: (N, 5) density young poisson mu restitution;
        walls: iterable of (mesh_id, pos(3), quat(4), mu, restitution)."""
        mesh_id = np.ascontiguousarray(mesh_id, dtype=np.int32)
        init = np.ascontiguousarray(init, dtype=np.float64)
        mat = np.zeros((len(mesh_id), 6))
        mat[:, :5] = material
        walls = list(walls)
        wm = np.ascontiguousarray([w[0] for w in walls] or [0], dtype=np.int32)
        wp = np.ascontiguousarray([list(w[1]) + list(w[2]) for w in walls] or [[0.0] * 7], dtype=np.float64)
        wmat = np.ascontiguousarray([[w[3], w[4]] for w in walls] or [[0.0, 0.0]], dtype=np.float64)
        d = SimDesc(len(mesh_id), mesh_id.ctypes.data, init.ctypes.data, mat.ctypes.data, len(walls), wm.ctypes.data,
                    wp.ctypes.data, wmat.ctypes.data, (C.c_double * 3)(*gravity), float(dt), int(split))
        h = C.c_void_p()
        ctx._check(ctx._L.zdx_sim_create(ctx.h, C.byref(d), C.byref(h)))
        return cls(ctx, h)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                self.ctx._L.zdx_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n: int = 1, mode: int = MODE_PLAIN) -> dict:
        ctr = np.zeros(len(COUNTERS), dtype=np.int64)
        self.ctx._check(self.ctx._L.zdx_sim_step(self.h, int(n), int(mode), ctr.ctypes.data))
        return dict(zip(COUNTERS, ctr.tolist()))

    def state(self) -> np.ndarray:
        s = np.zeros((self.n, 16))
        self.ctx._check(self.ctx._L.zdx_sim_get_state(self.h, s.ctypes.data))
        return s

    def set_state(self, s: np.ndarray) -> None:
        s = np.ascontiguousarray(s, dtype=np.float64)
        self.ctx._check(self.ctx._L.zdx_sim_set_state(self.h, s.ctypes.data))

    def forces(self):
        F, T = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        c = np.zeros(self.n, dtype=np.int32)
        self.ctx._check(self.ctx._L.zdx_sim_get_forces(self.h, F.ctypes.data, T.ctypes.data, c.ctypes.data))
        return F, T, c

    def pairs(self) -> np.ndarray:
        """Candidate pairs (i < j, ascending) of the last completed step, (P, 2) int32."""
        n = C.c_int64(0)
        st = self.ctx._L.zdx_sim_get_pairs(self.h, None, None, 0, C.byref(n))
        if st not in (OK, CAPACITY):
            self.ctx._check(st)
        pi = np.zeros(max(n.value, 1), dtype=np.int32)
        pj = np.zeros(max(n.value, 1), dtype=np.int32)
        self.ctx._check(self.ctx._L.zdx_sim_get_pairs(self.h, pi.ctypes.data, pj.ctypes.data, n.value, C.byref(n)))
        return np.stack([pi[:n.value], pj[:n.value]], axis=1)

    def write_vtk(self, path: str) -> None:
        """The reference's per-step particle VTK (write_vtk_particles)."""
        self.ctx._check(self.ctx._L.zdx_sim_write_vtk(self.h, path.encode()))

    def write_state_txt(self, path: str) -> None:
        """The reference's particle state file (write_particle_state_txt)."""
        self.ctx._check(self.ctx._L.zdx_sim_write_state_txt(self.h, path.encode()))

    def checkpoint(self, path: str) -> None:
        self.ctx._check(self.ctx._L.zdx_sim_checkpoint(self.h, path.encode()))

    def restore(self, path: str) -> None:
        self.ctx._check(self.ctx._L.zdx_sim_restore(self.h, path.encode()))

    def set_force_model(self, model: int, k: float = 0.0) -> None:
        """FORCE_ZDEM (the reference's particle-particle force), FORCE_MINIMAL
        (Fn = k/2 |Sn| at the loop centroid) or FORCE_ZDEM_DAMPED (the
        reference's force with its dormant normal damping)."""
        self.ctx._check(self.ctx._L.zdx_sim_set_force_model(self.h, int(model), float(k)))

    STAGES = ["transform", "broadphase", "narrow", "topology", "walls", "fold_integrate"]

    def stage_times(self) -> dict:
        """Mean per-stage device ms and counter sums over the steps since the
        previous call (then reset)."""
        v = np.zeros(10 + len(COUNTERS))
        self.ctx._check(self.ctx._L.zdx_sim_stage_times(self.h, v.ctypes.data, len(v)))
        out = {k: float(v[i]) for i, k in enumerate(self.STAGES)}
        out["steps"] = int(v[6])
        out["counters"] = {k: int(v[7 + i]) for i, k in enumerate(COUNTERS)}
        out["narrow_kernels"] = float(v[7 + len(COUNTERS)])
        out["retries"] = int(v[8 + len(COUNTERS)])
        out["graph_captures"] = int(v[9 + len(COUNTERS)])
        return out

    def set_tiles(self, block: int = 0, groups_per_cu: int = 0) -> None:
        self.ctx._check(self.ctx._L.zdx_sim_set_tiles(self.h, int(block), int(groups_per_cu)))

    def autotune(self):
        b, g = C.c_int32(0), C.c_int32(0)
        self.ctx._check(self.ctx._L.zdx_sim_autotune(self.h, C.byref(b), C.byref(g)))
        return b.value, g.value
