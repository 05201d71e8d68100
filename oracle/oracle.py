"""ctypes wrapper around the CPU oracle (test infrastructure only).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module.  It is the checker: the product path
never calls it.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_mesh.argtypes = [C.c_void_p, _ip, C.c_int, C.c_int]
        L.orc_num_edges.argtypes = [C.c_void_p]
        L.orc_get_edges.argtypes = [C.c_void_p, _ip]
        L.orc_build_springs.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.orc_build_stretch.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.orc_set_pins.argtypes = [C.c_void_p, _ip, _dp, C.c_int, C.c_double]
        L.orc_set_pin_projection.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_params.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_reset_warm_start.argtypes = [C.c_void_p]
        L.orc_set_dx.argtypes = [C.c_void_p, _dp]
        L.orc_get_dx.argtypes = [C.c_void_p, _dp]
        L.orc_step.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _ip, _dp, _ip]
        L.orc_nnz.argtypes = [C.c_void_p]
        L.orc_assemble.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double]
        L.orc_build_topology.argtypes = [C.c_void_p]
        L.orc_get_pattern.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_get_slots.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.orc_get_system.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get_rest.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_last_stats.argtypes = [C.c_void_p, _ip, _dp, _ip, _ip]
        L.orc_spring_element.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp]
        L.orc_triangle_element.argtypes = [_dp, _dp, C.c_double, C.c_float, _dp, _dp, _dp, _dp]
        L.orc_build_shear.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.orc_build_bending.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_num_dihedrals.argtypes = [C.c_void_p]
        L.orc_get_bending.argtypes = [C.c_void_p, _ip, _dp, _ip, _ip]
        L.orc_dihedral_element.argtypes = [_dp, _dp, _dp]
        L.orc_find_dihedrals.argtypes = [_ip, C.c_int, _ip]
        L.orc_add_plane.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_add_box.argtypes = [C.c_void_p, _dp]
        L.orc_clear_shapes.argtypes = [C.c_void_p]
        L.orc_set_contact.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_box_sdf.argtypes = [_dp, _dp, _dp]
        L.orc_box_sdf.restype = C.c_double
        L.orc_mat3_inverse.argtypes = [_dp, _dp]
        L.orc_pcg_solve.argtypes = [C.c_int, _dp, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, _ip, _dp]
        L.orc_time_steps.restype = C.c_double
        L.orc_time_steps.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, C.c_int, _ip]
        _LIB = L
    return _LIB


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def _check(rc):
    if rc != 0:
        msg = lib().orc_last_error().decode()
        if rc == 1:
            raise ValueError(msg)
        if rc == 3:
            raise IndexError(msg)
        raise RuntimeError(msg)


class OracleCloth:
    """One cloth in the oracle.  precision=64 (FP64) or 32 (FP32 matrix and
    vectors with FP64 reductions)."""

    def __init__(self, precision: int = 64):
        self._L = lib()
        self._h = self._L.orc_create(precision)
        self.n = 0

    def __del__(self):
        try:
            if self._h:
                self._L.orc_destroy(self._h)
        except Exception:
            pass

    def set_mesh(self, tris: np.ndarray, n: int):
        t = np.ascontiguousarray(tris, np.int32)
        self.n = n
        _check(self._L.orc_set_mesh(self._h, _i(t), t.shape[0], n))

    def edges(self) -> np.ndarray:
        e = np.zeros((self._L.orc_num_edges(self._h), 2), np.int32)
        _check(self._L.orc_get_edges(self._h, _i(e)))
        return e

    def build_springs(self, pos, k):
        p = _f64(pos)
        if np.ndim(k) == 0:
            _check(self._L.orc_build_springs(self._h, _d(p), None, float(k)))
        else:
            kk = _f64(k)
            _check(self._L.orc_build_springs(self._h, _d(p), _d(kk), 0.0))

    def build_stretch(self, pos, k):
        p = _f64(pos)
        if np.ndim(k) == 0:
            _check(self._L.orc_build_stretch(self._h, _d(p), None, float(k)))
        else:
            kk = _f64(k)
            _check(self._L.orc_build_stretch(self._h, _d(p), _d(kk), 0.0))

    def build_shear(self, pos, k):
        """FEM shear: the stretch element with the material axes rotated by pi/4."""
        p = _f64(pos)
        if np.ndim(k) == 0:
            _check(self._L.orc_build_shear(self._h, _d(p), None, float(k)))
        else:
            kk = _f64(k)
            _check(self._L.orc_build_shear(self._h, _d(p), _d(kk), 0.0))

    def build_bending(self, pos, k):
        _check(self._L.orc_build_bending(self._h, _d(_f64(pos)), float(k)))

    def num_dihedrals(self):
        return int(self._L.orc_num_dihedrals(self._h))

    def bending(self, ntri):
        """(dihedrals (D,4), rest angles (D,), shear slots (T,9), dihedral slots (D,16));
        slots need build_topology() first."""
        nd = self.num_dihedrals()
        d = np.zeros((nd, 4), np.int32)
        r = np.zeros(nd)
        ss = np.zeros((ntri, 9), np.int32)
        ds = np.zeros((nd, 16), np.int32)
        _check(self._L.orc_get_bending(self._h, _i(d), _d(r), _i(ss), _i(ds)))
        return d, r, ss, ds

    def add_plane(self, n, d):
        _check(self._L.orc_add_plane(self._h, _d(_f64(n).reshape(3)), float(d)))

    def add_box(self, center, half_ext, ex=(1, 0, 0), ey=(0, 1, 0), ez=(0, 0, 1)):
        b = np.concatenate([_f64(v).reshape(3) for v in (center, half_ext, ex, ey, ez)])
        _check(self._L.orc_add_box(self._h, _d(b)))

    def clear_shapes(self):
        _check(self._L.orc_clear_shapes(self._h))

    def set_contact(self, thickness, stiffness, friction=0.0, friction_epsilon=1e-4):
        _check(self._L.orc_set_contact(self._h, float(thickness), float(stiffness), float(friction),
                                       float(friction_epsilon)))

    def set_pins(self, idx, targets, k=1.0e6):
        i = np.ascontiguousarray(idx, np.int32)
        t = _f64(targets).reshape(-1, 3)
        _check(self._L.orc_set_pins(self._h, _i(i), _d(t), i.shape[0], float(k)))

    def set_pin_projection(self, on=True):
        """Pins as constraints (extension): pinned displacements fixed at
        target - x_n and eliminated; the particle ends each step at its target."""
        _check(self._L.orc_set_pin_projection(self._h, int(bool(on))))

    def set_params(self, gravity=(0.0, 0.0, -9.81), damping=0.0, springs_active=False,
                   force_active=False, tol=0.0, max_iter=50, threads=1):
        _check(self._L.orc_set_params(self._h, float(gravity[0]), float(gravity[1]), float(gravity[2]),
                                      float(damping), int(springs_active), int(force_active),
                                      float(tol), int(max_iter), int(threads)))

    def reset_warm_start(self):
        _check(self._L.orc_reset_warm_start(self._h))

    def set_dx(self, dx):
        d = np.ascontiguousarray(dx, np.float64).reshape(-1, 3)
        assert d.shape[0] == self.n
        _check(self._L.orc_set_dx(self._h, _d(d)))

    def get_dx(self):
        d = np.zeros((self.n, 3))
        _check(self._L.orc_get_dx(self._h, _d(d)))
        return d

    def step(self, pos, vel, inv_mass=None, force=None, dt=1.0 / 60.0, nsteps=1):
        """Advance in place (pos, vel: float64 (N,3) arrays).  Returns per-step
        (iters, relres, converged)."""
        assert pos.dtype == np.float64 and pos.flags.c_contiguous
        assert vel.dtype == np.float64 and vel.flags.c_contiguous
        im = _f64(inv_mass)
        f = _f64(force)
        it = np.zeros(nsteps, np.int32)
        rr = np.zeros(nsteps, np.float64)
        cv = np.zeros(nsteps, np.int32)
        _check(self._L.orc_step(self._h, _d(pos), _d(vel), _d(im), _d(f), float(dt), nsteps,
                                _i(it), _d(rr), _i(cv)))
        return it, rr, cv

    def time_steps(self, pos, vel, inv_mass=None, force=None, dt=1.0 / 60.0, nsteps=1):
        im = _f64(inv_mass)
        f = _f64(force)
        it = C.c_int(0)
        s = self._L.orc_time_steps(self._h, _d(pos), _d(vel), _d(im), _d(f), float(dt), nsteps, C.byref(it))
        if s < 0:
            _check(2)
        return s, it.value

    def assemble(self, pos, vel, inv_mass=None, force=None, dt=1.0 / 60.0):
        """Build the step's system at (pos, vel) without solving; read it
        with system()."""
        _check(self._L.orc_assemble(self._h, _d(_f64(pos)), _d(_f64(vel)), _d(_f64(inv_mass)),
                                    _d(_f64(force)), float(dt)))

    def build_topology(self):
        _check(self._L.orc_build_topology(self._h))

    def pattern(self):
        nnz = self._L.orc_nnz(self._h)
        ro = np.zeros(self.n + 1, np.int32)
        co = np.zeros(nnz, np.int32)
        _check(self._L.orc_get_pattern(self._h, _i(ro), _i(co)))
        return ro, co

    def slots(self, ntri, nspring, npin):
        ts = np.zeros(9 * ntri, np.int32)
        ss = np.zeros(4 * nspring, np.int32)
        ps = np.zeros(npin, np.int32)
        _check(self._L.orc_get_slots(self._h, _i(ts), _i(ss), _i(ps)))
        return ts, ss, ps

    def system(self):
        nnz = self._L.orc_nnz(self._h)
        diag = np.zeros((self.n, 3, 3))
        vals = np.zeros((nnz, 3, 3))
        rhs = np.zeros((self.n, 3))
        dx = np.zeros((self.n, 3))
        _check(self._L.orc_get_system(self._h, _d(diag), _d(vals), _d(rhs), _d(dx)))
        return diag, vals, rhs, dx

    def rest(self, nedge, ntri):
        sr = np.zeros(nedge)
        dm = np.zeros((ntri, 2, 2))
        ar = np.zeros(ntri)
        _check(self._L.orc_get_rest(self._h, _d(sr), _d(dm), _d(ar)))
        return sr, dm, ar

    def last_stats(self):
        it = C.c_int(0)
        rr = C.c_double(0)
        cv = C.c_int(0)
        rs = C.c_int(0)
        _check(self._L.orc_last_stats(self._h, C.byref(it), C.byref(rr), C.byref(cv), C.byref(rs)))
        return it.value, rr.value, bool(cv.value), rs.value


def oracle_from_scene(scene, precision=64, springs_active=True, force_active=True, tol=1e-6,
                      max_iter=10000, threads=1):
    """Build an oracle cloth from a chysx.scenes.Scene.  Returns (cloth, pos,
    vel, inv_mass, force) with float64 state arrays ready for step()."""
    oc = OracleCloth(precision)
    oc.set_mesh(scene.tris, scene.n)
    pos = np.ascontiguousarray(scene.pos, np.float64)
    vel = np.ascontiguousarray(scene.vel, np.float64)
    oc.build_springs(pos, scene.spring_k)
    oc.build_stretch(pos, scene.stretch_k)
    if getattr(scene, "shear_k", None) is not None:
        oc.build_shear(pos, scene.shear_k)
    if getattr(scene, "bend_k", None) is not None:
        oc.build_bending(pos, scene.bend_k)
    oc.set_pins(scene.pin_idx, scene.pin_targets, scene.pin_k)
    oc.set_params(scene.gravity, scene.damping, springs_active, force_active and scene.force is not None,
                  tol, max_iter, threads)
    return oc, pos, vel, _f64(scene.inv_mass), _f64(scene.force)


# ---- element-level helpers (FP64) ---------------------------------------

def spring_element(xa, xb, rest, k):
    x = np.ascontiguousarray(np.concatenate([xa, xb]), np.float64)
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_spring_element(_d(x), float(rest), float(k), _d(g), _d(H))
    return g, H.reshape(3, 3)


def triangle_element(x_rest, x, k, theta=0.0):
    """Stretch (theta = 0) or shear (theta = pi/4) triangle: gradient (3,3),
    Hessian blocks (3,3,3,3) [a][b], D_m^-1 R^-1 (2,2) and rest area."""
    xr = np.ascontiguousarray(x_rest, np.float64).reshape(9)
    xx = np.ascontiguousarray(x, np.float64).reshape(9)
    g = np.zeros(9)
    H = np.zeros(81)
    dm = np.zeros(4)
    ar = np.zeros(1)
    lib().orc_triangle_element(_d(xr), _d(xx), float(k), float(theta), _d(g), _d(H), _d(dm), _d(ar))
    return g.reshape(3, 3), H.reshape(3, 3, 3, 3), dm.reshape(2, 2), float(ar[0])


SHEAR_THETA = 0.7853981633974483  # pi/4, cloth_simulator.cu:428


def dihedral_element(x):
    """(theta, a (4,3), valid) of the dihedral x (4,3): v0, v1 hinge, v2, v3 wings."""
    xx = np.ascontiguousarray(x, np.float64).reshape(12)
    th = np.zeros(1)
    a = np.zeros(12)
    ok = lib().orc_dihedral_element(_d(xx), _d(th), _d(a))
    return float(th[0]), a.reshape(4, 3), bool(ok)


def box_sdf(center, half_ext, ex, ey, ez, x):
    b = np.concatenate([np.asarray(v, np.float64).reshape(3) for v in (center, half_ext, ex, ey, ez)])
    n = np.zeros(3)
    d = lib().orc_box_sdf(_d(b), _d(np.ascontiguousarray(x, np.float64).reshape(3)), _d(n))
    return float(d), n


def find_dihedrals(tris):
    t = np.ascontiguousarray(tris, np.int32)
    n = lib().orc_find_dihedrals(_i(t), t.shape[0], None)
    out = np.zeros((n, 4), np.int32)
    lib().orc_find_dihedrals(_i(t), t.shape[0], _i(out))
    return out


def mat3_inverse(m):
    a = np.ascontiguousarray(m, np.float64).reshape(9)
    o = np.zeros(9)
    lib().orc_mat3_inverse(_d(a), _d(o))
    return o.reshape(3, 3)


def pcg_solve(diag, row_offsets, cols, values, b, x0=None, tol=1e-10, max_iter=1000):
    n = diag.shape[0]
    x = np.zeros((n, 3)) if x0 is None else np.array(x0, np.float64)
    it = C.c_int(0)
    rr = C.c_double(0)
    _check(lib().orc_pcg_solve(n, _d(_f64(diag)), _i(np.ascontiguousarray(row_offsets, np.int32)),
                               _i(np.ascontiguousarray(cols, np.int32)), _d(_f64(values)), _d(_f64(b)),
                               _d(x), float(tol), int(max_iter), C.byref(it), C.byref(rr)))
    return x, it.value, rr.value
