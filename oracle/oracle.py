"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker of the HIP path. The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None
_P = ctypes.c_void_p
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int32)
_PL = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        import nairnmpm as nm  # structures shared with the product's C ABI
        _lib.oracle_create.argtypes = [ctypes.POINTER(nm.GridDesc), ctypes.POINTER(nm.Material), ctypes.c_int32,
                                       ctypes.POINTER(nm.RunParams), ctypes.c_int64, _PD, _PI]
        _lib.oracle_create.restype = _P
        _lib.oracle_destroy.argtypes = [_P]
        _lib.oracle_step.argtypes = [_P, ctypes.c_int32]
        _lib.oracle_error.argtypes = [_P]
        _lib.oracle_error.restype = ctypes.c_char_p
        _lib.oracle_run_task.argtypes = [_P, ctypes.c_int32]
        _lib.oracle_set_threads.argtypes = [_P, ctypes.c_int32]
        _lib.oracle_set_velocity_bcs.argtypes = [_P, ctypes.c_int64, ctypes.POINTER(nm.VelocityBC)]
        _lib.oracle_set_particle_loads.argtypes = [_P, ctypes.c_int64, ctypes.POINTER(nm.ParticleLoad)]
        _lib.oracle_set_particle_tractions.argtypes = [_P, ctypes.c_int64, ctypes.POINTER(nm.ParticleTraction)]
        _lib.oracle_xpic_order.argtypes = [_P]
        _lib.oracle_get_velocities.argtypes = [_P, _PD]
        _lib.oracle_surface_info.argtypes = [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                             _PD, _PD]
        _lib.oracle_get_particles.argtypes = [_P, _PD, _PI]
        _lib.oracle_num_nodes.argtypes = [_P]
        _lib.oracle_num_nodes.restype = ctypes.c_int64
        _lib.oracle_num_fields.argtypes = [_P]
        _lib.oracle_num_fields.restype = ctypes.c_int32
        _lib.oracle_get_grid.argtypes = [_P, _PD, _PD, _PD, _PI]
        _lib.oracle_get_contact.argtypes = [_P, _PD, _PD, _PD]
        _lib.oracle_counters.argtypes = [_P, _PL]
        _lib.oracle_regression_normal.argtypes = [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, _PD, _PD]
        _lib.oracle_regression_normal.restype = ctypes.c_int64
        _lib.oracle_deformed_radius.argtypes = [_P, ctypes.c_int64, _PD]
        _lib.oracle_deformed_radius.restype = ctypes.c_double
        _lib.oracle_shape.argtypes = [ctypes.POINTER(nm.GridDesc), ctypes.c_int32, ctypes.c_int64, _PD, _PD,
                                      ctypes.c_int32, _PL, _PD, _PD, _PD, _PD]
        _lib.oracle_constitutive.argtypes = [ctypes.c_int32, ctypes.POINTER(nm.Material), _PD, _PD]
        _lib.oracle_cpdi_shape.argtypes = [ctypes.POINTER(nm.GridDesc), ctypes.c_int64, _PD, ctypes.c_int32,
                                           _PL, _PD, _PD, _PD, _PD]
        _lib.oracle_contact_node.argtypes = [ctypes.POINTER(nm.Material), ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_double, _PD, _PI, _PD, _PD, _PD, _PD, _PD, _PD]
    return _lib


def _d(a):
    return a.ctypes.data_as(_PD)


class Oracle:
    """CPU restatement of NairnMPM's step on one problem (particles in mpm[] order)."""

    def __init__(self, grid, mats, run, dfields, ifields):
        import nairnmpm as nm
        self.dfields = np.ascontiguousarray(dfields, dtype=np.float64)
        self.ifields = np.ascontiguousarray(ifields, dtype=np.int32)
        self.n = self.dfields.shape[1]
        arr = (nm.Material * len(mats))(*mats)
        self.grid = grid
        self._h = lib().oracle_create(ctypes.byref(grid), arr, len(mats), ctypes.byref(run), self.n,
                                      _d(self.dfields), self.ifields.ctypes.data_as(_PI))

    @classmethod
    def from_problem(cls, prob):
        g, mats, run, _, _ = prob.describe()
        d, i = prob.particles()
        o = cls(g, mats, run, d, i)
        bcs = prob.velocity_bcs()
        if bcs:
            o.set_velocity_bcs(bcs)
        loads = prob.particle_loads()
        if loads:
            o.set_particle_loads(loads)
        tractions = prob.particle_tractions()
        if tractions:
            o.set_particle_tractions(tractions)
        return o

    def set_particle_tractions(self, tractions):
        """traction conditions in list order (nairnmpm.ParticleTraction list)"""
        import nairnmpm as nm
        arr = (nm.ParticleTraction * max(len(tractions), 1))(*tractions)
        if lib().oracle_set_particle_tractions(self._h, len(tractions), arr) != 0:
            raise RuntimeError("bad traction condition")

    def grid_velocities(self):
        """grid velocities vk [nf, 3, nn]"""
        nn = lib().oracle_num_nodes(self._h)
        nf = lib().oracle_num_fields(self._h)
        v = np.zeros(3 * nf * nn)
        lib().oracle_get_velocities(self._h, _d(v))
        return v.reshape(nf, nn, 3).transpose(0, 2, 1)

    def xpic_order(self):
        """the XPIC or FMPM order the last step ran with (0: FLIP)"""
        return int(lib().oracle_xpic_order(self._h))

    def surface_info(self, p, face, dir, thickness=1.0):
        """GetSurfaceInfo of particle p in its current state: (corners, weighted normal), corners
        an (n, 3) array in the reference's order"""
        c = np.zeros(12)
        wn = np.zeros(3)
        nc = lib().oracle_surface_info(self._h, p, face, dir, thickness,
                                       c.ctypes.data_as(_PD), wn.ctypes.data_as(_PD))
        return c[:3 * nc].reshape(nc, 3), wn

    def set_particle_loads(self, loads):
        """external particle loads in list order (nairnmpm.ParticleLoad list)"""
        import nairnmpm as nm
        arr = (nm.ParticleLoad * max(len(loads), 1))(*loads)
        if lib().oracle_set_particle_loads(self._h, len(loads), arr) != 0:
            raise RuntimeError("bad particle load")

    def set_velocity_bcs(self, bcs):
        """grid velocity conditions in list order (nairnmpm.VelocityBC list)"""
        import nairnmpm as nm
        arr = (nm.VelocityBC * max(len(bcs), 1))(*bcs)
        if lib().oracle_set_velocity_bcs(self._h, len(bcs), arr) != 0:
            raise RuntimeError("bad velocity condition")

    def set_threads(self, n):
        """Threads of the timing mode (the CPU baseline): particle loops split over n threads and
        the scatters summed through per-thread grid copies. 1 (the default) is the serial
        restatement that the parity checks use; n > 1 differs from it in grid-sum rounding."""
        lib().oracle_set_threads(self._h, int(n))

    def step(self, n=1):
        if lib().oracle_step(self._h, n) != 0:
            raise RuntimeError("oracle step failed: " + lib().oracle_error(self._h).decode())

    def run_task(self, t):
        if lib().oracle_run_task(self._h, t) != 0:
            raise RuntimeError("oracle task failed: " + lib().oracle_error(self._h).decode())

    def particles(self):
        import nairnmpm as nm
        d = np.zeros((nm.NUM_FIELDS, self.n))
        i = np.zeros((nm.NUM_IFIELDS, self.n), dtype=np.int32)
        lib().oracle_get_particles(self._h, _d(d), i.ctypes.data_as(_PI))
        return d, i

    def grid_fields(self):
        """mass [nf, nn], momentum and force [nf, 3, nn], point counts [nf, nn] of every field"""
        nn = lib().oracle_num_nodes(self._h)
        nf = lib().oracle_num_fields(self._h)
        m = np.zeros(nf * nn)
        pk = np.zeros(3 * nf * nn)
        f = np.zeros(3 * nf * nn)
        cnt = np.zeros(nf * nn, dtype=np.int32)
        lib().oracle_get_grid(self._h, _d(m), _d(pk), _d(f), cnt.ctypes.data_as(_PI))
        return (m.reshape(nf, nn), pk.reshape(nf, nn, 3).transpose(0, 2, 1),
                f.reshape(nf, nn, 3).transpose(0, 2, 1), cnt.reshape(nf, nn))

    def contact_values(self):
        """volume [nf, nn], displacement and volume gradient [nf, 3, nn] (multimaterial mode)"""
        nn = lib().oracle_num_nodes(self._h)
        nf = lib().oracle_num_fields(self._h)
        v = np.zeros(nf * nn)
        d = np.zeros(3 * nf * nn)
        g = np.zeros(3 * nf * nn)
        if lib().oracle_get_contact(self._h, _d(v), _d(d), _d(g)) != 0:
            raise RuntimeError("no contact extrapolations in single-material mode")
        return v.reshape(nf, nn), d.reshape(nf, nn, 3).transpose(0, 2, 1), g.reshape(nf, nn, 3).transpose(0, 2, 1)

    def grid_values(self):
        """field 0: mass [nn], momentum and force [3, nn], point count [nn]"""
        m, pk, f, cnt = self.grid_fields()
        return m[0], pk[0], f[0], cnt[0]

    def counters(self):
        c = np.zeros(3, dtype=np.int64)
        lib().oracle_counters(self._h, c.ctypes.data_as(_PL))
        return {"steps": int(c[0]), "leftGrid": int(c[1]), "nan": int(c[2])}

    def regression_normal(self, node: int, field: int, logistic: bool):
        """(unit normal, separation, list length) of a regression normal at `node` for material
        field `field` against the other particles listed there (normals methods 5 and 6)"""
        nv = np.zeros(3)
        d = np.zeros(1)
        k = lib().oracle_regression_normal(self._h, int(node), int(field), int(logistic), _d(nv), _d(d))
        return nv, float(d[0]), int(k)

    def deformed_radius(self, p: int, n):
        """distance from particle p's center to its deformed domain's surface along unit vector n"""
        v = np.ascontiguousarray(n, dtype=np.float64)
        return float(lib().oracle_deformed_radius(self._h, int(p), _d(v)))

    def __del__(self):
        try:
            if self._h:
                lib().oracle_destroy(self._h)
        except Exception:
            pass


def shape(grid, shape_kind, elem, pos, lp, deriv=True):
    """Nodes and shape functions of one particle: (nodes, N, dNdx, dNdy, dNdz)."""
    nodes = np.zeros(64, dtype=np.int64)
    fn, dx, dy, dz = (np.zeros(64) for _ in range(4))
    p = np.ascontiguousarray(pos, dtype=np.float64)
    l = np.ascontiguousarray(lp, dtype=np.float64)
    k = lib().oracle_shape(ctypes.byref(grid), shape_kind, elem, _d(p), _d(l), 1 if deriv else 0,
                           nodes.ctypes.data_as(_PL), _d(fn), _d(dx), _d(dy), _d(dz))
    return nodes[:k], fn[:k], dx[:k], dy[:k], dz[:k]


def constitutive(analysis, mat, du, state):
    """One constitutive update of one particle: du is 3x3, state is a NUM_FIELDS vector."""
    s = np.ascontiguousarray(state, dtype=np.float64).copy()
    d = np.ascontiguousarray(du, dtype=np.float64).reshape(9)
    lib().oracle_constitutive(analysis, ctypes.byref(mat), _d(d), _d(s))
    return s


def cpdi_shape(grid, elem, state, deriv=True):
    """lCPDI nodes and functions of one particle whose pos, lp and F are in the state vector."""
    nodes = np.zeros(64, dtype=np.int64)
    fn, dx, dy, dz = (np.zeros(64) for _ in range(4))
    st = np.ascontiguousarray(state, dtype=np.float64)
    k = lib().oracle_cpdi_shape(ctypes.byref(grid), elem, _d(st), 1 if deriv else 0,
                                nodes.ctypes.data_as(_PL), _d(fn), _d(dx), _d(dy), _d(dz))
    if k < 0:
        raise RuntimeError("a CPDI corner left the grid")
    return nodes[:k], fn[:k], dx[:k], dy[:k], dz[:k]


def contact_node(law, normals, call_type, dt, mass, npts, vol, disp, grad, pk, ftot, pkc=None):
    """One lumped contact pass on a single node of len(mass) material velocity fields. mass, npts
    and vol are per field; disp (mass-weighted displacement), grad (volume gradient), pk, ftot and
    pkc are [nf, 3]. Returns the updated (pk, ftot, pkc)."""
    nf = len(mass)
    m = np.ascontiguousarray(mass, dtype=np.float64)
    n = np.ascontiguousarray(npts, dtype=np.int32)
    v = np.ascontiguousarray(vol, dtype=np.float64)
    d = np.ascontiguousarray(disp, dtype=np.float64).reshape(nf * 3)
    g = np.ascontiguousarray(grad, dtype=np.float64).reshape(nf * 3)
    p = np.array(pk, dtype=np.float64).reshape(nf * 3)
    f = np.array(ftot, dtype=np.float64).reshape(nf * 3)
    c = np.zeros(nf * 3) if pkc is None else np.array(pkc, dtype=np.float64).reshape(nf * 3)
    if lib().oracle_contact_node(ctypes.byref(law), normals, call_type, nf, dt, _d(m), n.ctypes.data_as(_PI),
                                 _d(v), _d(d), _d(g), _d(p), _d(f), _d(c)) != 0:
        raise RuntimeError("bad contact node input")
    return p.reshape(nf, 3), f.reshape(nf, 3), c.reshape(nf, 3)
