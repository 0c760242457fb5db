"""Python binding of the MI355X explicit MPM step (libnairnmpm_gpu.so).

The library's C ABI is declared in include/nairnmpm_gpu.h. This module mirrors the reference's
structure: a :class:`Problem` is what NairnMPM holds after reading an XML input file
(MPMReadHandler, Read_MPM/Generators.cpp) and preparing particles (NairnMPM::CMPreparations,
NairnMPM_Class/NairnMPM.cpp:135-168); a :class:`Simulation` is the device state the MPMTask list
works on (NairnMPM::MPMStep, NairnMPM.cpp:284-335). Errors are raised as :class:`MPMError`, the
counterpart of the reference's CommonException.

There is no CPU fallback: if the HIP library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.dirname(_HERE)
LIBRARY_PATH = os.path.join(_LIBDIR, "libnairnmpm_gpu.so")

# analysis types, shape functions, material types (include/nairnmpm_gpu.h)
PLANE_STRAIN, PLANE_STRESS, THREED = 10, 11, 12
SHAPE_LINEAR, SHAPE_UGIMP, SHAPE_LCPDI = 0, 1, 10
MAT_ISOTROPIC, MAT_MOONEY, MAT_ISOPLASTICITY, MAT_NEOHOOKEAN, MAT_CONTACT, MAT_COULOMB = 1, 8, 9, 28, 60, 61

FIELDS = [
    "pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "mass", "lp_x", "lp_y", "lp_z",
    "ep_xx", "ep_yy", "ep_zz", "ep_yz", "ep_xz", "ep_xy", "wr_yz", "wr_xz", "wr_xy",
    "sp_xx", "sp_yy", "sp_zz", "sp_yz", "sp_xz", "sp_xy", "pressure",
    "alt_xx", "alt_yy", "alt_zz", "alt_yz", "alt_xz", "alt_xy", "hist0", "hist1",
    "orig_x", "orig_y", "orig_z",
]
F = {name: i for i, name in enumerate(FIELDS)}
NUM_FIELDS = len(FIELDS)
IFIELDS = ["elem", "mat", "id", "left"]
I = {name: i for i, name in enumerate(IFIELDS)}
NUM_IFIELDS = len(IFIELDS)

TASKS = [
    "Initialize", "Extrapolate Mass and Momentum", "Post Extrapolation Tasks",
    "Update Strains First", "Extrapolate Grid Forces", "Post Force Extrapolation Tasks",
    "Update Momenta", "Update Particles", "Update Strains Last with Extrapolation", "Reset Elements",
]


class MPMError(RuntimeError):
    """Failure reported by the library (the reference's CommonException)."""


class GridDesc(ctypes.Structure):
    _fields_ = [("analysis", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("xmin", ctypes.c_double), ("ymin", ctypes.c_double),
                ("zmin", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("dz", ctypes.c_double), ("thickness", ctypes.c_double)]


class Material(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("field", ctypes.c_int32), ("rho", ctypes.c_double),
                ("E", ctypes.c_double), ("nu", ctypes.c_double), ("G", ctypes.c_double),
                ("K", ctypes.c_double), ("C", (ctypes.c_double * 6) * 6), ("Gred", ctypes.c_double),
                ("Kred", ctypes.c_double), ("yield_", ctypes.c_double), ("Ep", ctypes.c_double),
                ("yldred", ctypes.c_double), ("yldredMin", ctypes.c_double), ("Epred", ctypes.c_double),
                ("alphaMax", ctypes.c_double), ("Gsp", ctypes.c_double), ("Lamesp", ctypes.c_double),
                ("Ksp", ctypes.c_double), ("G1sp", ctypes.c_double), ("G2sp", ctypes.c_double),
                ("waveSpeed", ctypes.c_double), ("friction", ctypes.c_double),
                ("frictionStatic", ctypes.c_double), ("ujOption", ctypes.c_int32), ("usePDamping", ctypes.c_int32),
                ("pdamping", ctypes.c_double)]


class RunParams(ctypes.Structure):
    _fields_ = [("timestep", ctypes.c_double), ("strainFirst", ctypes.c_double),
                ("strainLast", ctypes.c_double), ("shape", ctypes.c_int32), ("mpmMethod", ctypes.c_int32),
                ("multiMaterial", ctypes.c_int32), ("contactLaw", ctypes.c_int32),
                ("normalsMethod", ctypes.c_int32), ("deleteLeaving", ctypes.c_int32),
                ("maxLeaving", ctypes.c_int32), ("numFields", ctypes.c_int32),
                ("gridDamping", ctypes.c_double), ("particleDamping", ctypes.c_double),
                ("gravity", ctypes.c_double * 3), ("contactNormal", ctypes.c_double * 3),
                ("xpicOrder", ctypes.c_int32), ("xpicFMPM", ctypes.c_int32), ("xpicPeriodSteps", ctypes.c_int32),
                ("xpicReserved", ctypes.c_int32), ("xpicPeriodTime", ctypes.c_double)]


class InstanceDesc(ctypes.Structure):
    _fields_ = [("grid", GridDesc), ("run", RunParams), ("firstMaterial", ctypes.c_int32),
                ("numMaterials", ctypes.c_int32)]


class ServerParams(ctypes.Structure):
    _fields_ = [("memoryBudget", ctypes.c_int64), ("maxResident", ctypes.c_int32), ("graphSteps", ctypes.c_int32),
                ("analysis", ctypes.c_int32), ("shape", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("stepsTotal", ctypes.c_int32), ("stepsDone", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("nparticles", ctypes.c_int64), ("footprint", ctypes.c_int64),
                ("submitted", ctypes.c_double), ("admitted", ctypes.c_double), ("finished", ctypes.c_double)]


REQ_QUEUED, REQ_RESIDENT, REQ_DONE, REQ_COLLECTED = 0, 1, 2, 3


class Stats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("nparticles", ctypes.c_int64), ("leftGrid", ctypes.c_int64),
                ("nanParticles", ctypes.c_int64), ("time", ctypes.c_double), ("lastStepMs", ctypes.c_double),
                ("stopped", ctypes.c_int32), ("pad", ctypes.c_int32)]


# why a run stopped (MpmStats.stopped)
STOP_TOO_MANY_LEFT, STOP_NAN, STOP_DELETE = 1, 2, 3


class VelocityBC(ctypes.Structure):
    """one grid velocity condition (MpmVelocityBC): node (0-based), dir 1..3, style 1..4,
    value (mm/s, linear: mm/s^2) and ftime (s, or 1/s for the sine and cosine styles)"""
    _fields_ = [("node", ctypes.c_int64), ("dir", ctypes.c_int32), ("style", ctypes.c_int32),
                ("value", ctypes.c_double), ("ftime", ctypes.c_double)]


class ParticleLoad(ctypes.Structure):
    """one external particle load (MpmParticleLoad): particle (upload index), dir 1..3, style 1..4,
    value (uN; linear: uN/s) and ftime (s, or 1/s for the sine and cosine styles)"""
    _fields_ = [("particle", ctypes.c_int64), ("dir", ctypes.c_int32), ("style", ctypes.c_int32),
                ("value", ctypes.c_double), ("ftime", ctypes.c_double)]


class ParticleTraction(ctypes.Structure):
    """one traction condition (MpmParticleTraction): particle (upload index), dir (1..3, 11 normal,
    12 tangential in 2D), face (1..4 in 2D, 1..6 in 3D), style 1..4, value (Pa; linear: Pa/s),
    ftime (s, or 1/s for the sine and cosine styles) and the particle's thickness (2D, mm)"""
    _fields_ = [("particle", ctypes.c_int64), ("dir", ctypes.c_int32), ("face", ctypes.c_int32),
                ("style", ctypes.c_int32), ("reserved", ctypes.c_int32), ("value", ctypes.c_double),
                ("ftime", ctypes.c_double), ("thickness", ctypes.c_double)]


_lib = None

_P = ctypes.c_void_p
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int32)

_SIGNATURES = {
    "mpmgpu_global_error": ([], ctypes.c_char_p),
    "mpmgpu_last_error": ([_P], ctypes.c_char_p),
    "mpmgpu_problem_create": ([ctypes.c_int32, ctypes.POINTER(_P)], ctypes.c_int),
    "mpmgpu_problem_set_grid": ([_P, _PD, _PD, _PI, ctypes.c_double], ctypes.c_int),
    "mpmgpu_problem_set": ([_P, ctypes.c_char_p, ctypes.c_double], ctypes.c_int),
    "mpmgpu_problem_add_material": ([_P, ctypes.c_int32, ctypes.c_char_p, _PI], ctypes.c_int),
    "mpmgpu_problem_material_property": ([_P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_double], ctypes.c_int),
    "mpmgpu_problem_add_body": ([_P, ctypes.c_int32, ctypes.c_char_p, _PD, ctypes.c_int32, _PD, ctypes.c_double], ctypes.c_int),
    "mpmgpu_problem_add_point": ([_P, ctypes.c_int32, _PD, _PD, _PD, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "mpmgpu_problem_from_xml": ([ctypes.c_char_p, ctypes.POINTER(_P)], ctypes.c_int),
    "mpmgpu_problem_describe": ([_P, ctypes.POINTER(GridDesc), ctypes.POINTER(Material), _PI,
                                 ctypes.POINTER(RunParams), ctypes.POINTER(ctypes.c_int64), _PD], ctypes.c_int),
    "mpmgpu_problem_particles": ([_P, _PD, _PI], ctypes.c_int),
    "mpmgpu_problem_archive": ([_P, ctypes.c_char_p, _PD, _PD], ctypes.c_int),
    "mpmgpu_problem_velocity_bcs": ([_P, ctypes.POINTER(VelocityBC), ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "mpmgpu_problem_particle_loads": ([_P, ctypes.POINTER(ParticleLoad), ctypes.POINTER(ctypes.c_int64)],
                                      ctypes.c_int),
    "mpmgpu_problem_particle_tractions": ([_P, ctypes.POINTER(ParticleTraction), ctypes.POINTER(ctypes.c_int64)],
                                          ctypes.c_int),
    "mpmgpu_problem_destroy": ([_P], None),
    "mpmgpu_archive_order": ([ctypes.c_char_p, ctypes.c_char_p], ctypes.c_int),
    "mpmgpu_archive_record_size": ([ctypes.c_char_p, ctypes.c_int32], ctypes.c_int32),
    "mpmgpu_archive_write": ([ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_double, ctypes.c_int64,
                              _PD, _PI, ctypes.POINTER(Material), ctypes.c_int32, _PD, ctypes.c_double],
                             ctypes.c_int),
    "mpmgpu_archive_error": ([], ctypes.c_char_p),
    "mpmgpu_write_archive": ([_P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, _PD, ctypes.c_double],
                             ctypes.c_int),
    "mpmgpu_create": ([ctypes.POINTER(GridDesc), ctypes.POINTER(Material), ctypes.c_int32,
                       ctypes.POINTER(RunParams), ctypes.POINTER(_P)], ctypes.c_int),
    "mpmgpu_upload_particles": ([_P, ctypes.c_int64, _PD, _PI], ctypes.c_int),
    "mpmgpu_set_velocity_bcs": ([_P, ctypes.c_int64, ctypes.POINTER(VelocityBC)], ctypes.c_int),
    "mpmgpu_set_particle_loads": ([_P, ctypes.c_int64, ctypes.POINTER(ParticleLoad)], ctypes.c_int),
    "mpmgpu_set_particle_tractions": ([_P, ctypes.c_int64, ctypes.POINTER(ParticleTraction)], ctypes.c_int),
    "mpmgpu_batch_create": ([ctypes.c_int32, ctypes.POINTER(InstanceDesc), ctypes.POINTER(Material), ctypes.c_int32,
                             ctypes.POINTER(_P)], ctypes.c_int),
    "mpmgpu_run_task": ([_P, ctypes.c_int32], ctypes.c_int),
    "mpmgpu_step": ([_P, ctypes.c_int32], ctypes.c_int),
    "mpmgpu_set_graph_steps": ([_P, ctypes.c_int32], ctypes.c_int),
    "mpmgpu_set_sort_interval": ([_P, ctypes.c_int32], ctypes.c_int),
    "mpmgpu_autotune": ([_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)],
                        ctypes.c_int),
    "mpmgpu_set_tile_shape": ([_P, ctypes.c_int32], ctypes.c_int),
    "mpmgpu_tile_shape": ([_P], ctypes.c_int32),
    "mpmgpu_sync": ([_P], ctypes.c_int),
    "mpmgpu_download_particles": ([_P, _PD, _PI], ctypes.c_int),
    "mpmgpu_query": ([_P, ctypes.POINTER(Stats)], ctypes.c_int),
    "mpmgpu_download_grid": ([_P, _PD, _PD, _PD], ctypes.c_int),
    "mpmgpu_download_contact": ([_P, _PD, _PD, _PD], ctypes.c_int),
    "mpmgpu_destroy": ([_P], None),
    "mpmgpu_set_profiling": ([_P, ctypes.c_int32], ctypes.c_int),
    "mpmgpu_kernel_times": ([_P, _PD, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "mpmgpu_kernel_group_name": ([ctypes.c_int32], ctypes.c_char_p),
    "mpmgpu_footprint": ([ctypes.POINTER(GridDesc), ctypes.POINTER(RunParams), ctypes.c_int64], ctypes.c_int64),
    "mpmgpu_server_create": ([ctypes.POINTER(ServerParams), ctypes.POINTER(_P)], ctypes.c_int),
    "mpmgpu_server_submit": ([_P, ctypes.POINTER(GridDesc), ctypes.POINTER(Material), ctypes.c_int32,
                              ctypes.POINTER(RunParams), ctypes.c_int64, _PD, _PI, ctypes.c_int32,
                              ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "mpmgpu_server_advance": ([_P, ctypes.c_int32, _PI], ctypes.c_int),
    "mpmgpu_server_start": ([_P], ctypes.c_int),
    "mpmgpu_server_stop": ([_P], ctypes.c_int),
    "mpmgpu_server_wait": ([_P, ctypes.c_int64, ctypes.c_double, _PI], ctypes.c_int),
    "mpmgpu_server_info": ([_P, ctypes.c_int64, ctypes.POINTER(RequestInfo)], ctypes.c_int),
    "mpmgpu_server_collect": ([_P, ctypes.c_int64, _PD, _PI], ctypes.c_int),
    "mpmgpu_server_counts": ([_P, _PI, _PI, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                              _PD, _PD], ctypes.c_int),
    "mpmgpu_server_clock": ([_P], ctypes.c_double),
    "mpmgpu_server_error": ([_P], ctypes.c_char_p),
    "mpmgpu_server_destroy": ([_P], None),
    "mpmgpu_solve_static": ([ctypes.POINTER(GridDesc), ctypes.POINTER(Material), ctypes.c_int32,
                             ctypes.POINTER(RunParams), ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_int32, ctypes.c_void_p], ctypes.c_int),
}

# bytes of the static device array behind mpmgpu_solve_static (MPMGPU_STATIC_SCRATCH_BYTES)
STATIC_SCRATCH_BYTES = 256 << 20


def solve_static(grid, mats, run, n: int, dfields_dev: int, ifields_dev: int, steps: int, stream: int = 0):
    """mpmgpu_solve_static: advance n particles held in device memory (addresses of the
    [NUM_FIELDS, n] float64 and [NUM_IFIELDS, n] int32 rows) by `steps` steps without allocating;
    stream is a hipStream_t address (0: the null stream)."""
    arr = (Material * max(len(mats), 1))(*mats)
    _check(library().mpmgpu_solve_static(ctypes.byref(grid), arr, len(mats), ctypes.byref(run), int(n),
                                         ctypes.c_void_p(int(dfields_dev)), ctypes.c_void_p(int(ifields_dev)),
                                         int(steps), ctypes.c_void_p(int(stream)) if stream else None))

KERNEL_GROUPS = ["k_p2g_mass", "k_node_mass", "k_usf", "k_p2g_force", "k_node_force", "k_g2p_p2g",
                 "k_node_mom2", "k_usl_reset", "binning", "contact"]


def library() -> ctypes.CDLL:
    """Load libnairnmpm_gpu.so from this package's directory (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise MPMError(f"{LIBRARY_PATH} is missing: build it with `make -C nairn-mpm-fea_amd` "
                           "(or __graft_entry__.build())")
        lib = ctypes.CDLL(LIBRARY_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def _check(status: int, ctx=None):
    if status != 0:
        lib = library()
        msg = lib.mpmgpu_last_error(ctx) if ctx else lib.mpmgpu_global_error()
        raise MPMError(msg.decode() if msg else "unknown error")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_PD)


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(_PI)


class Problem:
    """Input and generated particles (MPMReadHandler + PreliminaryParticleCalcs)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_xml(cls, text: str) -> "Problem":
        h = _P()
        _check(library().mpmgpu_problem_from_xml(text.encode(), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, analysis: int) -> "Problem":
        h = _P()
        _check(library().mpmgpu_problem_create(analysis, ctypes.byref(h)))
        return cls(h)

    def set_grid(self, ext6, cell3=(-1., -1., -1.), n3=(0, 0, 0), thickness=1.):
        e = np.ascontiguousarray(ext6, dtype=np.float64)
        c = np.ascontiguousarray(cell3, dtype=np.float64)
        n = np.ascontiguousarray(n3, dtype=np.int32)
        _check(library().mpmgpu_problem_set_grid(self._h, _dptr(e), _dptr(c), _iptr(n), thickness))

    def set(self, key: str, value: float):
        _check(library().mpmgpu_problem_set(self._h, key.encode(), float(value)))

    def add_material(self, mtype: int, name: str = "", **props) -> int:
        mid = ctypes.c_int32()
        _check(library().mpmgpu_problem_add_material(self._h, mtype, name.encode(), ctypes.byref(mid)))
        for k, v in props.items():
            _check(library().mpmgpu_problem_material_property(self._h, mid.value, k.encode(), float(v)))
        return mid.value

    def add_body(self, mat: int, shape: str, ext6, vel=(0., 0., 0.), axis: int = 0, thick: float = -1.):
        e = np.ascontiguousarray(ext6, dtype=np.float64)
        v = np.ascontiguousarray(vel, dtype=np.float64)
        _check(library().mpmgpu_problem_add_body(self._h, mat, shape.encode(), _dptr(e), axis, _dptr(v), thick))

    def add_point(self, mat: int, pos, vel=(0., 0., 0.), lp=(0.5, 0.5, 0.5), thick: float = -1., mass: float = -1.):
        p = np.ascontiguousarray(pos, dtype=np.float64)
        v = np.ascontiguousarray(vel, dtype=np.float64)
        l = np.ascontiguousarray(lp, dtype=np.float64)
        _check(library().mpmgpu_problem_add_point(self._h, mat, _dptr(p), _dptr(v), _dptr(l), thick, mass))

    def describe(self):
        g, mats, nm, run = GridDesc(), (Material * 16)(), ctypes.c_int32(), RunParams()
        n, tmax = ctypes.c_int64(), ctypes.c_double()
        _check(library().mpmgpu_problem_describe(self._h, ctypes.byref(g), mats, ctypes.byref(nm),
                                                 ctypes.byref(run), ctypes.byref(n), ctypes.byref(tmax)))
        return g, [mats[i] for i in range(nm.value)], run, n.value, tmax.value

    def particles(self):
        """Initial particle state in mpm[] order: (double fields [NUM_FIELDS, n], int fields [4, n])."""
        _, _, _, n, _ = self.describe()
        d = np.zeros((NUM_FIELDS, n), dtype=np.float64)
        i = np.zeros((NUM_IFIELDS, n), dtype=np.int32)
        _check(library().mpmgpu_problem_particles(self._h, _dptr(d), _iptr(i)))
        return d, i

    def archive_settings(self):
        """(MPM archive order as given, archive interval in s or None, per-particle thickness)
        from <MPMArchiveOrder>, <MPMArchive> and <ArchiveTime>"""
        _, _, _, n, _ = self.describe()
        order = ctypes.create_string_buffer(32)
        t = ctypes.c_double()
        thick = np.zeros(n, dtype=np.float64)
        _check(library().mpmgpu_problem_archive(self._h, order, ctypes.byref(t), _dptr(thick)))
        return order.value.decode(), (t.value if t.value >= 0. else None), thick

    def velocity_bcs(self):
        """the <GridBCs> velocity conditions in the reference's list order (VelocityBC list)"""
        n = ctypes.c_int64(0)
        _check(library().mpmgpu_problem_velocity_bcs(self._h, None, ctypes.byref(n)))
        arr = (VelocityBC * max(n.value, 1))()
        _check(library().mpmgpu_problem_velocity_bcs(self._h, arr, ctypes.byref(n)))
        return [arr[k] for k in range(n.value)]

    def particle_loads(self):
        """the <ParticleBCs> loads in the reference's list order (ParticleLoad list)"""
        n = ctypes.c_int64(0)
        _check(library().mpmgpu_problem_particle_loads(self._h, None, ctypes.byref(n)))
        arr = (ParticleLoad * max(n.value, 1))()
        _check(library().mpmgpu_problem_particle_loads(self._h, arr, ctypes.byref(n)))
        return [arr[k] for k in range(n.value)]

    def particle_tractions(self):
        """the <ParticleBCs> tractions in the reference's list order (ParticleTraction list)"""
        n = ctypes.c_int64(0)
        _check(library().mpmgpu_problem_particle_tractions(self._h, None, ctypes.byref(n)))
        arr = (ParticleTraction * max(n.value, 1))()
        _check(library().mpmgpu_problem_particle_tractions(self._h, arr, ctypes.byref(n)))
        return [arr[k] for k in range(n.value)]

    def close(self):
        if self._h:
            library().mpmgpu_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """Device state of one simulation; step() runs whole NairnMPM steps on the GPU."""

    def __init__(self, grid: GridDesc, mats, run: RunParams):
        arr = (Material * len(mats))(*mats)
        h = _P()
        _check(library().mpmgpu_create(ctypes.byref(grid), arr, len(mats), ctypes.byref(run), ctypes.byref(h)))
        self._h = h
        self.grid, self.run = grid, run
        self.numFields = run.numFields if run.multiMaterial and run.numFields > 1 else 1
        self.n = 0

    @classmethod
    def from_problem(cls, prob: Problem) -> "Simulation":
        g, mats, run, _, _ = prob.describe()
        sim = cls(g, mats, run)
        d, i = prob.particles()
        sim.upload(d, i)
        bcs = prob.velocity_bcs()
        if bcs:
            sim.set_velocity_bcs(bcs)
        loads = prob.particle_loads()
        if loads:
            sim.set_particle_loads(loads)
        tractions = prob.particle_tractions()
        if tractions:
            sim.set_particle_tractions(tractions)
        return sim

    def set_particle_loads(self, loads):
        """external particle loads (ParticleLoad list, reference list order, upload indices);
        [] removes them; a later upload removes them too"""
        arr = (ParticleLoad * max(len(loads), 1))(*loads)
        _check(library().mpmgpu_set_particle_loads(self._h, len(loads), arr), self._h)

    def set_particle_tractions(self, tractions):
        """traction conditions (ParticleTraction list, reference list order, upload indices);
        [] removes them; a later upload removes them too"""
        arr = (ParticleTraction * max(len(tractions), 1))(*tractions)
        _check(library().mpmgpu_set_particle_tractions(self._h, len(tractions), arr), self._h)

    def set_velocity_bcs(self, bcs):
        """grid velocity conditions (VelocityBC list, reference list order); [] removes them"""
        arr = (VelocityBC * max(len(bcs), 1))(*bcs)
        _check(library().mpmgpu_set_velocity_bcs(self._h, len(bcs), arr), self._h)

    def upload(self, dfields: np.ndarray, ifields: np.ndarray):
        d = np.ascontiguousarray(dfields, dtype=np.float64)
        i = np.ascontiguousarray(ifields, dtype=np.int32)
        n = d.shape[1]
        _check(library().mpmgpu_upload_particles(self._h, n, _dptr(d), _iptr(i)), self._h)
        self.n = n

    def step(self, nsteps: int = 1):
        _check(library().mpmgpu_step(self._h, int(nsteps)), self._h)

    def set_graph_steps(self, steps: int):
        _check(library().mpmgpu_set_graph_steps(self._h, int(steps)), self._h)

    def set_sort_interval(self, steps: int):
        """steps between sorts of the particle rows into tile order (0: only at upload)"""
        _check(library().mpmgpu_set_sort_interval(self._h, int(steps)), self._h)

    def set_tile_shape(self, shape: int):
        """tile shape 0 (4x4x4 cells in 3D, 8x8 in 2D) or 1 (8x4x4, 16x8); the particles are binned again"""
        _check(library().mpmgpu_set_tile_shape(self._h, int(shape)), self._h)

    @property
    def tile_shape(self) -> int:
        return int(library().mpmgpu_tile_shape(self._h))

    def autotune(self, trial_steps: int = 32):
        """pick the tile shape and graph length with the fastest steps (the shape stays in
        tile_shape); the state is left unchanged"""
        g, ms = ctypes.c_int32(), ctypes.c_double()
        _check(library().mpmgpu_autotune(self._h, int(trial_steps), ctypes.byref(g), ctypes.byref(ms)), self._h)
        return g.value, ms.value

    def run_task(self, task: int):
        _check(library().mpmgpu_run_task(self._h, int(task)), self._h)

    def sync(self):
        _check(library().mpmgpu_sync(self._h), self._h)

    def particles(self):
        d = np.zeros((NUM_FIELDS, self.n), dtype=np.float64)
        i = np.zeros((NUM_IFIELDS, self.n), dtype=np.int32)
        _check(library().mpmgpu_download_particles(self._h, _dptr(d), _iptr(i)), self._h)
        return d, i

    def write_archive(self, path, order: str = "iY", time_ms: float = 0., thickness=None,
                      default_thickness: float = 1.):
        """NairnMPM archive file (format version 6) of the current particles; read it back with
        nairnmpm.archive.read"""
        t = None if thickness is None else np.ascontiguousarray(thickness, dtype=np.float64)
        _check(library().mpmgpu_write_archive(self._h, os.fsencode(path), order.encode(), float(time_ms),
                                              None if t is None else _dptr(t), float(default_thickness)), self._h)

    def _nodes(self):
        g = self.grid
        return (g.nx + 1) * (g.ny + 1) * ((g.nz + 1) if g.analysis == THREED else 1)

    def grid_fields(self):
        """mass [nf, nn], momentum and force [nf, 3, nn] of every material velocity field"""
        nn, nf = self._nodes(), self.numFields
        m = np.zeros(nf * nn)
        pk = np.zeros(3 * nf * nn)
        f = np.zeros(3 * nf * nn)
        _check(library().mpmgpu_download_grid(self._h, _dptr(m), _dptr(pk), _dptr(f)), self._h)
        return m.reshape(nf, nn), pk.reshape(nf, 3, nn), f.reshape(nf, 3, nn)

    def grid_values(self):
        """field 0: mass [nn], momentum and force [3, nn]"""
        m, pk, f = self.grid_fields()
        return m[0], pk[0], f[0]

    def contact_values(self):
        """volume [nf, nn], displacement and volume gradient [nf, 3, nn] (multimaterial mode)"""
        nn, nf = self._nodes(), self.numFields
        v = np.zeros(nf * nn)
        d = np.zeros(3 * nf * nn)
        gr = np.zeros(3 * nf * nn)
        _check(library().mpmgpu_download_contact(self._h, _dptr(v), _dptr(d), _dptr(gr)), self._h)
        return v.reshape(nf, nn), d.reshape(nf, 3, nn), gr.reshape(nf, 3, nn)

    def set_profiling(self, on: bool):
        """HIP-event timing of every kernel group (eager launches while on)."""
        _check(library().mpmgpu_set_profiling(self._h, 1 if on else 0), self._h)

    def kernel_times(self):
        """{group: (total ms, launches)} since profiling was turned on"""
        ms = np.zeros(len(KERNEL_GROUPS))
        n = np.zeros(len(KERNEL_GROUPS), dtype=np.int64)
        _check(library().mpmgpu_kernel_times(self._h, _dptr(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
               self._h)
        return {g: (float(ms[k]), int(n[k])) for k, g in enumerate(KERNEL_GROUPS)}

    def stats(self) -> Stats:
        s = Stats()
        _check(library().mpmgpu_query(self._h, ctypes.byref(s)), self._h)
        return s

    def close(self):
        if getattr(self, "_h", None):
            library().mpmgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(Simulation):
    """Independent problems advanced together on one device context (mpmgpu_batch_create): each
    keeps its own grid, time step, materials and particles; step() moves every instance one of its
    own time steps with one launch per kernel."""

    def __init__(self, problems, particles=None):
        """problems: Problem objects; particles: optional list of (dfields, ifields) per problem that
        replaces the problem's own initial particles (same count and order)"""
        descs, mats, ds, iss = [], [], [], []
        for k, prob in enumerate(problems):
            if prob.velocity_bcs() or prob.particle_loads() or prob.particle_tractions():
                raise MPMError("grid velocity conditions, particle loads and tractions are for single simulations, "
                               "not batches")
            g, m, run, n, _ = prob.describe()
            d, i = prob.particles() if particles is None or particles[k] is None else particles[k]
            i = i.copy()
            i[IFIELDS.index("mat")] += len(mats)
            descs.append(InstanceDesc(g, run, len(mats), len(m)))
            mats.extend(m)
            ds.append(d)
            iss.append(i)
        arr = (InstanceDesc * len(descs))(*descs)
        marr = (Material * len(mats))(*mats)
        h = _P()
        _check(library().mpmgpu_batch_create(len(descs), arr, marr, len(mats), ctypes.byref(h)))
        self._h = h
        self.grid, self.run = descs[0].grid, descs[0].run
        self.numFields = 1
        self.descs = descs
        self.counts = [d.shape[1] for d in ds]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        self.node_counts = [(g.nx + 1) * (g.ny + 1) * ((g.nz + 1) if g.analysis == THREED else 1)
                            for g in (d.grid for d in descs)]
        self.node_offsets = np.concatenate([[0], np.cumsum([(c + 255) // 256 * 256 for c in self.node_counts])])
        self.n = 0
        self.upload(np.concatenate(ds, axis=1), np.concatenate(iss, axis=1))

    def _nodes(self):
        return int(self.node_offsets[-1])

    def instance_grid(self, k: int):
        """mass [nn], momentum and force [3, nn] of instance k's own grid"""
        m, pk, f = self.grid_fields()
        a, b = self.node_offsets[k], self.node_offsets[k] + self.node_counts[k]
        return m[0, a:b], pk[0, :, a:b], f[0, :, a:b]

    def instance_particles(self, k: int):
        """particles of instance k (its materials renumbered from 0, as in its own problem)"""
        d, i = self.particles()
        a, b = self.offsets[k], self.offsets[k + 1]
        i = i[:, a:b].copy()
        i[IFIELDS.index("mat")] -= self.descs[k].firstMaterial
        return d[:, a:b], i


def footprint(grid: GridDesc, run: RunParams, nparticles: int) -> int:
    """device bytes a single-material solve holds (mpmgpu_footprint)"""
    return int(library().mpmgpu_footprint(ctypes.byref(grid), ctypes.byref(run), int(nparticles)))


class Server:
    """Solve server (mpmgpu_server_*): queued solves are admitted by device footprint, stay
    resident and advance together as one batch; each result equals the solve run alone."""

    def __init__(self, analysis: int = THREED, shape: int = SHAPE_UGIMP, memory_budget: int = 0,
                 max_resident: int = 0, graph_steps: int = 8):
        p = ServerParams(int(memory_budget), int(max_resident), int(graph_steps), int(analysis), int(shape))
        h = _P()
        if library().mpmgpu_server_create(ctypes.byref(p), ctypes.byref(h)) != 0:
            raise MPMError(library().mpmgpu_server_error(None).decode())
        self._h = h
        self._n = {}

    def _ok(self, status: int):
        if status != 0:
            msg = library().mpmgpu_server_error(self._h)
            raise MPMError(msg.decode() if msg else "server error")

    def submit(self, grid, mats, run, dfields, ifields, steps: int) -> int:
        d = np.ascontiguousarray(dfields, dtype=np.float64)
        i = np.ascontiguousarray(ifields, dtype=np.int32)
        arr = (Material * len(mats))(*mats)
        rid = ctypes.c_int64()
        self._ok(library().mpmgpu_server_submit(self._h, ctypes.byref(grid), arr, len(mats), ctypes.byref(run),
                                                d.shape[1], _dptr(d), _iptr(i), int(steps), ctypes.byref(rid)))
        self._n[rid.value] = d.shape[1]
        return rid.value

    def submit_problem(self, prob: Problem, steps: int, particles=None) -> int:
        if prob.velocity_bcs() or prob.particle_loads() or prob.particle_tractions():
            raise MPMError("grid velocity conditions, particle loads and tractions are for single simulations, "
                           "not served solves")
        g, mats, run, _, _ = prob.describe()
        d, i = prob.particles() if particles is None else particles
        return self.submit(g, mats, run, d, i, steps)

    def advance(self, max_steps: int = 8) -> int:
        """one round of the request loop on this thread; returns the number of solves retired"""
        c = ctypes.c_int32()
        self._ok(library().mpmgpu_server_advance(self._h, int(max_steps), ctypes.byref(c)))
        return c.value

    def start(self):
        self._ok(library().mpmgpu_server_start(self._h))

    def stop(self):
        self._ok(library().mpmgpu_server_stop(self._h))

    def wait(self, rid: int, timeout: float = -1.0) -> bool:
        done = ctypes.c_int32()
        self._ok(library().mpmgpu_server_wait(self._h, int(rid), float(timeout), ctypes.byref(done)))
        return bool(done.value)

    def info(self, rid: int) -> dict:
        r = RequestInfo()
        self._ok(library().mpmgpu_server_info(self._h, int(rid), ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in RequestInfo._fields_ if k != "pad"}

    def collect(self, rid: int):
        n = self._n[rid]
        d = np.zeros((NUM_FIELDS, n), dtype=np.float64)
        i = np.zeros((NUM_IFIELDS, n), dtype=np.int32)
        self._ok(library().mpmgpu_server_collect(self._h, int(rid), _dptr(d), _iptr(i)))
        return d, i

    def counts(self) -> dict:
        q, r = ctypes.c_int32(), ctypes.c_int32()
        b, lay = ctypes.c_int64(), ctypes.c_int64()
        ls, ss = ctypes.c_double(), ctypes.c_double()
        self._ok(library().mpmgpu_server_counts(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b),
                                                ctypes.byref(lay), ctypes.byref(ls), ctypes.byref(ss)))
        return {"queued": q.value, "resident": r.value, "resident_bytes": b.value, "layouts": lay.value,
                "layout_seconds": ls.value, "step_seconds": ss.value}

    def clock(self) -> float:
        return float(library().mpmgpu_server_clock(self._h))

    def close(self):
        if getattr(self, "_h", None):
            library().mpmgpu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
