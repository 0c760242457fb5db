"""Traction conditions (MatPtTractionBC) on the HIP path against the oracle (relative L2 <= 1e-9 of
particle positions, velocities and stresses; grid forces after the momentum update to 1e-9):
uGIMP, classic and lCPDI particles in 2D and 3D, every face and direction (axes, normal,
tangential), constant/linear/sine/cosine styles, tractions together with particle loads on the
same particles, multimaterial mode (the force goes to the particle's own material field), row
sorting (traction tables follow their particles), graphs and task by task."""
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-9

# elastic_block_2d(cells=16, block=8): particles at 4.25 .. 11.75 (jitter 0.05) on both axes
TRAC_2D = [
    ("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75, "tolerance": 0.3},
     [{"dir": 11, "face": 3, "stress": 20.}, {"dir": 12, "face": 3, "style": 3, "stress": 5., "ftime": 20.}]),
    ("BCLine", {"x1": 11.75, "y1": 4, "x2": 11.75, "y2": 12, "tolerance": 0.3},
     [{"dir": 1, "face": 2, "style": 2, "stress": 1000., "ftime": 0.}]),
    ("BCLine", {"x1": 4, "y1": 4.25, "x2": 12, "y2": 4.25, "tolerance": 0.3},
     [{"dir": 2, "face": 1, "style": 4, "stress": -10., "ftime": 10.}, {"dir": 11, "face": 1, "stress": 3.}]),
    ("BCLine", {"x1": 4.25, "y1": 4, "x2": 4.25, "y2": 12, "tolerance": 0.3},
     [{"dir": 12, "face": 4, "stress": 2., "style": 2, "ftime": 0.005}]),
]

# neohookean_block_3d(cells=16, lo=4, hi=12): particles at 4.25 .. 11.75 on every axis
TRAC_3D = [
    ("BCBox", {"xmin": 4, "xmax": 12, "ymin": 4, "ymax": 12, "zmin": 11.5, "zmax": 12},
     [{"dir": 11, "face": 6, "stress": 0.2}, {"dir": 1, "face": 6, "style": 3, "stress": 0.05, "ftime": 20.}]),
    ("BCBox", {"xmin": 11.5, "xmax": 12, "ymin": 4, "ymax": 12, "zmin": 4, "zmax": 12},
     [{"dir": 1, "face": 2, "style": 2, "stress": 5., "ftime": 0.}]),
    ("BCBox", {"xmin": 4, "xmax": 12, "ymin": 4, "ymax": 4.5, "zmin": 4, "zmax": 12},
     [{"dir": 11, "face": 1, "stress": 0.1}, {"dir": 3, "face": 1, "style": 4, "stress": 0.03, "ftime": 10.}]),
    ("BCBox", {"xmin": 4, "xmax": 4.5, "ymin": 4, "ymax": 12, "zmin": 4, "zmax": 12},
     [{"dir": 2, "face": 4, "stress": -0.05}]),
    ("BCBox", {"xmin": 4, "xmax": 12, "ymin": 11.5, "ymax": 12, "zmin": 4, "zmax": 12},
     [{"dir": 11, "face": 3, "stress": 0.08}]),
    ("BCBox", {"xmin": 4, "xmax": 12, "ymin": 4, "ymax": 12, "zmin": 4, "zmax": 4.5},
     [{"dir": 11, "face": 5, "stress": 0.15}]),
]


def _pair(xml):
    prob = nm.Problem.from_xml(xml)
    assert prob.particle_tractions(), "the input has tractions"
    return prob, nm.Simulation.from_problem(prob), orc.Oracle.from_problem(prob)


def _check(sim, ora, dim, label, tol=TOL):
    g, _ = sim.particles()
    r, _ = ora.particles()
    errs = compare(g, r, dim)
    print(label, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{label}: {k} relative L2 {v:.3e} > {tol}"


@pytest.mark.parametrize("shape", ["uGIMP", "Classic", "lCPDI"])
def test_tractions_2d(shape):
    print(f"case: 2D tractions, {shape}", flush=True)
    _, sim, ora = _pair(configs.traction_bcs(configs.elastic_block_2d(cells=16, block=8, vmax=2000., shape=shape),
                                             TRAC_2D))
    for n in (1, 19, 80):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 2, f"2D tractions {shape} +{n}")


@pytest.mark.parametrize("shape", ["uGIMP", "lCPDI"])
def test_tractions_3d_sorted_rows(shape):
    print(f"case: 3D neo-Hookean tractions on every face, {shape}, rows sorted every 5 steps", flush=True)
    xml = configs.traction_bcs(configs.neohookean_block_3d(cells=16, lo=4, hi=12, vx=2000., shape=shape), TRAC_3D)
    _, sim, ora = _pair(xml)
    sim.set_sort_interval(5)
    sim.set_graph_steps(8)
    for n in (1, 7, 32):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 3, f"3D tractions {shape} +{n}")


def test_tractions_with_loads_task_by_task_forces():
    print("case: 2D tractions and loads on the same particles, task by task", flush=True)
    loads = [("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75, "tolerance": 0.3},
              ["net", {"dir": 2, "style": 3, "load": 0.4, "ftime": 20.}])]
    xml = configs.particle_bcs(configs.elastic_block_2d(cells=16, block=8, vmax=2000.), loads)
    xml = configs.traction_bcs(xml, TRAC_2D)
    prob, sim, ora = _pair(xml)
    assert prob.particle_loads()
    for step in range(3):
        for task in range(10):
            sim.run_task(task)
            ora.run_task(task)
            if task == 6:
                sim.sync()
                m, pk, f = sim.grid_fields()
                rm, rpk, rf, _ = ora.grid_fields()
                e = rel_l2(f, rf)
                print("step", step, "ftot", f"{e:.3e}", flush=True)
                assert e <= TOL
    sim.step(40)
    ora.step(40)
    _check(sim, ora, 2, "tractions and loads")


def test_tractions_multimaterial_own_field():
    print("case: 2D contact blocks, tractions on each block", flush=True)
    xml = configs.contact_blocks_2d(cells=32, w=8, h=8, gap=1., vx=2000.)
    # block 1: x 7.5 .. 15.5, block 2: x 16.5 .. 24.5, both y 12 .. 20 (4 particles per cell)
    xml = configs.traction_bcs(xml, [
        ("BCLine", {"x1": 7, "y1": 19.75, "x2": 16, "y2": 19.75}, [{"dir": 11, "face": 3, "stress": 5.}]),
        ("BCLine", {"x1": 24.25, "y1": 11, "x2": 24.25, "y2": 21}, [{"dir": 1, "face": 2, "stress": -8.}]),
    ])
    _, sim, ora = _pair(xml)
    for n in (1, 30, 60):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 2, f"multimaterial tractions +{n}")


def test_removing_tractions_and_bad_input():
    print("case: tractions removed; bad conditions refused", flush=True)
    prob, sim, ora = _pair(configs.traction_bcs(configs.elastic_block_2d(cells=16, block=8), TRAC_2D))
    sim.step(5)
    ora.step(5)
    sim.set_particle_tractions([])
    ora.set_particle_tractions([])
    sim.step(5)
    ora.step(5)
    _check(sim, ora, 2, "tractions removed")
    for bad in (nm.ParticleTraction(particle=10 ** 8, dir=1, face=1, style=1, thickness=1.),
                nm.ParticleTraction(particle=0, dir=3, face=1, style=1, thickness=1.),
                nm.ParticleTraction(particle=0, dir=1, face=5, style=1, thickness=1.),
                nm.ParticleTraction(particle=0, dir=1, face=1, style=5, thickness=1.)):
        with pytest.raises(nm.MPMError):
            sim.set_particle_tractions([bad])
