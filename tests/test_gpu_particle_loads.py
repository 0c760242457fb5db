"""External particle loads on the HIP path against the oracle (relative L2 <= 1e-9 of particle
positions, velocities and stresses; grid forces after the force task to 1e-9): uGIMP, classic
and lCPDI scatters, constant/linear/sine/cosine loads, loads with grid velocity conditions, row
sorting (loads follow their particles through the permutations), graphs and task by task."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-9

LOADS_2D = [
    ("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75, "tolerance": 0.3},
     ["net", {"dir": 2, "style": 3, "load": 0.4, "ftime": 20.}, {"dir": 1, "style": 2, "load": 0.2, "ftime": 0.}]),
    ("BCLine", {"x1": 4, "y1": 4.25, "x2": 12, "y2": 4.25, "tolerance": 0.3},
     [{"dir": 1, "load": -0.002}, {"dir": 2, "style": 4, "load": 0.001, "ftime": 10.}]),
]


def _pair(xml):
    prob = nm.Problem.from_xml(xml)
    assert prob.particle_loads(), "the input has loads"
    return prob, nm.Simulation.from_problem(prob), orc.Oracle.from_problem(prob)


def _check(sim, ora, dim, label, tol=TOL):
    g, _ = sim.particles()
    r, _ = ora.particles()
    errs = compare(g, r, dim)
    print(label, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{label}: {k} relative L2 {v:.3e} > {tol}"


@pytest.mark.parametrize("shape", ["uGIMP", "Classic", "lCPDI"])
def test_loads_2d(shape):
    print(f"case: 2D loads, {shape}", flush=True)
    _, sim, ora = _pair(configs.particle_bcs(configs.elastic_block_2d(cells=16, block=8, shape=shape), LOADS_2D))
    for n in (1, 19, 80):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 2, f"2D loads {shape} +{n}")


@pytest.mark.parametrize("shape", ["uGIMP", "lCPDI"])
def test_loads_and_conditions_3d_sorted_rows(shape):
    print(f"case: 3D J2 clamped and pulled, {shape}, rows sorted every 5 steps", flush=True)
    xml = configs.j2_block_3d(cells=12, lo=3, hi=9, vx=5e4, shape=shape)
    xml = configs.grid_bcs(xml, [("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 2.9, "zmax": 3.1},
                                  [{"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}, {"dir": 3, "vel": 0.}])])
    xml = configs.particle_bcs(xml, [("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 8.5, "zmax": 9},
                                      ["net", {"dir": 3, "load": 4000.}, {"dir": 1, "style": 3, "load": 0.5, "ftime": 30.}])])
    _, sim, ora = _pair(xml)
    sim.set_sort_interval(5)
    sim.set_graph_steps(8)
    for n in (1, 7, 32):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 3, f"3D loads {shape} +{n}")


def test_loads_task_by_task_forces():
    print("case: 2D loads task by task", flush=True)
    _, sim, ora = _pair(configs.particle_bcs(configs.elastic_block_2d(cells=16, block=8), LOADS_2D))
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
    _check(sim, ora, 2, "loads task by task")


def test_removing_and_reuploading_clears_loads():
    print("case: loads removed, then particles uploaded again", flush=True)
    prob, sim, ora = _pair(configs.particle_bcs(configs.elastic_block_2d(cells=16, block=8), LOADS_2D))
    sim.step(5)
    ora.step(5)
    sim.set_particle_loads([])
    ora.set_particle_loads([])
    sim.step(5)
    ora.step(5)
    _check(sim, ora, 2, "loads removed")
    d, i = prob.particles()
    sim.upload(d, i)
    free = orc.Oracle.from_problem(nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8)))
    sim.step(10)
    free.step(10)
    _check(sim, free, 2, "re-uploaded without loads")
    bad = nm.ParticleLoad(particle=10 ** 8, dir=1, style=1, value=0., ftime=0.)
    with pytest.raises(nm.MPMError):
        sim.set_particle_loads([bad])
