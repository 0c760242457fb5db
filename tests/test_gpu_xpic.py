"""XPIC(k) and FMPM(k) (the PeriodicXPIC custom task) on the HIP path against the oracle (relative
L2 <= 1e-9 of particle positions, displacements, velocities and stresses; grid forces after the
particle update to 1e-9): orders 1 to 5, uGIMP, classic and lCPDI particles in 2D and 3D, grid
velocity conditions (each increment's held components zeroed; XPIC's particle pass keeps the
reactions in the forces), damping, schedules by steps and by time (a schedule by steps runs in
graphs whose length is a multiple of its period, captured for the phase of each call's first
step; by time, step by step), rows sorted into tile order, and the
task-by-task path. Each case also checks that the run differs from FLIP, so the order was used."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-9

HELD_2D = [
    ("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 0.}, {"dir": 1, "vel": 150.}]),
    ("BCLine", {"x1": 12, "y1": 4, "x2": 12, "y2": 12}, [{"dir": 1, "style": 3, "vel": 300., "time": 5.}]),
]
HELD_3D = [
    ("BCBox", {"xmin": 4, "xmax": 12, "ymin": 4, "ymax": 12, "zmin": 3.9, "zmax": 4.1},
     [{"dir": 3, "vel": 0.}, {"dir": 1, "vel": 100.}]),
]


def _damped(xml):
    return xml.replace("  </MPMHeader>", "    <Damping>2000</Damping>\n    <PDamping>500.5</PDamping>\n  </MPMHeader>", 1)


def _check(sim, ora, dim, label, tol=TOL):
    g, _ = sim.particles()
    r, _ = ora.particles()
    errs = compare(g, r, dim)
    print(label, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{label}: {k} relative L2 {v:.3e} > {tol}"


def _differs_from_flip(sim, flip, dim, label):
    g, _ = sim.particles()
    r, _ = flip.particles()
    errs = compare(g, r, dim)
    print(label, "vs FLIP", {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    assert errs["vel"] > 1e-6, f"{label}: the run is FLIP's"


CASES_2D = [
    # order, fmpm, shape, velocity conditions, damping
    (2, False, "uGIMP", True, False),
    (4, False, "Classic", True, True),
    (5, False, "lCPDI", True, False),
    (1, False, "uGIMP", False, True),
    (3, True, "uGIMP", True, False),
    (2, True, "lCPDI", True, True),
    (1, True, "Classic", True, False),
]


@pytest.mark.parametrize("order,fmpm,shape,held,damp", CASES_2D)
def test_every_step_2d(order, fmpm, shape, held, damp):
    name = f"{'FMPM' if fmpm else 'XPIC'}({order}) 2D {shape}{' held' if held else ''}{' damped' if damp else ''}"
    print("case:", name, flush=True)
    base = configs.elastic_block_2d(cells=16, block=8, vmax=3000., shape=shape, jitter=0.1)
    if held:
        base = configs.grid_bcs(base, HELD_2D)
    if damp:
        base = _damped(base)
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=order, fmpm=fmpm, steps=1))
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    flip = orc.Oracle.from_problem(nm.Problem.from_xml(base))
    sim.set_graph_steps(8)
    for n in (1, 19, 40):
        sim.step(n)
        ora.step(n)
        flip.step(n)
        assert ora.xpic_order() == order
        _check(sim, ora, 2, f"{name} +{n}")
    _differs_from_flip(sim, flip, 2, name)


@pytest.mark.parametrize("order,fmpm,shape", [(3, False, "uGIMP"), (2, True, "lCPDI"), (4, True, "Classic")])
def test_every_step_3d_sorted_rows(order, fmpm, shape):
    name = f"{'FMPM' if fmpm else 'XPIC'}({order}) 3D neo-Hookean {shape}"
    print("case:", name, "rows sorted every 5 steps", flush=True)
    base = configs.grid_bcs(configs.neohookean_block_3d(cells=16, lo=4, hi=12, vx=2000., shape=shape), HELD_3D)
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=order, fmpm=fmpm, steps=1))
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    flip = orc.Oracle.from_problem(nm.Problem.from_xml(base))
    sim.set_sort_interval(5)
    for n in (1, 7, 24):
        sim.step(n)
        ora.step(n)
        flip.step(n)
        _check(sim, ora, 3, f"{name} +{n}")
    _differs_from_flip(sim, flip, 3, name)


def test_schedules_by_steps_and_time():
    print("case: XPIC(3) every 3rd step, FMPM(2) every 2.5 time steps and every half time step", flush=True)
    base = configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8, vmax=3000.), HELD_2D)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    dt_ms = run.timestep * 1e3
    for xml, orders in (
            (configs.periodic_xpic(base, order=3, steps=3), [0, 0, 3, 0, 0, 3]),
            (configs.periodic_xpic(base, order=2, fmpm=True, steps=None, time_ms=2.5 * dt_ms), [0, 0, 0, 2, 0, 0]),
            (configs.periodic_xpic(base, order=2, steps=None, time_ms=0.5 * dt_ms), [2, 0, 2, 2, 2, 2])):
        prob = nm.Problem.from_xml(xml)
        sim = nm.Simulation.from_problem(prob)
        ora = orc.Oracle.from_problem(prob)
        seen = []
        for _ in range(6):
            sim.step(1)
            ora.step(1)
            seen.append(ora.xpic_order())
        assert seen == orders
        _check(sim, ora, 2, f"schedule {orders} +6")
        sim.step(31)
        ora.step(31)
        _check(sim, ora, 2, f"schedule {orders} +37")


@pytest.mark.parametrize("order,fmpm,period,graph", [(3, False, 3, 6), (2, True, 2, 4), (2, False, 3, 8)])
def test_periodic_schedule_in_graphs(order, fmpm, period, graph):
    # every period-th step in graphs of `graph` steps: with graph a multiple of the period the steps
    # run in graphs captured for the phase of each call's first step (calls below start at several
    # phases); otherwise step by step
    name = f"{'FMPM' if fmpm else 'XPIC'}({order}) every {period} steps, {graph}-step graphs"
    print("case:", name, flush=True)
    base = configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8, vmax=3000.), HELD_2D)
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=order, fmpm=fmpm, steps=period))
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    sim.set_graph_steps(graph)
    done = 0
    for n in (1, 13, 30, 7, 25):
        sim.step(n)
        ora.step(n)
        done += n
        _check(sim, ora, 2, f"{name} step {done}")


@pytest.mark.parametrize("fmpm", [False, True])
def test_task_by_task_forces(fmpm):
    print(f"case: {'FMPM' if fmpm else 'XPIC'}(3) task by task, held nodes", flush=True)
    base = configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8, vmax=3000., shape="Classic"), HELD_2D)
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=3, fmpm=fmpm, steps=2))
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    for step in range(4):
        for task in range(10):
            sim.run_task(task)
            ora.run_task(task)
            if task in (6, 7, 8):
                sim.sync()
                m, pk, f = sim.grid_fields()
                rm, rpk, rf, _ = ora.grid_fields()
                for nm_, a, b in (("pk", pk, rpk), ("ftot", f, rf)):
                    e = rel_l2(a, b)
                    print("step", step, "task", task, nm_, f"{e:.3e}", flush=True)
                    assert e <= TOL, f"step {step} task {task}: grid {nm_} relative L2 {e:.3e}"
        _check(sim, ora, 2, f"task by task, step {step}")


def test_refusals():
    print("case: XPIC settings the path refuses", flush=True)
    prob = nm.Problem.from_xml(configs.periodic_xpic(configs.elastic_block_2d(cells=16, block=8), order=2))
    g, mats, run, _, _ = prob.describe()
    for field, value, msg in (("xpicOrder", -1, "order"), ("xpicOrder", 33, "order"),
                              ("xpicPeriodSteps", 0, "period")):
        bad = nm.RunParams.from_buffer_copy(run)
        setattr(bad, field, value)
        with pytest.raises(nm.MPMError, match=msg):
            nm.Simulation(g, mats, bad)
    bad = nm.RunParams.from_buffer_copy(run)
    bad.multiMaterial = 1
    bad.numFields = 2
    with pytest.raises(nm.MPMError):
        nm.Simulation(g, mats, bad)
