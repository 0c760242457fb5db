"""Grid velocity conditions on the HIP path against the oracle (relative L2 <= 1e-9 of particle
positions, velocities and stresses; grid values after each task to 1e-9 of each array's norm):
2D lines and 3D boxes with constant, linear, sine and cosine styles, conditions that switch on
inside captured multi-step graphs, multimaterial contact, and the task-by-task path."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-9

SHAPES_2D = [
    ("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 0.}, {"dir": 1, "vel": 150.}]),
    ("BCLine", {"x1": 4, "y1": 12, "x2": 12, "y2": 12}, [{"dir": 2, "style": 2, "vel": -2.}]),
    ("BCLine", {"x1": 12, "y1": 4, "x2": 12, "y2": 12}, [{"dir": 1, "style": 3, "vel": 300., "time": 5.}]),
]


def _pair(xml):
    prob = nm.Problem.from_xml(xml)
    assert prob.velocity_bcs(), "the input has velocity conditions"
    return prob, nm.Simulation.from_problem(prob), orc.Oracle.from_problem(prob)


def _check(sim, ora, dim, label, tol=TOL):
    g, _ = sim.particles()
    r, _ = ora.particles()
    errs = compare(g, r, dim)
    print(label, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{label}: {k} relative L2 {v:.3e} > {tol}"


def _check_grid(sim, ora, label, tol=TOL):
    m, pk, f = sim.grid_fields()
    rm, rpk, rf, _ = ora.grid_fields()
    for name, a, b in (("mass", m, rm), ("pk", pk, rpk), ("ftot", f, rf)):
        e = rel_l2(a, b)
        assert e <= tol, f"{label}: grid {name} relative L2 {e:.3e}"


def test_lines_2d_graph_steps():
    print("case: 2D lines, constant/linear/sine, 200 steps", flush=True)
    _, sim, ora = _pair(configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8), SHAPES_2D))
    for n in (1, 9, 40, 150):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 2, f"2D lines +{n}")
    _check_grid(sim, ora, "2D lines end")


def test_boxes_3d_j2():
    print("case: 3D J2 block, clamped bottom box and a cosine side box", flush=True)
    xml = configs.grid_bcs(configs.j2_block_3d(cells=16, lo=4, hi=12, vx=2e5), [
        ("BCBox", {"xmin": 4, "xmax": 12, "ymin": 4, "ymax": 12, "zmin": 3.9, "zmax": 4.1},
         [{"dir": 3, "vel": 0.}, {"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}]),
        ("BCBox", {"xmin": 11.9, "xmax": 12.1, "ymin": 4, "ymax": 12, "zmin": 6, "zmax": 12},
         [{"dir": 1, "style": 4, "vel": -3000., "time": 50.}]),
    ])
    _, sim, ora = _pair(xml)
    for n in (1, 19, 40):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 3, f"3D boxes +{n}")
    d, _ = sim.particles()
    assert np.any(d[nm.F["hist0"]] > 0.), "the block yields"


def test_activation_inside_a_captured_graph():
    # constant and linear conditions switching on after 13 steps, run in 16-step graphs: the
    # device's step start time must advance inside the graph
    base = configs.elastic_block_2d(cells=16, block=8)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    t_ms = 12.5 * run.timestep * 1e3
    xml = configs.grid_bcs(base, [
        ("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 80., "time": t_ms}]),
        ("BCLine", {"x1": 4, "y1": 8, "x2": 12, "y2": 8}, [{"dir": 1, "style": 2, "vel": 40., "time": t_ms}]),
    ])
    print("case: activation at step 13 inside 16-step graphs", flush=True)
    _, sim, ora = _pair(xml)
    sim.set_graph_steps(16)
    sim.step(32)
    ora.step(32)
    _check(sim, ora, 2, "activation in graphs")
    assert sim.stats().steps == 32


def test_contact_blocks_with_conditions():
    # frictionless law: the lines hold both blocks' edges at one vertical velocity, so the relative
    # tangential momentum at the interface nodes is zero up to rounding, and Coulomb friction would
    # normalise that rounding residue into a unit tangent (CoulombFriction.cpp:202-206) on one side
    # and not on the other
    print("case: 2D contact blocks with a moving top line", flush=True)
    base = configs.contact_blocks_2d(cells=32, w=8, h=8, gap=1.0, friction=0.0)
    xml = configs.grid_bcs(base, [
        ("BCLine", {"x1": 7, "y1": 20, "x2": 25, "y2": 20}, [{"dir": 2, "vel": -400.}]),
        ("BCLine", {"x1": 7, "y1": 12, "x2": 25, "y2": 12}, [{"dir": 2, "vel": 0.}]),
    ])
    _, sim, ora = _pair(xml)
    for n in (1, 29, 70):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, 2, f"contact + conditions +{n}")


def test_task_by_task_grid_values():
    print("case: 2D lines task by task", flush=True)
    _, sim, ora = _pair(configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8), SHAPES_2D))
    for step in range(3):
        for task in range(10):
            sim.run_task(task)
            ora.run_task(task)
            if task in (1, 2, 6, 8):     # the device fuses POST_FORCES into UPDATE_MOMENTA
                sim.sync()
                _check_grid(sim, ora, f"step {step} task {task}")
    _check(sim, ora, 2, "task by task")


def test_autotune_and_reset_keep_the_condition_time():
    print("case: autotune with time-dependent conditions", flush=True)
    _, sim, ora = _pair(configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8), SHAPES_2D))
    sim.step(3)
    ora.step(3)
    sim.autotune(8)
    sim.step(37)
    ora.step(37)
    _check(sim, ora, 2, "after autotune")
    # removing the conditions returns the free path
    sim.set_velocity_bcs([])
    ora.set_velocity_bcs([])
    sim.step(10)
    ora.step(10)
    _check(sim, ora, 2, "conditions removed")


def test_batch_refuses_conditions():
    prob = nm.Problem.from_xml(configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8), SHAPES_2D))
    with pytest.raises(nm.MPMError):
        nm.Batch([prob])
    sim = nm.Simulation.from_problem(nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8)))
    bad = nm.VelocityBC(node=10 ** 9, dir=1, style=1, value=0., ftime=0.)
    with pytest.raises(nm.MPMError):
        sim.set_velocity_bcs([bad])
