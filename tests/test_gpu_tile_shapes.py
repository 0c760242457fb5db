"""Tile shape 1 (8x4x4 cells per tile in 3D, 16x8 in 2D; shape 0 is 4x4x4 and 8x8) against the
oracle (relative L2 <= 1e-9 of positions, displacements, velocities and stresses): elastic and J2
blocks in 2D and 3D, lCPDI, multimaterial contact with regression normals, grid velocity
conditions with particle loads and tractions, XPIC; the shape switched between steps (the
particles binned again) in both directions; a shape-1 run repeats bit for bit and its graph replay
equals its eager steps; the tuner picks a shape and leaves the state where it was."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _check(sim, ora, dim, label, tol=TOL):
    g, _ = sim.particles()
    r, _ = ora.particles()
    errs = compare(g, r, dim)
    print(label, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{label}: {k} relative L2 {v:.3e} > {tol}"


def _loaded_2d():
    xml = configs.grid_bcs(configs.elastic_block_2d(cells=24, block=12, vmax=3000.),
                           [("BCLine", {"x1": 6, "y1": 6, "x2": 18, "y2": 6}, [{"dir": 2, "vel": 0.}])])
    xml = configs.particle_bcs(xml, [("BCLine", {"x1": 6, "y1": 17.75, "x2": 18, "y2": 17.75, "tolerance": 0.3},
                                      ["net", {"dir": 1, "style": 2, "load": 0.5, "ftime": 0.}])])
    return configs.traction_bcs(xml, [("BCLine", {"x1": 17.75, "y1": 6, "x2": 17.75, "y2": 18, "tolerance": 0.3},
                                       [{"dir": 11, "face": 2, "stress": 20.}])])


# label, dimension, input, steps at which to compare. The contact blocks translate rigidly until
# they meet (about step 165): their stresses are rounding noise before, so they are compared after.
CASES = [
    ("2D J2 block", 2, lambda: configs.j2_block_2d(cells=40, block=20), (1, 14, 54)),
    ("2D lCPDI elastic", 2, lambda: configs.elastic_block_2d(cells=24, block=12, vmax=3000., shape="lCPDI"), (1, 14, 54)),
    ("2D held, loads and tractions", 2, _loaded_2d, (1, 14, 54)),
    ("2D contact, linear regression normals", 2, lambda: configs.contact_blocks_2d(normals=5, vy=3000.), (200, 300)),
    ("2D XPIC(3)", 2, lambda: configs.periodic_xpic(configs.elastic_block_2d(cells=24, block=12, vmax=3000.), order=3),
     (1, 14, 54)),
    ("3D J2 block", 3, lambda: configs.j2_block_3d(cells=20, lo=4, hi=15, vx=2e5), (1, 14, 54)),
    ("3D neo-Hookean lCPDI", 3, lambda: configs.neohookean_block_3d(cells=20, lo=5, hi=14, vx=2000., shape="lCPDI"),
     (1, 14, 54)),
    ("3D contact cylinders, frictionless", 3,
     lambda: configs.contact_cylinders_3d(cells=24, radius=5., length=8., gap=1., shape="uGIMP", law_type=61,
                                          friction=0.), (40, 100)),
]


@pytest.mark.parametrize("label,dim,make,checks", CASES, ids=[c[0] for c in CASES])
def test_shape_one_against_oracle(label, dim, make, checks):
    print("case: tile shape 1,", label, flush=True)
    prob = nm.Problem.from_xml(make())
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    assert sim.tile_shape == 0
    sim.set_tile_shape(1)
    assert sim.tile_shape == 1
    sim.set_sort_interval(7)
    done = 0
    for target in checks:
        sim.step(target - done)
        ora.step(target - done)
        done = target
        _check(sim, ora, dim, f"shape 1 {label} step {target}")


@pytest.mark.parametrize("dim", [2, 3])
def test_switching_shapes_between_steps(dim):
    print(f"case: {dim}D shape switched 0 -> 1 -> 0 between steps", flush=True)
    xml = configs.j2_block_2d(cells=32, block=16) if dim == 2 else configs.j2_block_3d(cells=16, lo=4, hi=12, vx=2e5)
    prob = nm.Problem.from_xml(xml)
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    for shape, n in ((0, 10), (1, 11), (0, 9), (1, 10)):
        sim.set_tile_shape(shape)
        sim.step(n)
        ora.step(n)
        _check(sim, ora, dim, f"{dim}D after {n} steps in shape {shape}")
    assert sim.stats().steps == 40


def test_shape_one_is_deterministic_and_graph_equals_eager():
    print("case: shape 1 repeats bit for bit; graphs equal eager steps", flush=True)
    prob = nm.Problem.from_xml(configs.j2_block_3d(cells=16, lo=4, hi=12, vx=2e5))
    runs = []
    for graph in (8, None, 8):
        sim = nm.Simulation.from_problem(prob)
        sim.set_tile_shape(1)
        if graph:
            sim.set_graph_steps(graph)
            sim.step(24)
        else:
            for _ in range(24):
                sim.step(1)
        runs.append(sim.particles()[0])
    assert np.array_equal(runs[0], runs[2]), "a repeated shape-1 run differs"
    assert np.array_equal(runs[0], runs[1]), "graph replay differs from eager steps"


def test_tuner_picks_a_shape_and_keeps_the_state():
    print("case: tuner over tile shapes and graph lengths", flush=True)
    prob = nm.Problem.from_xml(configs.j2_block_2d(cells=40, block=20))
    a = nm.Simulation.from_problem(prob)
    b = nm.Simulation.from_problem(prob)
    a.step(5)
    b.step(5)
    g, ms = a.autotune(trial_steps=8)
    print("tuner:", a.tile_shape, g, f"{ms:.4f} ms/step", flush=True)
    assert a.tile_shape in (0, 1) and g in (2, 4, 8, 16, 32) and ms > 0
    assert a.stats().steps == 5
    assert np.array_equal(a.particles()[0], b.particles()[0])
    b.set_tile_shape(a.tile_shape)
    a.step(30)
    b.step(30)
    assert np.array_equal(a.particles()[0], b.particles()[0])


def test_bad_shape_and_half_step_refused():
    print("case: shape requests the context refuses", flush=True)
    sim = nm.Simulation.from_problem(nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8)))
    for bad in (-1, 2):
        with pytest.raises(nm.MPMError, match="tile shape"):
            sim.set_tile_shape(bad)
    sim.run_task(0)
    with pytest.raises(nm.MPMError, match="half done"):
        sim.set_tile_shape(1)
