"""External particle loads: <ParticleBCs> particle sets and values as the reference's reader builds
them (Read_MPM/Generators.cpp:1012-1023,1105-1188; MatPtLoadBC.cpp:202-221), and the oracle's
loads pinned by Newton's second law: with partition of unity the internal forces cancel over the
grid, so the total particle momentum of a free body grows by the sum of the applied loads times
the time step, step by step, with each load's value at the step's start time."""
import math

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def _block2d(shapes, **kw):
    kw.setdefault("vmax", 0.)
    return configs.particle_bcs(configs.elastic_block_2d(cells=16, block=8, **kw), shapes)


def test_load_particle_sets_and_units():
    xml = _block2d([
        ("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75, "tolerance": 0.1},
         ["net", {"dir": 2, "load": 2.0}, "perParticle", {"dir": 1, "style": 2, "load": 0.5, "time": 1.}]),
        ("BCLine", {"x1": 4, "y1": 4.25, "x2": 12, "y2": 4.25, "tolerance": 0.1},
         [{"dir": 1, "style": 3, "load": -1.5, "ftime": 2.}]),
    ], jitter=0.)
    prob = nm.Problem.from_xml(xml)
    d, _ = prob.particles()
    loads = prob.particle_loads()
    top = np.nonzero(np.abs(d[F["pos_y"]] - 11.75) <= 0.1)[0]
    bottom = np.nonzero(np.abs(d[F["pos_y"]] - 4.25) <= 0.1)[0]
    assert len(top) == 16 and len(bottom) == 16
    assert len(loads) == 3 * 16
    # list order: load by load, particles in mpm[] order
    assert [l.particle for l in loads[:16]] == top.tolist()
    assert [l.particle for l in loads[16:32]] == top.tolist()
    assert [l.particle for l in loads[32:]] == bottom.tolist()
    # net: N -> uN shared by the 16 particles; per particle: linear per ms -> per s, ms -> s;
    # sine: frequency 1/ms -> 1/s
    assert all(l.dir == 2 and l.style == 1 and l.value == 1e6 * (2.0 / 16) for l in loads[:16])
    assert all(l.dir == 1 and l.style == 2 and l.value == 1e6 * 0.5 * 1e3 and l.ftime == 1e-3 for l in loads[16:32])
    assert all(l.style == 3 and l.value == -1.5e6 and l.ftime == 2000. for l in loads[32:])


def test_load_box_3d():
    xml = configs.particle_bcs(configs.j2_block_3d(cells=12, lo=3, hi=9), [
        ("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 8.5, "zmax": 9}, ["net", {"dir": 3, "load": -4.}]),
    ])
    prob = nm.Problem.from_xml(xml)
    d, _ = prob.particles()
    loads = prob.particle_loads()
    inside = np.nonzero(d[F["pos_z"]] >= 8.5)[0]
    assert [l.particle for l in loads] == inside.tolist()
    assert sum(l.value for l in loads) == pytest.approx(-4e6, rel=1e-14)


@pytest.mark.parametrize("child, msg", [
    ({"dir": 3, "load": 1.}, "must be 1 or 2"),
    ({"dir": 1, "style": 5, "load": 1.}, "styles"),
    ({"dir": 1, "function": "t"}, "function"),
])
def test_load_refusals(child, msg):
    with pytest.raises(nm.MPMError, match=msg):
        nm.Problem.from_xml(_block2d([("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [child])]))


def _momentum(o, dim):
    d, _ = o.particles()
    return np.array([np.sum(d[F["mass"]] * d[F["vel_" + a]]) for a in "xyz"[:dim]])


@pytest.mark.parametrize("style, load, ftime_ms", [(1, 2.0, 0.), (1, 2.0, None), (2, 3.0, 0.), (3, 1.5, 20.), (4, -1.0, 30.)])
def test_total_momentum_follows_the_loads(style, load, ftime_ms):
    # a block at rest under a net load on its top row, no other forces: sum mp v grows by the
    # load times dt each step (the internal forces sum to zero over the grid)
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0.)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    dt = run.timestep
    if ftime_ms is None:                       # a constant load that starts at the 6th step
        ftime_ms = 5.5 * dt * 1e3
    xml = configs.particle_bcs(base, [
        ("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75, "tolerance": 0.3},
         ["net", {"dir": 2, "style": style, "load": load, "ftime": ftime_ms}]),
    ])
    prob = nm.Problem.from_xml(xml)
    loads = prob.particle_loads()
    o = orc.Oracle.from_problem(prob)
    t = 0.
    expect = 0.
    for step in range(12):
        total = 0.
        for l in loads:                        # BCValue at the step's start time
            if l.style == 1:
                total += l.value if t >= l.ftime else 0.
            elif l.style == 2:
                total += l.value * (t - l.ftime) if t >= l.ftime else 0.
            elif l.style == 3:
                total += l.value * math.sin(t * l.ftime)
            else:
                total += l.value * math.cos(t * l.ftime)
        expect += total * dt
        o.step(1)
        px, py = _momentum(o, 2)
        scale = abs(load) * 1e6 * dt * (step + 1) * (1 + t * 1e3)
        assert px == pytest.approx(0., abs=1e-12 * scale)
        assert py == pytest.approx(expect, rel=1e-11, abs=1e-12 * scale), (step, py, expect)
        t += dt
    if style == 1 and ftime_ms > 0.:
        assert expect == pytest.approx(6 * loads[0].value * len(loads) * dt, rel=1e-12)


def test_loads_and_conditions_together_3d():
    # a J2 block clamped at the bottom and pulled up on its top face: oracle passes in sequence
    xml = configs.j2_block_3d(cells=12, lo=3, hi=9, vx=0.)
    xml = configs.grid_bcs(xml, [("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 2.9, "zmax": 3.1},
                                  [{"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}, {"dir": 3, "vel": 0.}])])
    xml = configs.particle_bcs(xml, [("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 8.5, "zmax": 9},
                                      ["net", {"dir": 3, "load": 50.}])])
    prob = nm.Problem.from_xml(xml)
    o = orc.Oracle.from_problem(prob)
    d0, _ = o.particles()
    o.step(30)
    d, _ = o.particles()
    top = d0[F["pos_z"]] > 8.5
    bottom = d0[F["pos_z"]] < 3.5
    dz = d[F["pos_z"]] - d0[F["pos_z"]]
    assert dz[top].mean() > 0. and abs(dz[bottom].mean()) < 0.1 * dz[top].mean()
