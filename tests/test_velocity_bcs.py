"""Grid velocity conditions: <GridBCs> node sets and values as the reference's reader builds them
(Read_MPM/Generators.cpp:831-1010, LineController.cpp:33-105, BoxController.cpp:20-74,
BoundaryCondition.cpp:72-206), and the oracle's four passes pinned by what they are meant to
achieve (NodalVelBC.cpp:321-400, MatVelocityField.cpp:490-575): after every pass but the force
pass, the controlled momentum component is mass * value; the force pass sets the force so the
momentum update lands on mass * value from the restored momenta; uncontrolled components and
inactive conditions change nothing."""
import math

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc


def _block2d(shapes, **kw):
    return configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8, **kw), shapes)


def _nodes_xy(g, bcs):
    nnx = g.nx + 1
    return [(g.xmin + (b.node % nnx) * g.dx, g.ymin + (b.node // nnx) * g.dy) for b in bcs]


def test_bcline_node_set_and_units():
    xml = _block2d([
        ("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 0.}]),
        ("BCLine", {"x1": "min", "y1": 12, "x2": "max", "y2": 12, "tolerance": "*0.5"},
         [{"dir": 1, "style": 3, "vel": 100., "time": 2.}, {"dir": 2, "style": 2, "vel": 5., "ftime": 0.5}]),
        ("BCLine", {"x1": 0, "y1": 0, "x2": 16, "y2": 16}, [{"dir": 12, "vel": 7.}]),
    ])
    prob = nm.Problem.from_xml(xml)
    g, _, _, _, _ = prob.describe()
    bcs = prob.velocity_bcs()
    xy = _nodes_xy(g, bcs)
    # 9 nodes on y = 4 from x = 4 to 12; 17 nodes on y = 12 over the mesh x range [0, 16] (the
    # min/max keywords are the limits before the extra GIMP elements), twice; 17 on the diagonal
    assert len(bcs) == 9 + 17 + 17 + 17
    assert all(y == 4. and 4. <= x <= 12. for x, y in xy[:9])
    assert [x for x, _ in xy[9:26]] == [float(k) for k in range(17)]
    assert all(y == 12. for _, y in xy[9:43])
    assert all(x == y for x, y in xy[43:])
    # list order: DisBC by DisBC, nodes in node order
    assert [b.node for b in bcs[9:26]] == [b.node for b in bcs[26:43]]
    assert all(bcs[k].node < bcs[k + 1].node for k in range(8))
    # units: sine frequency 1/ms -> 1/s, linear value per ms -> per s and start ms -> s,
    # skewed dir 12 at angle 0 -> x (Generators.cpp:951-960)
    b = bcs[9]
    assert (b.dir, b.style, b.value, b.ftime) == (1, 3, 100., 2000.)
    b = bcs[26]
    assert (b.dir, b.style, b.value, b.ftime) == (2, 2, 5000., 0.5e-3)
    assert all(b.dir == 1 and b.style == 1 and b.value == 7. and b.ftime == 0. for b in bcs[43:])


def test_bcline_tolerance():
    # LineController::ContainsPoint: distance to the line within tolerance, projection within the
    # segment (no tolerance past its ends); the default tolerance is 0.1 of the cell size
    def count(attrs):
        prob = nm.Problem.from_xml(_block2d([("BCLine", attrs, [{"dir": 1, "vel": 0.}])]))
        return len(prob.velocity_bcs())
    assert count({"x1": 4.05, "y1": 4, "x2": 4.05, "y2": 8}) == 5        # 0.05 mm off the x = 4 column
    assert count({"x1": 4.15, "y1": 4, "x2": 4.15, "y2": 8}) == 0        # beyond 0.1 mm
    assert count({"x1": 4.15, "y1": 4, "x2": 4.15, "y2": 8, "tolerance": 0.2}) == 5
    assert count({"x1": 4.15, "y1": 4, "x2": 4.15, "y2": 8, "tolerance": "*0.2"}) == 5
    assert count({"x1": 4, "y1": 4.05, "x2": 4, "y2": 7.95}) == 3        # ends not widened


def test_bcbox_node_set():
    xml = configs.grid_bcs(configs.j2_block_3d(cells=12, lo=3, hi=9), [
        ("BCBox", {"xmin": 9, "xmax": 3, "ymin": 3, "ymax": 9, "zmin": 2.9, "zmax": 3.1}, [{"dir": 3, "vel": 0.}]),
        ("BCBox", {"xmin": "max-", "xmax": "max+", "ymin": "min", "ymax": "max", "zmin": "min", "zmax": "max"},
         [{"dir": 1, "style": 4, "vel": -20., "time": 1.}]),
    ])
    prob = nm.Problem.from_xml(xml)
    g, _, _, _, _ = prob.describe()
    bcs = prob.velocity_bcs()
    nnx, nny = g.nx + 1, g.ny + 1
    pts = [(g.xmin + (b.node % nnx) * g.dx, g.ymin + (b.node // nnx % nny) * g.dy,
            g.zmin + (b.node // (nnx * nny)) * g.dz) for b in bcs]
    assert len(bcs) == 49 + 13 * 13 * 3
    assert all(3 <= x <= 9 and 3 <= y <= 9 and z == 3. for x, y, z in pts[:49])    # reversed x swapped
    assert sorted({x for x, _, _ in pts[49:]}) == [11., 12., 13.]                # max -/+ one cell
    assert bcs[49].style == 4 and bcs[49].ftime == 1000. and bcs[49].value == -20.


@pytest.mark.parametrize("shape, msg", [
    (("BCArc", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 1}]), "not on this path"),
    (("BCLine", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 3}]), "must be 1, 2, or 12"),
    (("BCLine", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 1, "id": 2}]), "'id'"),
    (("BCLine", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 1, "function": "t"}]), "function"),
    (("BCLine", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 12, "angle": 30.}]), "skewed"),
    (("BCLine", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 1, "style": 6}]), "styles"),
    (("BCBox", {"xmin": 0, "xmax": 1, "ymin": 0, "ymax": 1, "zmin": 0, "zmax": 1}, [{"dir": 1}]), "3D"),
])
def test_refusals_2d(shape, msg):
    with pytest.raises(nm.MPMError, match=msg):
        nm.Problem.from_xml(_block2d([shape]))


def test_refusals_3d():
    base = configs.j2_block_3d(cells=12, lo=3, hi=9)
    with pytest.raises(nm.MPMError, match="xmax cannot equal"):
        nm.Problem.from_xml(configs.grid_bcs(base, [
            ("BCBox", {"xmin": 3, "xmax": 3, "ymin": 0, "ymax": 1, "zmin": 0, "zmax": 1}, [{"dir": 1}])]))
    with pytest.raises(nm.MPMError, match="cylinder"):
        nm.Problem.from_xml(configs.grid_bcs(base, [
            ("BCBox", {"xmin": 0, "xmax": 1, "ymin": 0, "ymax": 1, "zmin": 0, "zmax": 1, "axis": "z"}, [{"dir": 1}])]))
    with pytest.raises(nm.MPMError, match="BCLine"):
        nm.Problem.from_xml(configs.grid_bcs(base, [
            ("BCLine", {"x1": 0, "y1": 0, "x2": 1, "y2": 1}, [{"dir": 1}])]))
    with pytest.raises(nm.MPMError, match="batches"):
        prob = nm.Problem.from_xml(configs.grid_bcs(base, [
            ("BCBox", {"xmin": 0, "xmax": 1, "ymin": 0, "ymax": 1, "zmin": 0, "zmax": 1}, [{"dir": 1}])]))
        nm.Batch([prob])


# ---- the oracle's passes ---------------------------------------------------------------------

def _bc_value(b, t):
    """BoundaryCondition::BCValue written from its definition: (acts, value) at time t"""
    if b.style in (1, 2):
        if t < b.ftime:
            return False, 0.
        return True, b.value if b.style == 1 else b.value * (t - b.ftime)
    return True, b.value * (math.sin(t * b.ftime) if b.style == 3 else math.cos(t * b.ftime))


def _controlled(bcs, t):
    """node, axis -> summed value of the acting conditions at time t"""
    out = {}
    for b in bcs:
        on, v = _bc_value(b, t)
        if on:
            key = (b.node, b.dir - 1)
            out[key] = out.get(key, 0.) + v
    return out


SHAPES_2D = [
    ("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 0.}, {"dir": 1, "vel": 150.}]),
    ("BCLine", {"x1": 4, "y1": 12, "x2": 12, "y2": 12}, [{"dir": 2, "style": 2, "vel": -2.}]),
    ("BCLine", {"x1": 12, "y1": 4, "x2": 12, "y2": 12}, [{"dir": 1, "style": 3, "vel": 300., "time": 5.}]),
]


def _oracle_pair(shapes, base=None):
    base = base or configs.elastic_block_2d(cells=16, block=8)
    with_bc = nm.Problem.from_xml(configs.grid_bcs(base, shapes))
    plain = nm.Problem.from_xml(base)
    return with_bc, orc.Oracle.from_problem(with_bc), orc.Oracle.from_problem(plain)


def _run_until(o, task):
    for t in range(task + 1):
        o.run_task(t)


def test_oracle_passes_land_on_the_prescribed_momenta():
    prob, o, free = _oracle_pair(SHAPES_2D)
    bcs = prob.velocity_bcs()
    _, _, run, _, _ = prob.describe()
    dt = run.timestep
    t = 0.
    for step in range(3):
        ctrl = _controlled(bcs, t)
        assert ctrl, "some conditions act"
        for task in range(10):           # MPMGPU_TASK_* ids
            o.run_task(task)
            free.run_task(task)
            m, pk, f, cnt = o.grid_values()
            on = [(n, a, v) for (n, a), v in ctrl.items() if cnt[n] > 0]
            if task in (2, 8):           # POST_EXTRAPOLATION, UPDATE_STRAINS_LAST: pk = m v
                for n, a, v in on:
                    assert pk[a, n] == pytest.approx(m[n] * v, rel=1e-13, abs=1e-300), (task, n, a)
            if task == 5:                # POST_FORCES: momenta restored, forces give m v after dt
                if step == 0:            # same particle state as the free run in the first step
                    _, pk_free, f_free, _ = free.grid_values()
                    assert np.array_equal(pk, pk_free)
                scale = max(np.abs(f).max(), np.abs(pk).max() / dt)
                for n, a, v in on:
                    assert f[a, n] == pytest.approx((m[n] * v - pk[a, n]) / dt, abs=1e-12 * scale), (n, a)
                pk0, f0 = pk.copy(), f.copy()
            if task == 6:                # UPDATE_MOMENTA: pk = m v, forces unchanged up to rounding
                for n, a, v in on:
                    assert pk[a, n] == pytest.approx(m[n] * v, rel=1e-13, abs=1e-300), (n, a)
                    assert f[a, n] == pytest.approx(f0[a, n], abs=1e-12 * scale), (n, a)
                free_mask = (cnt > 0)[None, :].repeat(3, axis=0)
                for (n, a) in ctrl:
                    free_mask[a, n] = False
                # every other component: pk + f dt with the forces the force pass left
                assert np.array_equal(pk[free_mask], (pk0 + f0 * dt)[free_mask])
                assert np.array_equal(f[free_mask], f0[free_mask])
        t += dt


def test_inactive_conditions_change_nothing():
    # constant and linear conditions starting after the run, on every node of the block
    shapes = [("BCLine", {"x1": 4, "y1": y, "x2": 12, "y2": y},
               [{"dir": 1, "vel": 100., "time": 1.}, {"dir": 2, "style": 2, "vel": 3., "time": 1.}])
              for y in range(4, 13)]
    _, o, free = _oracle_pair(shapes)
    o.step(20)
    free.step(20)
    a, _ = o.particles()
    b, _ = free.particles()
    assert np.array_equal(a, b)


def test_activation_time_follows_the_step_start_time():
    # a constant condition switches on at the first step whose start time reaches ftime
    prob0 = nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8))
    _, _, run, _, _ = prob0.describe()
    dt = run.timestep
    ftime_ms = 5.5 * dt * 1e3
    shapes = [("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 2, "vel": 40., "time": ftime_ms}])]
    prob, o, _ = _oracle_pair(shapes)
    bcs = prob.velocity_bcs()
    t = 0.
    for step in range(8):
        for task in range(10):
            o.run_task(task)
            if task == 8:
                m, pk, _, cnt = o.grid_values()
                on = t >= bcs[0].ftime
                for b in bcs:
                    if cnt[b.node] > 0:
                        if on:
                            assert pk[1, b.node] == pytest.approx(m[b.node] * 40., rel=1e-14)
                        elif step == 0:
                            assert pk[1, b.node] != pytest.approx(m[b.node] * 40., rel=1e-6)
        assert (t >= bcs[0].ftime) == (step >= 6)
        t += dt


def test_clamped_edge_holds_the_particles():
    # physics: with the bottom edge fixed in y and x, the particles next to it barely move while
    # the rest of the block swings
    shapes = [("BCLine", {"x1": 0, "y1": 4, "x2": 16, "y2": 4}, [{"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}])]
    prob, o, free = _oracle_pair(shapes)
    d0, _ = o.particles()
    o.step(100)
    free.step(100)
    d1, _ = o.particles()
    f1, _ = free.particles()
    near = d0[nm.F["pos_y"]] < 4.5
    mv_bc = np.abs(d1[nm.F["pos_x"]] - d0[nm.F["pos_x"]])[near].max()
    mv_free = np.abs(f1[nm.F["pos_x"]] - d0[nm.F["pos_x"]])[near].max()
    assert mv_bc < 0.2 * mv_free


def test_multimaterial_fields_each_take_the_condition():
    base = configs.contact_blocks_2d(cells=32, w=8, h=8, gap=1.0)
    shapes = [("BCLine", {"x1": 7, "y1": 12, "x2": 25, "y2": 12}, [{"dir": 2, "vel": 25.}])]
    prob, o, _ = _oracle_pair(shapes, base)
    bcs = prob.velocity_bcs()
    for task in range(7):
        o.run_task(task)
    m, pk, f, cnt = o.grid_fields()
    hits = 0
    for b in bcs:
        for fl in range(m.shape[0]):
            if cnt[fl, b.node] > 0:
                assert pk[fl, 1, b.node] == m[fl, b.node] * 25.
                hits += 1
    assert hits > 0
