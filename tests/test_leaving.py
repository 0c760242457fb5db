"""Particles that leave the grid (ResetElementsTask.cpp:44-265) in the oracle and the XML reader:
<LeaveLimit> resolved as NairnMPM::CreateWarnings does (NairnMPM.cpp:814-848), particles pushed
back into their element, the run stopped once the limit is reached or at a NaN position, and
deletion into the reservoir refused with a message."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def _header(xml, extra):
    return xml.replace("  </MPMHeader>", extra + "  </MPMHeader>", 1)


def moving_block(limit=None, block=8, cells=16, speed=0.3, seed=5, vmax=0.):
    """A 2D elastic block (block x block cells, 4 particles per cell, so 4 block^2 particles)
    moving toward the right edge of the grid by `speed` cells per step: rigidly when vmax is 0,
    else with a shearing velocity profile of amplitude vmax (mm/s) on top of the drift."""
    xml = configs.elastic_block_2d(cells=cells, block=block, vmax=vmax, jitter=0.02, seed=seed)
    if limit is not None:
        xml = _header(xml, f"    <LeaveLimit>{limit}</LeaveLimit>\n")
    prob = nm.Problem.from_xml(xml)
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    d[F["vel_x"]] += speed * g.dx / run.timestep
    d[F["vel_y"]] += 0.1 * speed * g.dx / run.timestep
    return g, mats, run, d, i


@pytest.mark.parametrize("limit, n, want", [
    (None, 256, (0, 2)),       # default: 1% of 256 particles = 2
    (None, 1024, (0, 10)),
    (5, 256, (0, 5)),
    (1, 256, (0, 0)),          # 1: pushed back, never stops
    (-3, 256, (1, 3)),         # negative: deleted (not supported), stop after 3
    (None, 64, (1, 0)),        # fewer than 100 particles: 1% rounds to 0, which deletes
])
def test_leave_limit_resolution(limit, n, want):
    block = int(round(np.sqrt(n / 4)))
    _, _, run, _, _ = moving_block(limit, block=block)
    assert (run.deleteLeaving, run.maxLeaving) == want


def test_delete_limit_refused():
    xml = _header(configs.elastic_block_2d(cells=16, block=8), "    <DeleteLimit>5</DeleteLimit>\n")
    with pytest.raises(nm.MPMError, match="DeleteLimit"):
        nm.Problem.from_xml(xml)


def test_pushed_back_particles_stay_in_the_grid():
    g, mats, run, d, i = moving_block(limit=100000)
    o = orc.Oracle(g, mats, run, d, i)
    o.step(30)
    r, ri = o.particles()
    c = o.counters()
    assert c["leftGrid"] > 0 and c["leftGrid"] == int(ri[nm.I["left"]].sum())
    # every particle is inside an interior cell (not an edge cell of the padded grid)
    x = (r[F["pos_x"]] - g.xmin) / g.dx
    assert x.max() < g.nx - 1 and x.min() >= 1
    e = ri[nm.I["elem"]]
    assert np.array_equal(e % g.nx, np.floor(x).astype(int))


def test_run_stops_when_the_limit_is_reached():
    g, mats, run, d, i = moving_block(limit=5)
    o = orc.Oracle(g, mats, run, d, i)
    for k in range(60):
        try:
            o.step(1)
        except RuntimeError as e:
            assert "Too many particles have left the grid" in str(e)
            break
    else:
        pytest.fail("the run did not stop")
    c = o.counters()
    assert c["leftGrid"] == 5 and c["steps"] == k


def test_nan_position_stops_the_run():
    g, mats, run, d, i = moving_block(limit=100000, speed=0.01)
    d[F["vel_y"]][17] = np.nan
    o = orc.Oracle(g, mats, run, d, i)
    with pytest.raises(RuntimeError, match="position nan"):
        o.step(1)
    assert o.counters()["nan"] == 1 and o.counters()["steps"] == 0


def test_deleting_a_leaving_particle_is_refused():
    g, mats, run, d, i = moving_block(limit=-5)
    o = orc.Oracle(g, mats, run, d, i)
    with pytest.raises(RuntimeError, match="deleted into the reservoir"):
        o.step(60)
