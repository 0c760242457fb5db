"""Particles that leave the grid on the HIP path against the oracle: pushed back into their
element (ReturnToElement, ResetElementsTask.cpp:232-265) with the same left-grid counts and
flags, and the run stopped on the device when the <LeaveLimit> count is reached, at a NaN
position, or when a leaving particle would be deleted (ResetElementsTask.cpp:71-151)."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare
from test_leaving import moving_block, _header

pytestmark = pytest.mark.gpu

F, I = nm.F, nm.I
TOL = 1e-9


def moving_block_3d(limit, speed=0.3, vx=20000.):
    """The 3D J2 block's halves colliding at +-vx (mm/s) while the block drifts toward the +x face
    of the grid by `speed` cells per step"""
    xml = _header(configs.j2_block_3d(cells=12, lo=3, hi=9, vx=vx), f"    <LeaveLimit>{limit}</LeaveLimit>\n")
    prob = nm.Problem.from_xml(xml)
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    d[F["vel_x"]] += speed * g.dx / run.timestep
    d[F["vel_z"]] += -0.1 * speed * g.dx / run.timestep
    return g, mats, run, d, i


def _both(g, mats, run, d, i):
    sim = nm.Simulation(g, mats, run)
    sim.upload(d, i)
    return sim, orc.Oracle(g, mats, run, d, i)


@pytest.mark.parametrize("dim", [2, 3])
def test_pushed_back_particles_match_oracle(dim):
    # the blocks deform while they drift (a shearing profile in 2D, colliding halves in 3D): a rigid
    # drift leaves only rounding noise in the stresses, which no relative bound can compare
    case = moving_block(limit=100000, vmax=2000.) if dim == 2 else moving_block_3d(100000)
    sim, ora = _both(*case)
    print(f"leaving block {dim}D, {case[3].shape[1]} particles", flush=True)
    for n in (10, 20):
        sim.step(n)
        ora.step(n)
    g, gi = sim.particles()
    r, ri = ora.particles()
    errs = compare(g, r, dim)
    print(errs, flush=True)
    for k, v in errs.items():
        assert v <= TOL, (k, v)
    st = sim.stats()
    c = ora.counters()
    assert c["leftGrid"] > 0 and st.leftGrid == c["leftGrid"] and st.stopped == 0
    assert np.array_equal(gi[I["left"]], ri[I["left"]])
    assert np.array_equal(gi[I["elem"]], ri[I["elem"]])


def test_run_stops_when_the_limit_is_reached():
    case = moving_block(limit=5, vmax=2000.)
    sim, ora = _both(*case)
    print("leaving block, LeaveLimit 5", flush=True)
    for k in range(60):
        try:
            ora.step(1)
        except RuntimeError as e:
            assert "Too many particles" in str(e)
            break
    sim.step(60)
    with pytest.raises(nm.MPMError, match="Too many particles have left the grid"):
        sim.sync()
    st = sim.stats()
    c = ora.counters()
    assert st.stopped == nm.STOP_TOO_MANY_LEFT and st.steps == c["steps"] == k
    assert st.leftGrid >= 5 and c["leftGrid"] == 5
    with pytest.raises(nm.MPMError, match="Too many particles"):
        sim.step(1)
    # the stopped state is the stopping step's: velocities and stresses as the oracle's; positions
    # of the particles that did not leave as well (the oracle stops pushing back at the fifth)
    g, gi = sim.particles()
    r, ri = ora.particles()
    for name in ("vel_x", "vel_y", "sp_xx", "sp_xy"):
        a, b = g[F[name]], r[F[name]]
        assert np.linalg.norm(a - b) <= TOL * np.linalg.norm(b), name
    stay = (gi[I["left"]] == 0) & (ri[I["left"]] == 0)
    for name in ("pos_x", "pos_y"):
        a, b = g[F[name]][stay], r[F[name]][stay]
        assert np.linalg.norm(a - b) <= TOL * np.linalg.norm(b), name
    # a new upload starts again
    sim.upload(case[3], case[4])
    sim.step(2)
    sim.sync()
    assert sim.stats().steps == 2 and sim.stats().stopped == 0


def test_nan_position_stops_the_run():
    g, mats, run, d, i = moving_block(limit=100000, speed=0.01)
    d[F["vel_y"]][17] = np.nan
    sim, ora = _both(g, mats, run, d, i)
    print("NaN velocity on particle 17", flush=True)
    with pytest.raises(RuntimeError, match="position nan"):
        ora.step(1)
    sim.step(4)
    with pytest.raises(nm.MPMError, match="position nan"):
        sim.sync()
    st = sim.stats()
    assert st.stopped == nm.STOP_NAN and st.steps == 0 and st.nanParticles >= 1


def test_deleting_a_leaving_particle_is_refused():
    sim, ora = _both(*moving_block(limit=-5))
    print("leaving block, LeaveLimit -5 (delete)", flush=True)
    sim.step(60)
    with pytest.raises(nm.MPMError, match="deleted into the reservoir"):
        sim.sync()
    assert sim.stats().stopped == nm.STOP_DELETE
