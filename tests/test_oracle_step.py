"""Pins the oracle's whole step by invariants the mathematics fixes: partition of unity on the
grid, conservation of momentum, exact rigid translation, energy exchange in an elastic vibration,
and agreement of the task-by-task path with the fused step."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def small_block(**kw):
    return nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8, **kw))


def momentum(d):
    m = d[F["mass"]]
    return np.array([(m * d[F["vel_x"]]).sum(), (m * d[F["vel_y"]]).sum(), (m * d[F["vel_z"]]).sum()])


def test_grid_mass_and_momentum_match_particles():
    prob = small_block()
    o = orc.Oracle.from_problem(prob)
    d0, _ = o.particles()
    o.run_task(0)                                    # Initialize
    o.run_task(1)                                    # Extrapolate Mass and Momentum
    mass, pk, _, cnt = o.grid_values()
    np.testing.assert_allclose(mass.sum(), d0[F["mass"]].sum(), rtol=1e-13)
    np.testing.assert_allclose(pk.sum(axis=1), momentum(d0), rtol=1e-12, atol=1e-12 * np.abs(momentum(d0)).max())
    assert ((mass > 0) == (cnt > 0)).all()
    # every particle of the 2x2-per-cell lattice touches 9 nodes except on regular sites
    assert cnt.sum() >= 4 * d0.shape[1]


@pytest.mark.parametrize("builder", ["elastic_2d", "j2_3d"])
def test_momentum_is_conserved(builder):
    """Internal forces sum to zero over the nodes (sum_i grad N_i = 0), and the FLIP update
    moves exactly the grid's momentum change onto the particles: total momentum is constant."""
    if builder == "elastic_2d":
        prob = small_block()
        steps = 40
    else:
        prob = nm.Problem.from_xml(configs.j2_block_3d(cells=12, lo=3, hi=9))
        steps = 10
    o = orc.Oracle.from_problem(prob)
    d0, _ = o.particles()
    p0 = momentum(d0)
    scale = (d0[F["mass"]] * np.abs(d0[F["vel_x"]])).sum()
    o.step(steps)
    d1, _ = o.particles()
    np.testing.assert_allclose(momentum(d1), p0, rtol=0, atol=1e-11 * scale)
    # while the stresses are far from zero (the test is not vacuous)
    assert np.abs(d1[F["sp_xx"]]).max() > 0


def test_rigid_translation_is_exact():
    """Uniform velocity and no stress: grid velocities equal the particle velocity, the velocity
    gradient vanishes, and every particle moves by v dt per step."""
    prob = small_block(vmax=0.)
    d, iv = prob.particles()
    d[F["vel_x"]] = 1234.5
    d[F["vel_y"]] = -321.0
    g, mats, run, _, _ = prob.describe()
    o = orc.Oracle(g, mats, run, d, iv)
    o.step(5)
    d1, _ = o.particles()
    dt = run.timestep
    np.testing.assert_allclose(d1[F["pos_x"]] - d[F["pos_x"]], 5 * 1234.5 * dt, rtol=1e-9)
    np.testing.assert_allclose(d1[F["pos_y"]] - d[F["pos_y"]], 5 * -321.0 * dt, rtol=1e-9)
    np.testing.assert_allclose(d1[F["vel_x"]], 1234.5, rtol=1e-12)
    assert np.abs(d1[F["sp_xx"]]).max() < 1e-15 * mats[0].C[1][1]   # roundoff strains only


def energies(d):
    """kinetic energy about the centre of mass (the sine profile also translates the block)
    and the small-strain elastic energy sum m sp:eps / 2"""
    m = d[F["mass"]]
    vc = momentum(d) / m.sum()
    ke = 0.5 * (m * ((d[F["vel_x"]] - vc[0]) ** 2 + (d[F["vel_y"]] - vc[1]) ** 2 + (d[F["vel_z"]] - vc[2]) ** 2)).sum()
    se = 0.5 * (m * (d[F["sp_xx"]] * d[F["ep_xx"]] + d[F["sp_yy"]] * d[F["ep_yy"]] +
                     d[F["sp_zz"]] * d[F["ep_zz"]] + d[F["sp_xy"]] * d[F["ep_xy"]])).sum()
    return ke, se


def test_elastic_vibration_exchanges_energy():
    """The sine shear mode of the free block trades kinetic for strain energy and keeps the sum
    within a small drift; a flipped force sign would grow the kinetic energy instead."""
    prob = small_block(vmax=100.)
    o = orc.Oracle.from_problem(prob)
    d0, _ = o.particles()
    ke0, se0 = energies(d0)
    assert se0 == 0.
    kes, ses = [], []
    for _ in range(30):
        o.step(10)
        d, _ = o.particles()
        ke, se = energies(d)
        kes.append(ke)
        ses.append(se)
    kes, ses = np.array(kes), np.array(ses)
    assert kes.min() < 0.5 * ke0                     # energy moved into strain
    assert ses.max() > 0.4 * ke0
    total = kes + ses
    assert np.abs(total - ke0).max() < 0.01 * ke0


def test_task_by_task_equals_fused_step():
    prob = nm.Problem.from_xml(configs.j2_block_3d(cells=10, lo=3, hi=7))
    a = orc.Oracle.from_problem(prob)
    b = orc.Oracle.from_problem(prob)
    a.step(2)
    for _ in range(2):
        for t in range(10):
            b.run_task(t)
    da, ia = a.particles()
    db, ib = b.particles()
    assert np.array_equal(da, db) and np.array_equal(ia, ib)


@pytest.mark.parametrize("xml", [
    configs.j2_block_3d(cells=12, lo=3, hi=9),
    configs.contact_cylinders_3d(cells=20, radius=4., length=6., gap=1., friction=0.),
    configs.neohookean_block_3d(cells=14, lo=3, hi=11, shape="lCPDI"),
    configs.elastic_block_2d(cells=24, block=12),
], ids=["j2_3d", "contact_3d", "lcpdi_3d", "elastic_2d"])
def test_threaded_timing_mode_matches_serial(xml):
    """The threaded mode timed as the CPU baseline computes the same step: its grid sums differ
    from the serial restatement's only in rounding (relative L2 of particle state ~1e-15 after 20
    steps), two threaded runs agree bitwise, and a lost per-thread sum would show as O(1)."""
    prob = nm.Problem.from_xml(xml)
    ser = orc.Oracle.from_problem(prob)
    ser.step(20)
    ds, _ = ser.particles()
    runs = []
    for _ in range(2):
        o = orc.Oracle.from_problem(prob)
        o.set_threads(4)
        o.step(20)
        runs.append(o.particles()[0])
    assert np.array_equal(runs[0], runs[1])
    for f in ("pos_x", "pos_y", "vel_x", "vel_y"):
        a, b = runs[0][F[f]], ds[F[f]]
        assert np.linalg.norm(a - b) <= 1e-12 * max(np.linalg.norm(b), 1e-300), f
    disp = runs[0][F["pos_x"]] - runs[0][F["orig_x"]]
    disp_s = ds[F["pos_x"]] - ds[F["orig_x"]]
    assert np.linalg.norm(disp - disp_s) <= 1e-10 * np.linalg.norm(disp_s)
