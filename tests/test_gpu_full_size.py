"""The benchmark configurations at full size. The 3D J2 GIMP block (128^3 cells, 7,077,888 particles, halves colliding
at +-50 m/s): two steps against the oracle in its threaded timing mode (relative L2 <= 1e-9 of
positions, displacements, velocities and stresses), and size-independent properties over longer
runs: total momentum conserved (no conditions or body forces act), grid mass equal to particle
mass, and the same bits from a second context that sorts its rows every step and steps without
graphs. The contact cylinders (160^3), the 256-instance batched sweep and served requests of the 3D
stream at their own sizes follow below."""
import os

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare

pytestmark = pytest.mark.gpu

F = nm.F


def _momentum(d):
    m = d[F["mass"]]
    return np.array([np.sum(m * d[F["vel_" + a]]) for a in "xyz"]), np.sum(m * np.abs(d[F["vel_x"]]))


def test_full_size_block():
    print("case: 3D J2 block, 128^3 cells, full size", flush=True)
    prob = nm.Problem.from_xml(configs.j2_block_3d(cells=128, lo=16, hi=112))
    g, mats, run, n, _ = prob.describe()
    assert n == 7077888
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    ora.set_threads(min(16, os.cpu_count() or 1))
    d0, _ = sim.particles()
    p0, scale = _momentum(d0)

    sim.step(2)
    ora.step(2)
    gd, _ = sim.particles()
    od, _ = ora.particles()
    errs = compare(gd, od, 3)
    print("full size, 2 steps vs oracle", {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-9, f"{k} relative L2 {v:.3e}"
    del ora, od

    sim.step(38)
    d, _ = sim.particles()
    p, _ = _momentum(d)
    print("momentum change / scale", np.abs(p - p0).max() / scale, flush=True)
    assert np.abs(p - p0).max() <= 1e-12 * scale
    mass, _, _ = sim.grid_fields()
    assert abs(mass.sum() - d[F["mass"]].sum()) <= 1e-12 * d[F["mass"]].sum()

    other = nm.Simulation.from_problem(prob)
    other.set_sort_interval(1)
    for _ in range(40):
        other.step(1)
    assert np.array_equal(other.particles()[0], d), "row order or graphs changed the bits"


def test_full_size_contact_cylinders():
    # the 160^3 two-body contact configuration (lCPDI neo-Hookean cylinders, Coulomb mu = 0.3, AVGG
    # normals): positions and velocities against the oracle while the bodies close rigidly (their
    # stresses are rounding noise then), and total momentum through the head-on impact (about step
    # 96), where frictional stick or slip follows rounding and is no parity case
    print("case: 3D contact cylinders, 160^3 cells, full size", flush=True)
    prob = nm.Problem.from_xml(configs.contact_cylinders_3d(cells=160, radius=20., length=60., gap=4.))
    g, mats, run, n, _ = prob.describe()
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    ora.set_threads(min(16, os.cpu_count() or 1))
    d0, _ = sim.particles()
    p0, scale = _momentum(d0)
    sim.step(2)
    ora.step(2)
    errs = compare(sim.particles()[0], ora.particles()[0], 3)
    print(f"contact full size ({n} particles), 2 steps vs oracle", {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    assert errs["pos"] <= 1e-9 and errs["vel"] <= 1e-9
    del ora
    sim.step(148)
    d, _ = sim.particles()
    p, _ = _momentum(d)
    print("momentum change / scale after impact", np.abs(p - p0).max() / scale, flush=True)
    assert np.abs(p - p0).max() <= 1e-11 * scale
    assert sim.stats().leftGrid == 0


def test_full_size_sweep_members_equal_alone():
    # all 256 members of the batched 2D sweep stepped together; a sample across sizes and both
    # materials equals its solo run bit for bit, and member 0 matches the oracle
    print("case: batched 2D sweep, 256 instances", flush=True)
    members = []
    for k in range(256):
        xml, fix, _ = configs.sweep_instance_2d(k)
        prob = nm.Problem.from_xml(xml)
        d, i = prob.particles()
        fix(d)
        members.append((prob, (d, i)))
    batch = nm.Batch([m[0] for m in members], [m[1] for m in members])
    print("particles", int(sum(batch.counts)), flush=True)
    batch.step(12)
    bd, _ = batch.particles()                      # once: rows of all instances, instance after instance
    sizes = np.array(batch.counts)
    sample = sorted(set([0, 255, int(np.argmin(sizes)), int(np.argmax(sizes))] + list(range(17, 256, 37))))
    for k in sample:
        prob, (d0, i0) = members[k]
        g, mats, run, _, _ = prob.describe()
        sim = nm.Simulation(g, mats, run)
        sim.upload(d0, i0)
        sim.set_tile_shape(batch.tile_shape)
        sim.step(12)
        d = bd[:, batch.offsets[k]:batch.offsets[k + 1]]
        assert np.array_equal(d, sim.particles()[0]), f"member {k} differs from its solo run"
        sim.close()
    prob, (d0, i0) = members[0]
    g, mats, run, _, _ = prob.describe()
    ora = orc.Oracle(g, mats, run, d0, i0)
    ora.step(12)
    errs = compare(bd[:, batch.offsets[0]:batch.offsets[1]], ora.particles()[0], 2)
    print("member 0 vs oracle", {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= 1e-9, f"member 0: {k} relative L2 {v:.3e}"


def test_full_size_served_requests_equal_alone():
    # four requests of the served 3D stream at its own sizes (50k-2M particles, 200-2000 steps,
    # J2 and neo-Hookean), two at once and two while those run: every served result equals the
    # same solve alone bit for bit, and the smallest matches the oracle over its first steps
    print("case: served 3D stream, full-size requests", flush=True)
    reqs = []
    for k in (3, 4, 10, 12):        # 55k and 1.9M neo-Hookean, 475k and 111k J2 particles
        xml, steps, _ = configs.serve_request_3d(k)
        reqs.append((nm.Problem.from_xml(xml), steps))
    print("requests", [(p.describe()[3], s) for p, s in reqs], flush=True)
    srv = nm.Server(analysis=nm.THREED, graph_steps=8)
    ids = [srv.submit_problem(*reqs[0]), srv.submit_problem(*reqs[1])]
    srv.advance(40)
    ids += [srv.submit_problem(*reqs[2]), srv.submit_problem(*reqs[3])]
    rounds = 0
    while any(srv.info(r)["state"] < nm.REQ_DONE for r in ids):
        srv.advance(8)
        rounds += 1
        assert rounds < 2000
    for (prob, steps), rid in zip(reqs, ids):
        d, i = srv.collect(rid)
        g, mats, run, _, _ = prob.describe()
        sim = nm.Simulation(g, mats, run)
        sim.upload(*prob.particles())
        sim.step(steps)
        assert np.array_equal(d, sim.particles()[0]), f"request of {prob.describe()[3]} particles differs alone"
        sim.close()
    srv.close()
    k = int(np.argmin([p.describe()[3] for p, _ in reqs]))
    prob, _ = reqs[k]
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    ora.set_threads(min(16, os.cpu_count() or 1))
    sim.step(20)
    ora.step(20)
    errs = compare(sim.particles()[0], ora.particles()[0], 3)
    print(f"request of {prob.describe()[3]} particles, 20 steps vs oracle", {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for key, v in errs.items():
        assert v <= 1e-9, f"{key} relative L2 {v:.3e}"
