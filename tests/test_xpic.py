"""XPIC(k) and FMPM(k) through the PeriodicXPIC custom task (Custom_Tasks/PeriodicXPIC.cpp,
XPICExtrapolationTask.cpp, MatVelocityField::XPICSupport, MatPoint3D::MoveParticle), the
oracle's restatement pinned by an independent formulation:

* the extrapolation is a matrix series. With S the particles' shape functions, Mp their masses,
  m = S^T Mp 1 the lumped node masses and A = diag(1/m) S^T Mp S, the reference's double loop over
  each particle's node pairs gives v(k) = sum_{r<k} (P (I - A))^r v0 with v0 the lumped velocity
  and P zeroing the velocity-condition components after each increment; it tends to the
  consistent-mass velocity A^-1 v0 = (S^T Mp S)^-1 p as k grows. Checked with numpy on classic
  (bilinear) shape functions computed here from the particle positions;
* FMPM takes the strain update's grid velocities from that series, XPIC adds a*dt to it for the
  particle update, and the three MoveParticle branches give the particle velocity and position;
* the corrections carry no momentum, so a free block keeps its total momentum;
* the schedule follows the reference's timing: decisions are made before the first step and at
  the end of each step for the next, by step count or by time."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F
I = nm.I


def _classic_block(**kw):
    kw.setdefault("cells", 12)
    kw.setdefault("block", 4)
    kw.setdefault("vmax", 3000.)
    kw.setdefault("jitter", 0.12)
    return configs.elastic_block_2d(shape="Classic", **kw)


def _shape_matrix(prob_or_ora, d, iv, g):
    """S[p, node] of classic bilinear shape functions on each particle's element"""
    nnx = g.nx + 1
    n = d.shape[1]
    S = np.zeros((n, (g.nx + 1) * (g.ny + 1)))
    for p in range(n):
        e = int(iv[I["elem"]][p])
        ei, ej = e % g.nx, e // g.nx
        xmn, xmx = g.xmin + ei * g.dx, g.xmin + (ei + 1) * g.dx
        ymn, ymx = g.ymin + ej * g.dy, g.ymin + (ej + 1) * g.dy
        xi = (2. * d[F["pos_x"]][p] - xmn - xmx) / (xmx - xmn)
        eta = (2. * d[F["pos_y"]][p] - ymn - ymx) / (ymx - ymn)
        for a, b in ((0, 0), (1, 0), (1, 1), (0, 1)):
            w = 0.25 * (1. + (2 * a - 1) * xi) * (1. + (2 * b - 1) * eta)
            S[p, (ej + b) * nnx + ei + a] = w
    return S


def _series(S, mp, v0, k, fixed=None):
    """v(k) = sum_{r<k} (P (I - A))^r v0 on the active nodes (A = diag(1/m) S^T Mp S)"""
    m = S.T @ mp
    act = m > 0.
    A = (S[:, act].T * mp) @ S[:, act] / m[act][:, None]
    out = v0[act].copy()
    d = v0[act].copy()
    for _ in range(k - 1):
        d = d - A @ d
        if fixed is not None:
            d[fixed[act]] = 0.
        out = out + d
    full = np.zeros_like(v0)
    full[act] = out
    return full, act, A


def _run_tasks(o, upto):
    for t in range(upto + 1):
        o.run_task(t)


def test_reader_and_refusals():
    base = _classic_block()
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=3, steps=2))
    _, _, run, _, _ = prob.describe()
    assert (run.xpicOrder, run.xpicFMPM, run.xpicPeriodSteps, run.xpicPeriodTime) == (3, 0, 2, 0.)
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=4, fmpm=True, steps=None, time_ms=0.5))
    _, _, run, _, _ = prob.describe()
    assert (run.xpicOrder, run.xpicFMPM, run.xpicPeriodSteps) == (4, 1, 0)
    assert run.xpicPeriodTime == 0.5e-3
    # periodicCFL: times the mechanics CFL time step (dcell / wave speed, before the CFL factor)
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=2, steps=None, cfl=2.0))
    _, _, run, _, _ = prob.describe()
    assert run.xpicPeriodTime == pytest.approx(2.0 * run.timestep / 0.2, rel=1e-14)   # the block's CFL is 0.2
    # steps override a time
    prob = nm.Problem.from_xml(configs.periodic_xpic(base, order=2, steps=3, time_ms=1.))
    _, _, run, _, _ = prob.describe()
    assert run.xpicPeriodSteps == 3 and run.xpicPeriodTime == 0.
    for xml, msg in [
        (configs.periodic_xpic(base, order=0), "order to 1 or higher"),
        (configs.periodic_xpic(base, order=2, steps=None), "did not activate"),
        (configs.periodic_xpic(base, order=2).replace('<Parameter name="periodicSteps">',
                                                        '<Parameter name="FMPMOrder">2</Parameter><Parameter name="periodicSteps">'),
         "but not both"),
        (configs.periodic_xpic(base, order=2).replace('name="periodicSteps"', 'name="periodicStepsConduction"'),
         "conduction and diffusion"),
        (configs.periodic_xpic(base, order=2).replace('name="periodicSteps"', 'name="bogus"'), "Unrecognized"),
        (configs.periodic_xpic(configs.contact_blocks_2d(cells=32), order=2), "multimaterial"),
    ]:
        with pytest.raises(nm.MPMError, match=msg):
            nm.Problem.from_xml(xml)


@pytest.mark.parametrize("k", [2, 3, 5])
def test_fmpm_strain_update_velocities_are_the_series(k):
    prob = nm.Problem.from_xml(configs.periodic_xpic(_classic_block(), order=k, fmpm=True))
    g, mats, run, n, _ = prob.describe()
    d, iv = prob.particles()
    o = orc.Oracle.from_problem(prob)
    _run_tasks(o, 3)                        # through UpdateStrainsFirst
    assert o.xpic_order() == k
    S = _shape_matrix(prob, d, iv, g)
    mp = d[F["mass"]]
    m = S.T @ mp
    for c, a in ((0, "x"), (1, "y")):
        pvec = S.T @ (mp * d[F["vel_" + a]])
        v0 = np.where(m > 0., pvec / np.where(m > 0., m, 1.), 0.)
        want, act, _ = _series(S, mp, v0, k)
        got = o.grid_velocities()[0][c]
        assert np.allclose(got[act], want[act], rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_fmpm_approaches_the_consistent_mass_velocity():
    # the residual of the consistent-mass equations, p - (S^T Mp S) v(k), falls as k grows; the
    # series at k = 25 still matches numpy
    residuals = []
    for k in (1, 2, 5, 25):
        prob = nm.Problem.from_xml(configs.periodic_xpic(_classic_block(jitter=0.05), order=k, fmpm=True))
        g, mats, run, n, _ = prob.describe()
        d, iv = prob.particles()
        o = orc.Oracle.from_problem(prob)
        _run_tasks(o, 3)
        S = _shape_matrix(prob, d, iv, g)
        mp = d[F["mass"]]
        m = S.T @ mp
        act = m > 0.
        Mc = (S[:, act].T * mp) @ S[:, act]
        pvec = S[:, act].T @ (mp * d[F["vel_x"]])
        got = o.grid_velocities()[0][0]
        if k > 1:
            want, _, _ = _series(S, mp, np.where(act, S.T @ (mp * d[F["vel_x"]]) / np.where(act, m, 1.), 0.), k)
            assert np.allclose(got[act], want[act], rtol=1e-11, atol=1e-11 * np.abs(want).max())
        residuals.append(np.linalg.norm(pvec - Mc @ got[act]))
    assert residuals[0] > residuals[1] > residuals[2] > residuals[3]
    assert residuals[3] < 0.5 * residuals[0]


@pytest.mark.parametrize("fmpm", [False, True])
@pytest.mark.parametrize("k", [1, 2, 4])
def test_particle_update_moves_particles_by_the_series(fmpm, k):
    # one step through UpdateMomenta, then the particle update checked against numpy
    xml = configs.periodic_xpic(_classic_block(), order=k, fmpm=fmpm)
    prob = nm.Problem.from_xml(xml)
    g, mats, run, n, _ = prob.describe()
    o = orc.Oracle.from_problem(prob)
    _run_tasks(o, 6)
    d0, iv = o.particles()
    mass, pk, ftot, cnt = o.grid_values()
    dt = run.timestep
    S = _shape_matrix(prob, d0, iv, g)
    mp = d0[F["mass"]]
    o.run_task(7)
    d1, _ = o.particles()
    act = mass > 0.
    safe = np.where(act, mass, 1.)
    for c, a in ((0, "x"), (1, "y")):
        if k > 1:
            start = pk[c] / safe if fmpm else (pk[c] + ftot[c] * -dt) * (1. / safe)
            vk, _, _ = _series(S, mp, np.where(act, start, 0.), k)
            if not fmpm:
                vk = vk + np.where(act, ftot[c] * (dt / safe), 0.)
        else:
            vk = np.where(act, pk[c] / safe, 0.)          # lumped v+ (PIC-like XPIC(1) / FMPM(1))
        Svk = S @ vk
        vold = d0[F["vel_" + a]]
        vnew = Svk                                        # no damping
        x = d0[F["pos_" + a]] + ((vold + vnew) * (0.5 * dt) if fmpm else (Svk + 0.5 * (vnew - vold)) * dt)
        scale = np.abs(vnew).max()
        assert np.allclose(d1[F["vel_" + a]], vnew, rtol=1e-12, atol=1e-12 * scale), (a, k, fmpm)
        assert np.allclose(d1[F["pos_" + a]], x, rtol=0, atol=1e-12 * scale * dt + 1e-15), (a, k, fmpm)


def test_velocity_conditions_zero_each_increment():
    # the bottom node row held at zero x and y velocity: each increment is zeroed there
    base = _classic_block(cells=12, block=4)
    xml = configs.grid_bcs(base, [("BCLine", {"x1": 3, "y1": 4, "x2": 9, "y2": 4}, [{"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}])])
    xml = configs.periodic_xpic(xml, order=4, fmpm=True)
    prob = nm.Problem.from_xml(xml)
    g, mats, run, n, _ = prob.describe()
    bcs = prob.velocity_bcs()
    assert bcs
    d, iv = prob.particles()
    o = orc.Oracle.from_problem(prob)
    _run_tasks(o, 3)
    mass, pk, ftot, cnt = o.grid_values()
    S = _shape_matrix(prob, d, iv, g)
    mp = d[F["mass"]]
    fixed = np.zeros(S.shape[1], bool)
    fixed[[b.node for b in bcs]] = True
    act = mass > 0.
    for c in (0, 1):
        v0 = np.where(act, pk[c] / np.where(act, mass, 1.), 0.)     # after the mass-momentum pass's conditions
        want, a, _ = _series(S, mp, v0, 4, fixed)
        got = o.grid_velocities()[0][c]
        assert np.allclose(got[a], want[a], rtol=1e-12, atol=1e-12 * max(np.abs(want).max(), 1.))
        assert np.all(got[fixed & act] == 0.)


@pytest.mark.parametrize("fmpm", [False, True])
def test_free_block_keeps_its_momentum(fmpm):
    xml = configs.periodic_xpic(configs.elastic_block_2d(cells=16, block=8, vmax=4000.), order=3, fmpm=fmpm)
    o = orc.Oracle.from_problem(nm.Problem.from_xml(xml))

    def momentum():
        d, _ = o.particles()
        return np.array([np.sum(d[F["mass"]] * d[F["vel_" + a]]) for a in "xy"])

    p0 = momentum()
    total = np.sum(np.abs(o.particles()[0][F["mass"]] * o.particles()[0][F["vel_x"]]))
    for _ in range(10):
        o.step(1)
        assert o.xpic_order() == 3
    assert np.allclose(momentum(), p0, rtol=0, atol=1e-11 * total)


def test_schedule_by_steps_and_by_time():
    base = _classic_block()
    o = orc.Oracle.from_problem(nm.Problem.from_xml(configs.periodic_xpic(base, order=2, steps=3)))
    orders = []
    for _ in range(10):
        o.step(1)
        orders.append(o.xpic_order())
    assert orders == [0, 0, 2, 0, 0, 2, 0, 0, 2, 0]
    # by time, T = 2.5 dt: the decision before step 1 tests mtime + dt = dt; the one at the end of
    # step s (mtime still s's start time) tests the next step's start time; each hit moves the
    # next time to the tested time + T: steps 4 (3 dt), 7 (6 dt >= 5.5 dt), 10 (9 dt >= 8.5 dt)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    xml = configs.periodic_xpic(base, order=2, steps=None, time_ms=2.5 * run.timestep * 1e3)
    o = orc.Oracle.from_problem(nm.Problem.from_xml(xml))
    orders = []
    for _ in range(10):
        o.step(1)
        orders.append(o.xpic_order())
    assert orders == [0, 0, 0, 2, 0, 0, 2, 0, 0, 2]
    # a period shorter than a step: the first two decisions both test dt (the second is made while
    # mtime is still 0), so step 2 is skipped after step 1 moved the next time to 1.5 dt; from then
    # on every step
    xml = configs.periodic_xpic(base, order=2, steps=None, time_ms=0.5 * run.timestep * 1e3)
    o = orc.Oracle.from_problem(nm.Problem.from_xml(xml))
    orders = []
    for _ in range(6):
        o.step(1)
        orders.append(o.xpic_order())
    assert orders == [2, 0, 2, 2, 2, 2]


def test_xpic_particle_pass_with_velocity_conditions():
    # bottom row held in x: the particle pass of XPIC(k > 1) zeroes each increment's x component at
    # those nodes and adds -m dotn / dt to their force (MatVelocityField.cpp:523-534); the held
    # nodes' updated momentum is already zero from the grid-forces pass
    base = _classic_block(cells=12, block=4)
    xml = configs.grid_bcs(base, [("BCLine", {"x1": 3, "y1": 4, "x2": 9, "y2": 4}, [{"dir": 1, "vel": 0.}])])
    k = 3
    results = {}
    for order in (0, k):
        x = configs.periodic_xpic(xml, order=k, steps=1) if order else xml
        prob = nm.Problem.from_xml(x)
        g, mats, run, n, _ = prob.describe()
        bcs = prob.velocity_bcs()
        o = orc.Oracle.from_problem(prob)
        _run_tasks(o, 6)
        results[order] = (prob, g, run, bcs, o, o.grid_values())
    held = [b.node for b in results[k][3]]
    mass0, pk0, _, _ = results[0][5]
    massk, pkk, ftotk, _ = results[k][5]
    act_held = [nd for nd in held if massk[nd] > 0.]
    assert act_held
    # FLIP's momentum pass zeroes them exactly; XPIC(3) skips that pass and keeps the grid-forces
    # pass's zero, up to rounding
    assert np.all(pk0[0][act_held] == 0.)
    assert np.abs(pkk[0][act_held]).max() <= 1e-12 * np.abs(pkk[0]).max()
    prob, g, run, bcs, o, _ = results[k]
    dt = run.timestep
    d0, iv = o.particles()
    S = _shape_matrix(prob, d0, iv, g)
    mp = d0[F["mass"]]
    o.run_task(7)
    d1, _ = o.particles()
    _, _, ftot_after, _ = o.grid_values()
    act = massk > 0.
    safe = np.where(act, massk, 1.)
    fixed = np.zeros(S.shape[1], bool)
    fixed[held] = True
    # x: the series with the held components zeroed after each increment, and the force the
    # zeroing leaves at the held nodes
    v0 = np.where(act, (pkk[0] + ftotk[0] * -dt) * (1. / safe), 0.)
    m = S.T @ mp
    A = (S[:, act].T * mp) @ S[:, act] / m[act][:, None]
    dlt = v0[act].copy()
    vs = v0[act].copy()
    dF = np.zeros(act.sum())
    fa = fixed[act]
    for _ in range(k - 1):
        dlt = dlt - A @ dlt
        dF[fa] += -m[act][fa] * dlt[fa] / dt
        dlt[fa] = 0.
        vs = vs + dlt
    vstar = np.zeros_like(v0)
    vstar[act] = vs + ftotk[0][act] * (dt / massk[act])
    vnew = S @ vstar
    scale = np.abs(vnew).max()
    assert np.allclose(d1[F["vel_x"]], vnew, rtol=1e-12, atol=1e-12 * scale)
    want_f = ftotk[0].copy()
    want_f[act] += dF
    fscale = np.abs(want_f).max()
    assert np.allclose(ftot_after[0][act_held], want_f[act_held], rtol=1e-12, atol=1e-12 * fscale)
    assert np.any(dF[fa] != 0.)


def test_library_checks_the_settings():
    # the C library's own checks (no device needed): an order > 1 holds three more doubles per
    # node for the series increment; orders outside 0..32 and an order without a period are refused
    prob = nm.Problem.from_xml(configs.periodic_xpic(_classic_block(), order=3))
    g, _, run, n, _ = prob.describe()
    flip = nm.RunParams.from_buffer_copy(run)
    flip.xpicOrder = 0
    nodes = ((g.nx + 1) * (g.ny + 1) + 255) // 256 * 256
    assert nm.footprint(g, run, n) - nm.footprint(g, flip, n) == 3 * 8 * nodes
    one = nm.RunParams.from_buffer_copy(run)
    one.xpicOrder = 1
    assert nm.footprint(g, one, n) == nm.footprint(g, flip, n)
    for field, value in (("xpicOrder", -1), ("xpicOrder", 33), ("xpicPeriodSteps", 0)):
        bad = nm.RunParams.from_buffer_copy(run)
        setattr(bad, field, value)
        assert nm.footprint(g, bad, n) == -1
