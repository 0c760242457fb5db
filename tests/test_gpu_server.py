"""Solve server: requests admitted by device footprint, resident and stepped together, moved
between batch layouts as others finish or arrive; every result must equal the same solve run
alone, bit for bit."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs

pytestmark = pytest.mark.gpu


def _requests(ks, nmin=2000, nmax=20000, smin=5, smax=40):
    out = []
    for k in ks:
        xml, steps, _ = configs.serve_request_3d(k, pad=3, nmin=nmin, nmax=nmax, smin=smin, smax=smax)
        prob = nm.Problem.from_xml(xml)
        out.append((prob, steps))
    return out


def _alone(prob, steps, particles=None):
    g, mats, run, _, _ = prob.describe()
    sim = nm.Simulation(g, mats, run)
    d, i = prob.particles() if particles is None else particles
    sim.upload(d, i)
    sim.step(steps)
    out = sim.particles()
    sim.close()
    return out


def _assert_alone(srv, rid, prob, steps, label, particles=None):
    d, i = srv.collect(rid)
    da, ia = _alone(prob, steps, particles)
    assert np.array_equal(d, da), f"{label}: served result differs from the solo run"
    assert np.array_equal(i, ia), f"{label}: int fields differ"


def test_served_results_equal_alone_with_staggered_arrivals():
    reqs = _requests(range(7))
    srv = nm.Server(analysis=nm.THREED, graph_steps=4)
    ids = []
    # two at once, then one more after a few rounds, and the rest while the first ones run
    for prob, steps in reqs[:2]:
        ids.append(srv.submit_problem(prob, steps))
    srv.advance(3)
    ids.append(srv.submit_problem(*reqs[2]))
    srv.advance(5)
    srv.advance(2)
    for prob, steps in reqs[3:]:
        ids.append(srv.submit_problem(prob, steps))
        srv.advance(3)
    rounds = 0
    while any(srv.info(r)["state"] < nm.REQ_DONE for r in ids):
        srv.advance(8)
        rounds += 1
        assert rounds < 1000
    c = srv.counts()
    assert c["layouts"] >= 4, c           # arrivals and completions changed the layout
    for k, (rid, (prob, steps)) in enumerate(zip(ids, reqs)):
        inf = srv.info(rid)
        assert inf["stepsDone"] == steps and inf["finished"] >= inf["admitted"] >= inf["submitted"]
        _assert_alone(srv, rid, prob, steps, f"request {k}")
    srv.close()


def test_memory_budget_bounds_residency():
    reqs = _requests(range(10, 16))
    fps = []
    for prob, _ in reqs:
        g, _, run, n, _ = prob.describe()
        fps.append(nm.footprint(g, run, n))
    budget = int(2.2 * max(fps))           # about two resident at a time, moves included
    srv = nm.Server(analysis=nm.THREED, memory_budget=budget, graph_steps=2)
    ids = [srv.submit_problem(prob, steps) for prob, steps in reqs]
    most = 0
    while any(srv.info(r)["state"] < nm.REQ_DONE for r in ids):
        srv.advance(6)
        c = srv.counts()
        assert c["resident_bytes"] <= budget
        most = max(most, c["resident"])
    assert 1 <= most < len(reqs)
    # first come, first admitted
    adm = [srv.info(r)["admitted"] for r in ids]
    assert adm == sorted(adm)
    for k, (rid, (prob, steps)) in enumerate(zip(ids, reqs)):
        _assert_alone(srv, rid, prob, steps, f"request {k}")
    with pytest.raises(nm.MPMError):
        small = nm.Server(analysis=nm.THREED, memory_budget=min(fps) // 2)
        small.submit_problem(*reqs[0])
    srv.close()


def test_request_loop_thread_serves_arrivals():
    reqs = _requests(range(20, 26))
    srv = nm.Server(analysis=nm.THREED, graph_steps=4)
    srv.start()
    ids = []
    for prob, steps in reqs:
        ids.append(srv.submit_problem(prob, steps))
        srv.wait(ids[0], timeout=0.005)     # let the loop run between arrivals
    for rid in ids:
        assert srv.wait(rid, timeout=60.0)
    srv.stop()
    for k, (rid, (prob, steps)) in enumerate(zip(ids, reqs)):
        _assert_alone(srv, rid, prob, steps, f"request {k}")
    with pytest.raises(nm.MPMError):
        srv.collect(ids[0])                # collected once
    srv.close()


def test_served_2d_sweep_members_equal_alone():
    srv = nm.Server(analysis=nm.PLANE_STRAIN, graph_steps=2)
    members = []
    for k, steps in zip([12, 30, 13, 35, 0], [9, 17, 4, 25, 11]):
        xml, fix, _ = configs.sweep_instance_2d(k)
        prob = nm.Problem.from_xml(xml)
        d, i = prob.particles()
        fix(d)
        members.append((prob, steps, (d, i), srv.submit_problem(prob, steps, (d, i))))
        srv.advance(2)
    while any(srv.info(m[3])["state"] < nm.REQ_DONE for m in members):
        srv.advance(4)
    for prob, steps, parts, rid in members:
        _assert_alone(srv, rid, prob, steps, f"request {rid}", parts)
    srv.close()


def test_server_rejects_mismatched_requests():
    srv = nm.Server(analysis=nm.THREED)
    prob = nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=6))
    with pytest.raises(nm.MPMError):
        srv.submit_problem(prob, 5)          # a 2D request on a 3D server
    prob3 = _requests([1])[0][0]
    with pytest.raises(nm.MPMError):
        srv.submit_problem(prob3, 0)         # no steps
    srv.close()


def test_served_results_match_oracle_at_50k_particles():
    """A J2 and a neo-Hookean block of ~47k particles each (the served stream's smallest size),
    served together with a third request that joins after a few steps, against the oracle."""
    import oracle as orc
    from helpers import compare
    srv = nm.Server(analysis=nm.THREED, graph_steps=4)
    cases = []
    for k in (10, 0):                        # request 10 is J2, request 0 neo-Hookean
        xml, _, j2 = configs.serve_request_3d(k, pad=3, nmin=50000, nmax=50000, smin=20, smax=20)
        prob = nm.Problem.from_xml(xml)
        cases.append((prob, 20, j2, srv.submit_problem(prob, 20)))
    assert [c[2] for c in cases] == [True, False]
    srv.advance(4)
    late = _requests([3], nmin=5000, nmax=5000, smin=12, smax=12)[0]
    late_id = srv.submit_problem(*late)
    while any(srv.info(r)["state"] < nm.REQ_DONE for r in [c[3] for c in cases] + [late_id]):
        srv.advance(4)
    for prob, steps, j2, rid in cases:
        d, _ = srv.collect(rid)
        g, mats, run, n, _ = prob.describe()
        dd, ii = prob.particles()
        print(f"served {'J2' if j2 else 'neo-Hookean'} block, {n} particles, {steps} steps", flush=True)
        ora = orc.Oracle(g, mats, run, dd, ii)
        ora.step(steps)
        r, _ = ora.particles()
        errs = compare(d, r, 3)
        print(errs, flush=True)
        for key, v in errs.items():
            assert v <= 1e-9, (j2, key, v)
        if j2:
            assert np.count_nonzero(r[nm.F["hist0"]]) > 0      # some particles yielded
    srv.close()
