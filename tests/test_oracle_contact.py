"""Multimaterial contact in the oracle, pinned by what the mechanics fixes: total momentum is
conserved by every contact law (the momentum changes at a node are equal and opposite), bodies
that separate exchange no momentum under a frictional law, the stick law drives the two fields of
a node to the center-of-mass velocity, and friction only adds tangential momentum exchange."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def body_momenta(d, iv):
    out = []
    for m in (0, 1):
        sel = iv[1] == m
        mass = d[F["mass"]][sel]
        out.append(np.array([(mass * d[F[c]][sel]).sum() for c in ("vel_x", "vel_y", "vel_z")]))
    return out


def run(xml, steps):
    prob = nm.Problem.from_xml(xml)
    g, mats, r, _, _ = prob.describe()
    assert r.multiMaterial == 1 and r.numFields == 2
    o = orc.Oracle.from_problem(prob)
    d0, i0 = o.particles()
    o.step(steps)
    d1, i1 = o.particles()
    return d0, i0, d1, i1


@pytest.mark.parametrize("law_type, friction", [(61, 0.3), (61, 0.0), (61, -1.0), (60, 0.0)])
def test_total_momentum_conserved_through_impact(law_type, friction):
    xml = configs.contact_blocks_2d(law_type=law_type, friction=friction, vy=2000.)
    d0, i0, d1, i1 = run(xml, 60)
    a0, b0 = body_momenta(d0, i0)
    a1, b1 = body_momenta(d1, i1)
    scale = np.abs(a0).sum() + np.abs(b0).sum()
    np.testing.assert_allclose(a1 + b1, a0 + b0, rtol=0, atol=1e-12 * scale)
    # the bodies did meet: x momentum moved between them
    assert abs(a1[0] - a0[0]) > 0.1 * abs(a0[0])


def test_separating_bodies_exchange_nothing():
    """Blocks moving apart share nodes for a few steps but never press: each keeps its momentum."""
    xml = configs.contact_blocks_2d(vx=-3000., gap=0.5, vy=1000.)
    d0, i0, d1, i1 = run(xml, 20)
    a0, b0 = body_momenta(d0, i0)
    a1, b1 = body_momenta(d1, i1)
    np.testing.assert_allclose(a1, a0, rtol=0, atol=1e-12 * np.abs(a0).sum())
    np.testing.assert_allclose(b1, b0, rtol=0, atol=1e-12 * np.abs(b0).sum())


def test_stick_law_moves_contact_nodes_to_center_of_mass():
    """With ContactLaw (type 60) the two fields at a shared node leave the contact pass with the
    same velocity (the center-of-mass velocity)."""
    xml = configs.contact_blocks_2d(law_type=60, gap=0.5, vx=1000., vy=500.)
    prob = nm.Problem.from_xml(xml)
    o = orc.Oracle.from_problem(prob)
    for t in range(3):
        o.run_task(t)               # Initialize, MassAndMomentum, PostExtrapolation (contact)
    m, pk, _, cnt = o.grid_fields()
    both = (cnt[0] > 0) & (cnt[1] > 0)
    assert both.sum() > 0
    v0 = pk[0][:, both] / m[0][both]
    v1 = pk[1][:, both] / m[1][both]
    np.testing.assert_allclose(v0, v1, rtol=1e-12, atol=1e-9)


def test_friction_adds_tangential_exchange_only():
    """Frictionless contact (mu = 0) leaves the y momentum of each block nearly unchanged while
    mu = 0.3 transfers y momentum from the sliding block to the other one."""
    res = {}
    for mu in (0.0, 0.3):
        xml = configs.contact_blocks_2d(friction=mu, vy=4000., vx=3000.)
        d0, i0, d1, i1 = run(xml, 60)
        a0, b0 = body_momenta(d0, i0)
        a1, b1 = body_momenta(d1, i1)
        res[mu] = (a1[1] - a0[1]) / b0[1]
    assert abs(res[0.0]) < 0.02
    assert res[0.3] > 5 * max(abs(res[0.0]), 1e-3)
