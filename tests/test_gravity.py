"""Gravity (constant body force, <Gravity><BodyXForce/>...</Gravity>, MPMReadHandler.cpp:1627-1660)
added as mass * g to every active field's force after RestoreMomenta (PostForcesTask.cpp:62-91,
MatVelocityField.cpp:254-258), pinned by free fall: with no other forces the total particle
momentum after k steps is M g k dt, and a body at rest stays undeformed."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def _gravity(xml, gx=0., gy=0., gz=0.):
    body = f"  <Gravity>\n    <BodyXForce>{gx!r}</BodyXForce>\n    <BodyYForce>{gy!r}</BodyYForce>\n"
    body += f"    <BodyZForce>{gz!r}</BodyZForce>\n  </Gravity>\n"
    return xml.replace("</JANFEAInput>", body + "</JANFEAInput>", 1)


def test_gravity_input():
    prob = nm.Problem.from_xml(_gravity(configs.elastic_block_2d(cells=16, block=8), 0., -9810.))
    _, _, run, _, _ = prob.describe()
    assert list(run.gravity) == [0., -9810., 0.]
    xml = configs.elastic_block_2d(cells=16, block=8).replace(
        "</JANFEAInput>", "  <Gravity>\n    <GridBodyXForce>x</GridBodyXForce>\n  </Gravity>\n</JANFEAInput>")
    with pytest.raises(nm.MPMError, match="GridBodyXForce"):
        nm.Problem.from_xml(xml)


@pytest.mark.parametrize("dim, g", [(2, (3.0e5, -9.81e5, 0.)), (3, (0., 2.0e5, -9.81e5))])
def test_free_fall(dim, g):
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0., jitter=0.) if dim == 2 else \
        configs.j2_block_3d(cells=12, lo=3, hi=9, vx=0.)
    prob = nm.Problem.from_xml(_gravity(base, *g))
    o = orc.Oracle.from_problem(prob)
    _, _, run, _, _ = prob.describe()
    dt = run.timestep
    for k in range(1, 16):
        o.step(1)
        d, _ = o.particles()
        M = d[F["mass"]].sum()
        for a in range(dim):
            p = np.sum(d[F["mass"]] * d[F["vel_" + "xyz"[a]]])
            assert p == pytest.approx(M * g[a] * k * dt, rel=1e-12, abs=1e-12 * M * 1e6 * k * dt), (k, a)
            # a uniform acceleration: every particle has the same velocity
            np.testing.assert_allclose(d[F["vel_" + "xyz"[a]]], g[a] * k * dt, rtol=1e-10, atol=1e-9 * 1e6 * k * dt)
        assert np.max(np.abs(d[F["sp_xx"]])) < 1e-6
