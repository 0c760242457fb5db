"""Constant grid and particle damping of the FLIP update: the <Damping>/<PDamping> header settings
(MPMReadHandler.cpp:311-350) and the oracle's update (MatPoint3D::MoveParticle,
MatPoint3D.cpp:104-160) pinned by a closed form: a block translating rigidly has zero internal
forces and grid velocity equal to its own, so each step multiplies its velocity by
1 - (alpha_p + alpha_g) dt."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def _with_header(xml, extra):
    return xml.replace("  </MPMHeader>", extra + "  </MPMHeader>", 1)


def test_header_settings():
    xml = _with_header(configs.elastic_block_2d(cells=16, block=8), "    <Damping>2000</Damping>\n    <PDamping>500.5</PDamping>\n")
    _, _, run, _, _ = nm.Problem.from_xml(xml).describe()
    assert run.gridDamping == 2000. and run.particleDamping == 500.5
    _, _, run, _, _ = nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8)).describe()
    assert run.gridDamping == 0. and run.particleDamping == 0.


@pytest.mark.parametrize("extra, msg", [
    ('    <Damping function="t">1</Damping>\n', "function"),
    ("    <FeedbackDamping>1</FeedbackDamping>\n", "FeedbackDamping"),
])
def test_header_refusals(extra, msg):
    with pytest.raises(nm.MPMError, match=msg):
        nm.Problem.from_xml(_with_header(configs.elastic_block_2d(cells=16, block=8), extra))


def test_material_damping_refused():
    xml = configs.elastic_block_2d(cells=16, block=8).replace("<nu>", "<PDamping>10</PDamping>\n    <nu>", 1)
    with pytest.raises(nm.MPMError, match="PDamping"):
        nm.Problem.from_xml(xml)


@pytest.mark.parametrize("ga, pa", [(3000., 0.), (0., 4000.), (1500., 2500.)])
def test_rigid_translation_decays_geometrically(ga, pa):
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0., jitter=0.)
    xml = _with_header(base, f"    <Damping>{ga!r}</Damping>\n    <PDamping>{pa!r}</PDamping>\n")
    prob = nm.Problem.from_xml(xml)
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    d[F["vel_x"]] = 3000.
    d[F["vel_y"]] = -1000.
    o = orc.Oracle(g, mats, run, d, i)
    dt = run.timestep
    for k in range(1, 21):
        o.step(1)
        r, _ = o.particles()
        f = (1. - (pa + ga) * dt) ** k
        np.testing.assert_allclose(r[F["vel_x"]], 3000. * f, rtol=1e-11)
        np.testing.assert_allclose(r[F["vel_y"]], -1000. * f, rtol=1e-11)
        assert np.max(np.abs(r[F["sp_xx"]])) < 1e-6 * 3000.      # no deformation


def test_undamped_is_unchanged():
    # zero damping takes the plain FLIP update: same numbers as before damping existed
    base = configs.elastic_block_2d(cells=16, block=8)
    a = orc.Oracle.from_problem(nm.Problem.from_xml(base))
    b = orc.Oracle.from_problem(nm.Problem.from_xml(_with_header(base, "    <Damping>0</Damping>\n")))
    a.step(10)
    b.step(10)
    assert np.array_equal(a.particles()[0], b.particles()[0])
