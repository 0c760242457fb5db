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
I = nm.I


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


def _material_damping(xml, value, block=1):
    """xml with <PDamping> in the material named 'block <block>' (or the first material)"""
    tag = f'Name="block {block}">' if f'Name="block {block}">' in xml else "<Material "
    i = xml.index(tag)
    j = xml.index("<nu>", i)
    return xml[:j] + f"<PDamping>{value!r}</PDamping>\n    " + xml[j:]


def test_material_damping_settings():
    # <PDamping> in <Material> (MPMReadHandler.cpp:319-324, MaterialBase::SetDamping): a constant
    # in 1/s that the material's particles take in place of the run's particle damping
    xml = _material_damping(configs.elastic_block_2d(cells=16, block=8), 125.5)
    _, mats, run, _, _ = nm.Problem.from_xml(xml).describe()
    assert mats[0].usePDamping == 1 and mats[0].pdamping == 125.5 and run.particleDamping == 0.
    _, mats, _, _, _ = nm.Problem.from_xml(_material_damping(configs.elastic_block_2d(cells=16, block=8), 0.)).describe()
    assert mats[0].usePDamping == 1 and mats[0].pdamping == 0.
    _, mats, _, _, _ = nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8)).describe()
    assert mats[0].usePDamping == 0
    bad = configs.elastic_block_2d(cells=16, block=8).replace("<nu>", '<PDamping function="t">10</PDamping>\n    <nu>', 1)
    with pytest.raises(nm.MPMError, match="Material damping function of time is not allowed"):
        nm.Problem.from_xml(bad)


def _two_blocks(ga, pa, ma):
    """the contact blocks in single-material mode (one velocity field), translating rigidly toward
    each other; block 1's material has its own particle damping ma"""
    xml = configs.contact_blocks_2d(cells=40, gap=6., vx=3000., vy=-1000.)   # no shared nodes
    i, j = xml.index("    <MultiMaterialMode"), xml.index("</MultiMaterialMode>\n") + len("</MultiMaterialMode>\n")
    xml = _with_header(xml[:i] + xml[j:], f"    <Damping>{ga!r}</Damping>\n    <PDamping>{pa!r}</PDamping>\n")
    return _material_damping(xml, ma, block=1)


def test_material_damping_replaces_the_run_particle_damping():
    # two rigidly translating blocks: block 1 (material PDamping ma) decays as (1 - (ma + ga) dt)^k,
    # block 2 as (1 - (pa + ga) dt)^k with the run's particle damping
    ga, pa, ma = 700., 2500., 300.
    prob = nm.Problem.from_xml(_two_blocks(ga, pa, ma))
    g, mats, run, _, _ = prob.describe()
    assert run.multiMaterial == 0 and mats[0].usePDamping == 1 and mats[1].usePDamping == 0
    o = orc.Oracle.from_problem(prob)
    d0, i0 = o.particles()
    mat = i0[I["mat"]]
    dt = run.timestep
    for k in range(1, 16):
        o.step(1)
        r, _ = o.particles()
        for m, a in ((0, ma), (1, pa)):
            sel = mat == m
            f = (1. - (a + ga) * dt) ** k
            # block 1 has no y velocity: rounding of order 1e-13 mm/s there
            np.testing.assert_allclose(r[F["vel_x"]][sel], d0[F["vel_x"]][sel] * f, rtol=1e-11, atol=1e-9)
            np.testing.assert_allclose(r[F["vel_y"]][sel], d0[F["vel_y"]][sel] * f, rtol=1e-11, atol=1e-9)


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
