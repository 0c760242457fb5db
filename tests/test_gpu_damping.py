"""Constant grid and particle damping on the HIP path against the oracle (relative L2 <= 1e-9 of
particle positions, velocities and stresses), and in a batch whose instances carry different
damping: every instance equals its solo run bit for bit."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _with_header(xml, extra):
    return xml.replace("  </MPMHeader>", extra + "  </MPMHeader>", 1)


def _check(sim, ora, dim, label, tol=TOL):
    g, _ = sim.particles()
    r, _ = ora.particles()
    errs = compare(g, r, dim)
    print(label, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, f"{label}: {k} relative L2 {v:.3e} > {tol}"


@pytest.mark.parametrize("base, dim, label", [
    (configs.elastic_block_2d(cells=16, block=8), 2, "2D uGIMP"),
    (configs.elastic_block_2d(cells=16, block=8, shape="lCPDI"), 2, "2D lCPDI"),
    (configs.j2_block_3d(cells=16, lo=4, hi=12, vx=2e5), 3, "3D J2"),
])
def test_damped_parity(base, dim, label):
    print(f"case: damping, {label}", flush=True)
    xml = _with_header(base, "    <Damping>4000</Damping>\n    <PDamping>1500</PDamping>\n")
    prob = nm.Problem.from_xml(xml)
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    for n in (1, 19, 40):
        sim.step(n)
        ora.step(n)
        _check(sim, ora, dim, f"{label} damped +{n}")


def test_batch_with_damped_and_undamped_instances():
    print("case: batch of damped and undamped instances", flush=True)
    probs = [nm.Problem.from_xml(_with_header(configs.elastic_block_2d(cells=16, block=6, seed=3),
                                              "    <Damping>3000</Damping>\n")),
             nm.Problem.from_xml(configs.elastic_block_2d(cells=20, block=8, seed=4)),
             nm.Problem.from_xml(_with_header(configs.j2_block_2d(cells=20, block=8, seed=6),
                                              "    <PDamping>2000</PDamping>\n"))]
    batch = nm.Batch(probs)
    batch.step(30)
    for k, prob in enumerate(probs):
        d, i = batch.instance_particles(k)
        sim = nm.Simulation.from_problem(prob)
        sim.step(30)
        da, ia = sim.particles()
        assert np.array_equal(d, da), f"instance {k} differs from its solo run"
        ora = orc.Oracle.from_problem(prob)
        ora.step(30)
        errs = compare(d, ora.particles()[0], 2)
        assert max(errs.values()) <= TOL, (k, errs)


def _material_damping(xml, value, block=1):
    """xml with <PDamping> in the material named 'block <block>' (or the first material)"""
    tag = f'Name="block {block}">' if f'Name="block {block}">' in xml else "<Material "
    i = xml.index(tag)
    j = xml.index("<nu>", i)
    return xml[:j] + f"<PDamping>{value!r}</PDamping>\n    " + xml[j:]


@pytest.mark.parametrize("label, make, checks", [
    ("2D block, material damping only", lambda: _material_damping(configs.elastic_block_2d(cells=16, block=8), 900.),
     (1, 20, 60)),
    ("2D block, material damping over the run's", lambda: _material_damping(
        _with_header(configs.elastic_block_2d(cells=16, block=8, shape="lCPDI"),
                     "    <Damping>2000</Damping>\n    <PDamping>1500</PDamping>\n"), 250.), (1, 20, 60)),
    ("2D contact blocks, block 1 damped", lambda: _material_damping(
        _with_header(configs.contact_blocks_2d(normals=2, law_type=61, friction=0., vy=3000.),
                     "    <PDamping>800</PDamping>\n"), 100.), (60, 200)),
])
def test_material_damping_parity(label, make, checks):
    # <PDamping> in <Material> replaces the run's particle damping for that material's particles
    print(f"case: {label}", flush=True)
    prob = nm.Problem.from_xml(make())
    assert any(m.usePDamping for m in prob.describe()[1])
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    done = 0
    for target in checks:
        sim.step(target - done)
        ora.step(target - done)
        done = target
        _check(sim, ora, 2, f"{label} step {target}")
