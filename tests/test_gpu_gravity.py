"""Gravity on the HIP path against the oracle (relative L2 <= 1e-9 of particle positions,
velocities and stresses): a 2D block clamped on its bottom edge sagging under gravity, and the
3D J2 block falling with damping."""
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from helpers import compare

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _gravity(xml, gx=0., gy=0., gz=0.):
    body = f"  <Gravity>\n    <BodyXForce>{gx!r}</BodyXForce>\n    <BodyYForce>{gy!r}</BodyYForce>\n"
    body += f"    <BodyZForce>{gz!r}</BodyZForce>\n  </Gravity>\n"
    return xml.replace("</JANFEAInput>", body + "</JANFEAInput>", 1)


def _run(xml, dim, label):
    prob = nm.Problem.from_xml(xml)
    sim = nm.Simulation.from_problem(prob)
    ora = orc.Oracle.from_problem(prob)
    for n in (1, 19, 40):
        sim.step(n)
        ora.step(n)
        g, _ = sim.particles()
        r, _ = ora.particles()
        errs = compare(g, r, dim)
        print(label, n, {k: f"{v:.3e}" for k, v in errs.items()}, flush=True)
        for k, v in errs.items():
            assert v <= TOL, f"{label}: {k} relative L2 {v:.3e}"


def test_clamped_block_under_gravity_2d():
    print("case: 2D block clamped on its bottom edge under gravity", flush=True)
    xml = configs.grid_bcs(configs.elastic_block_2d(cells=16, block=8, vmax=0.),
                           [("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [{"dir": 1, "vel": 0.}, {"dir": 2, "vel": 0.}])])
    _run(_gravity(xml, 0., -5.0e6), 2, "clamped 2D gravity")


def test_j2_block_falling_with_damping_3d():
    print("case: 3D J2 block falling, damped", flush=True)
    xml = configs.j2_block_3d(cells=16, lo=4, hi=12, vx=2e5)
    xml = xml.replace("  </MPMHeader>", "    <Damping>1000</Damping>\n  </MPMHeader>", 1)
    _run(_gravity(xml, 1.0e6, 0., -9.81e6), 3, "J2 3D gravity")
