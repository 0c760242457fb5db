"""Archives written by this project, decoded by the reference's own reader: NairnMPM's ExtractMPM
(NairnMPM/tools/ExtractMPM.cpp), compiled from its source into oracle/_ref by oracle/Makefile's
`ref` target. Its little-endian binary-double output (-d) must give back, bit for bit, the state
that was written: positions, mass, material, velocity, stress in Pa (current density times the
specific stress, the pressure folded in), strain, plastic strain, history 1 and displacement."""
import os
import subprocess

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import archive, configs
import oracle as orc

F = nm.F
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXTRACT = os.path.join(ROOT, "oracle", "_ref", "ExtractMPM")
ORDER = "iYYYYYNNNNNYNY"          # velocity, stress, strain, plastic strain, shear components, history

pytestmark = pytest.mark.skipif(not os.path.exists(EXTRACT),
                                reason="the reference's ExtractMPM is built only where its sources are present")


def _extract(path, quantities, dim, out):
    args = [EXTRACT, "-d"]
    for q in quantities:
        args += ["-q", q]
    subprocess.run(args + ["-o", out, path], check=True, capture_output=True)
    data = np.fromfile(out + ".data", dtype="<f8")
    ncol = dim + len(quantities)
    assert data.size % ncol == 0
    return data.reshape(-1, ncol).T


@pytest.mark.parametrize("xml, dim", [
    (configs.j2_block_3d(cells=12, lo=3, hi=9, vx=2e5), 3),
    (configs.neohookean_block_3d(cells=12, lo=3, hi=9), 3),
    (configs.elastic_block_2d(cells=16, block=8), 2),
], ids=["j2_3d", "neohookean_3d", "elastic_2d"])
def test_reference_reader_decodes_our_archives(tmp_path, xml, dim):
    prob = nm.Problem.from_xml(xml)
    ora = orc.Oracle.from_problem(prob)
    ora.step(20)
    d, iv = ora.particles()
    g, mats, _, n, _ = prob.describe()
    _, _, thick = prob.archive_settings()
    path = os.fspath(tmp_path / "run.mpm.20")          # archives end in .<step>
    archive.write(path, dim, ORDER, 1.25, d, iv, mats, thickness=thick)

    axes = "xyz"[:dim]
    comps = ["xx", "yy", "zz", "xy"] + (["xz", "yz"] if dim == 3 else [])
    qs = (["mass", "mat"] + ["vel" + a for a in axes] + ["s" + c for c in comps] + ["e" + c for c in comps]
          + ["pe" + c for c in comps] + ["hist1"] + ["disp" + a for a in axes])
    cols = _extract(path, qs, dim, os.fspath(tmp_path / "out"))
    assert cols.shape[1] == n
    col = dict(zip(list(axes) + qs, cols))

    for a in axes:
        assert np.array_equal(col[a], d[F["pos_" + a]]), a
        assert np.array_equal(col["vel" + a], d[F["vel_" + a]]), a
        assert np.array_equal(col["disp" + a], d[F["pos_" + a]] - d[F["orig_" + a]]), a
    assert np.array_equal(col["mass"], d[F["mass"]])
    assert np.array_equal(col["mat"], (iv[1] + 1).astype(np.float64))
    mat = mats[0]
    P = d[F["pressure"]] if mat.type in (nm.MAT_ISOPLASTICITY, nm.MAT_NEOHOOKEAN) else 0.
    rho = mat.rho / d[F["hist0"]] if mat.type == nm.MAT_NEOHOOKEAN else mat.rho
    for c in comps:
        sp = d[F["sp_" + c]] - (P if c in ("xx", "yy", "zz") else 0.)
        assert np.array_equal(col["s" + c], rho * sp), c
        assert np.array_equal(col["e" + c], d[F["ep_" + c]]), c
        assert np.array_equal(col["pe" + c], d[F["alt_" + c]]), c
    hist = d[F["hist0"]] if mat.type in (nm.MAT_ISOPLASTICITY, nm.MAT_NEOHOOKEAN) else np.zeros(n)
    assert np.array_equal(col["hist1"], hist)
    assert np.any(col["sxx"] != 0.)


def test_reference_reader_minimal_order(tmp_path):
    """defaults and velocity only (the smallest archive order): record padding to the crack record"""
    prob = nm.Problem.from_xml(configs.elastic_block_2d(cells=16, block=8))
    d, iv = prob.particles()
    g, mats, _, n, _ = prob.describe()
    path = os.fspath(tmp_path / "start.mpm.0")
    archive.write(path, 2, "iYY", 0., d, iv, mats)
    cols = _extract(path, ["velx", "vely", "mass"], 2, os.fspath(tmp_path / "o"))
    assert cols.shape[1] == n
    assert np.array_equal(cols[0], d[F["pos_x"]]) and np.array_equal(cols[1], d[F["pos_y"]])
    assert np.array_equal(cols[2], d[F["vel_x"]]) and np.array_equal(cols[4], d[F["mass"]])
