"""Archive of a device simulation: the file written from the context equals the one written from
its downloaded particles, and the records hold the stepped state."""
import os

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import archive, configs

pytestmark = pytest.mark.gpu

ORDER = "iYYYYYNNNNNYNY"


@pytest.mark.parametrize("xml, dim", [
    (configs.j2_block_3d(cells=12, lo=3, hi=9), 3),
    (configs.elastic_block_2d(cells=16, block=8), 2),
], ids=["j2_3d", "elastic_2d"])
def test_device_archive_matches_download(tmp_path, xml, dim):
    prob = nm.Problem.from_xml(xml)
    g, mats, _, _, _ = prob.describe()
    _, _, thick = prob.archive_settings()
    sim = nm.Simulation.from_problem(prob)
    sim.step(30)
    a = os.fspath(tmp_path / "device.30")
    b = os.fspath(tmp_path / "host.30")
    sim.write_archive(a, ORDER, time_ms=0.03, thickness=thick)
    d, i = sim.particles()
    archive.write(b, dim, ORDER, 0.03, d, i, mats, thickness=thick)
    with open(a, "rb") as fa, open(b, "rb") as fb:
        assert fa.read() == fb.read()
    _, r = archive.read(a)
    assert np.array_equal(r["pos_x"], d[nm.F["pos_x"]])
    assert np.array_equal(r["vel_y"], d[nm.F["vel_y"]])
    sim.close()


EXTRACT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ExtractMPM")


@pytest.mark.skipif(not os.path.exists(EXTRACT), reason="the reference's ExtractMPM was not built with this tree")
def test_device_archive_through_the_reference_reader(tmp_path):
    """the file written from device rows, decoded by NairnMPM's ExtractMPM (built from its source)"""
    import subprocess
    prob = nm.Problem.from_xml(configs.j2_block_3d(cells=12, lo=3, hi=9))
    g, mats, _, n, _ = prob.describe()
    sim = nm.Simulation.from_problem(prob)
    sim.step(30)
    path = os.fspath(tmp_path / "device.30")
    sim.write_archive(path, ORDER, time_ms=0.03)
    d, i = sim.particles()
    sim.close()
    out = os.fspath(tmp_path / "x")
    subprocess.run([EXTRACT, "-d", "-q", "velx", "-q", "sxy", "-q", "hist1", "-o", out, path], check=True,
                   capture_output=True)
    cols = np.fromfile(out + ".data", dtype="<f8").reshape(-1, 6).T
    assert np.array_equal(cols[0], d[nm.F["pos_x"]]) and np.array_equal(cols[2], d[nm.F["pos_z"]])
    assert np.array_equal(cols[3], d[nm.F["vel_x"]])
    assert np.array_equal(cols[4], mats[0].rho * d[nm.F["sp_xy"]])
    assert np.array_equal(cols[5], d[nm.F["hist0"]])
