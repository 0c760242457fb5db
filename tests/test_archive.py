"""NairnMPM archive files (format version 6): record sizes by CalcArchiveSize's rules
(ArchiveData.cpp:266-424), the 64-byte header (SetArchiveHeader, :443-470), records written by
the ArchiveResults rules (:721-1000) from an oracle-stepped state and read back, the XML archive
settings, and refusal of items this implementation does not carry."""
import math
import os

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import archive, configs
import oracle as orc

F = nm.F
ORDER = "iYYYYYNNNNNYNY"          # velocity, stress, strain, plastic strain, shear components, history


@pytest.mark.parametrize("order, dim, size", [
    ("iY", 3, 120),                   # defaults 88 B < the crack record (16 + 8 + 12 doubles)
    ("iY", 2, 88),                    # defaults 64 B < the crack record (16 + 8 + 8 doubles)
    ("iYY", 3, 120),                  # + velocity 24 = 112, still the crack record
    ("iYYYY", 3, 88 + 24 + 48 + 48 + 24),          # + stress, strain and its rotation angles
    ("iYYYY", 2, 64 + 16 + 32 + 32 + 8),
    (ORDER, 3, 88 + 24 + 48 + 48 + 48 + 16 + 8 + 24),
    (ORDER, 2, 64 + 16 + 32 + 32 + 32 + 16 + 8 + 8),
    ("iYNNNNNNNNNNN" + chr(0x33), 3, 120),            # histories 1 and 2 by bit flags: 88 + 16 < 120
    ("iYYYYYNNNNNNN" + chr(0x3F), 3, 88 + 24 + 48 + 48 + 48 + 4 * 8 + 24),
])
def test_record_sizes(order, dim, size):
    assert archive.record_size(order, dim) == size


def test_order_normalisation():
    o = archive.normalize_order("iYYYN")
    assert len(o) == 25
    assert o[:5] == "iYYYN" and o[17] == "N"           # rotation angles follow the strain item
    assert archive.normalize_order("mNNNYNYNNNY")[:11] == "iYNNYNNNNNN"   # byte order, defaults, retired items
    with pytest.raises(nm.MPMError):
        archive.normalize_order("iYNNNNNY")            # work energy is not carried
    with pytest.raises(nm.MPMError):
        archive.normalize_order("iY" + "N" * 22 + "Y")  # particle size is not carried


def _stepped(xml, steps=20):
    prob = nm.Problem.from_xml(xml)
    o = orc.Oracle.from_problem(prob)
    o.step(steps)
    d, i = o.particles()
    g, mats, _, _, _ = prob.describe()
    return prob, g, mats, d, i


@pytest.mark.parametrize("xml, dim", [
    (configs.j2_block_3d(cells=12, lo=3, hi=9, vx=2e5), 3),
    (configs.neohookean_block_3d(cells=12, lo=3, hi=9), 3),
    (configs.elastic_block_2d(cells=16, block=8), 2),
], ids=["j2_3d", "neohookean_3d", "elastic_2d"])
def test_archive_records(tmp_path, xml, dim):
    prob, g, mats, d, i = _stepped(xml)
    _, _, thick = prob.archive_settings()
    path = os.fspath(tmp_path / "run.1")
    archive.write(path, dim, ORDER, 1.25, d, i, mats, thickness=thick)
    n = d.shape[1]
    rec = archive.record_size(ORDER, dim)
    assert os.path.getsize(path) == 64 + n * rec

    with open(path, "rb") as fh:
        head = fh.read(64)
    order = archive.normalize_order(ORDER)
    expect = b"ver6" + bytes([25]) + order.encode() + bytes([7]) + b"iYNNNNN" + str(dim).encode() + b"1"
    assert head[:len(expect)] == expect
    assert np.frombuffer(head[len(expect):len(expect) + 4], "<f4")[0] == np.float32(1.25)
    assert head[len(expect) + 4:] == bytes(64 - len(expect) - 4)

    h, r = archive.read(path)
    assert h["dim"] == dim and h["mpm_order"] == order and h["structured"] and len(r) == n
    assert np.array_equal(r["elem"], i[0] + 1) and np.array_equal(r["mat"], i[1] + 1)
    assert np.array_equal(r["mass"], d[F["mass"]])
    axes = "xyz"[:dim]
    for a in axes:
        assert np.array_equal(r["pos_" + a], d[F["pos_" + a]])
        assert np.array_equal(r["orig_" + a], d[F["orig_" + a]])
        assert np.array_equal(r["vel_" + a], d[F["vel_" + a]])
    # rotation angle about z in degrees from the rotation strain: 180 (0 - wxy/2) / pi
    assert np.array_equal(r["angle_z"], 180. * (0. - 0.5 * d[F["wr_xy"]]) / math.pi)
    if dim == 2:
        assert np.array_equal(r["thickness"], thick)
    # stress in Pa: current density times specific stress, the pressure folded in for J2 and
    # hyperelastic materials, whose density is rho0/J (hyperelastic) or rho0 (J2, elastic)
    mat = mats[0]
    P = d[F["pressure"]] if mat.type in (nm.MAT_ISOPLASTICITY, nm.MAT_NEOHOOKEAN) else 0.
    rho = mat.rho / d[F["hist0"]] if mat.type == nm.MAT_NEOHOOKEAN else mat.rho
    comps = ["xx", "yy", "zz", "xy"] + (["xz", "yz"] if dim == 3 else [])
    for c in comps:
        sp = d[F["sp_" + c]] - (P if c in ("xx", "yy", "zz") else 0.)
        assert np.array_equal(r["stress_" + c], rho * sp), c
        assert np.array_equal(r["strain_" + c], d[F["ep_" + c]]), c
        assert np.array_equal(r["plastic_" + c], d[F["alt_" + c]]), c
    assert np.array_equal(r["dudy"], 0.5 * (d[F["ep_xy"]] - d[F["wr_xy"]]))
    assert np.array_equal(r["dvdx"], 0.5 * (d[F["ep_xy"]] + d[F["wr_xy"]]))
    hist = d[F["hist0"]] if mat.type in (nm.MAT_ISOPLASTICITY, nm.MAT_NEOHOOKEAN) else np.zeros(n)
    assert np.array_equal(r["history1"], hist)
    assert np.all(r["angle0_z"] == 0.)
    if mat.type == nm.MAT_ISOPLASTICITY:
        assert np.any(r["history1"] > 0.)               # the block has yielded: alpha is archived


def test_xml_archive_settings():
    extra = """    <ArchiveTime units="ms">0.5</ArchiveTime>
    <MPMArchiveOrder>iYYYNN</MPMArchiveOrder>
    <MPMArchive result="plasticstrain" setting="Y"/>
    <MPMArchive result="history2" setting="Y"/>
    <MPMArchive result="history1" setting="Y"/>
    <MPMArchive result="history2" setting="N"/>
    <MPMArchive result="stress" setting="N"/>
"""
    xml = configs.elastic_block_2d(cells=16, block=8).replace("  </MPMHeader>", extra + "  </MPMHeader>", 1)
    prob = nm.Problem.from_xml(xml)
    order, t, thick = prob.archive_settings()
    assert t == pytest.approx(0.5e-3)
    # MPMReadHandler.cpp:769-793: history bits on '0' (0x30), history1 only -> '1'
    assert order[:6] == "iYYNNY" and order[13] == "1"
    assert archive.record_size(order, 2) == 64 + 16 + 32 + 8    # velocity, plastic strain, history 1
    assert np.all(thick == 1.)
