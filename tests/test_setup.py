"""Problem set-up on the host: the XML reader, grid construction, particle generation, masses,
material constants and the CFL time step, checked against values worked out by hand from the
reference's rules (Read_MPM/Generators.cpp grid padding, MPMBase volume, CFLTimeStep)."""
import math

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs

F = nm.F


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def test_elastic_block_2d_setup():
    prob = nm.Problem.from_xml(configs.elastic_block_2d())
    g, mats, run, n, tmax = prob.describe()
    # 64 cells of 1 mm plus one padding element on each side
    assert (g.nx, g.ny, g.nz) == (66, 66, 0)
    assert (g.xmin, g.ymin, g.dx, g.dy) == (-1., -1., 1., 1.)
    assert n == 32 * 32 * 4
    d, iv = prob.particles()
    # mass = rho * 4 psx psy thickness, rho in g/mm^3, lp = 1/2
    np.testing.assert_allclose(d[F["mass"]], 1e-3 * 4 * 0.25 * 0.25 * 1.0, rtol=1e-15)
    # element of each particle holds it
    ex = iv[0] % g.nx
    ey = iv[0] // g.nx
    assert ((g.xmin + ex * g.dx <= d[F["pos_x"]]) & (d[F["pos_x"]] < g.xmin + (ex + 1) * g.dx)).all()
    assert ((g.ymin + ey * g.dy <= d[F["pos_y"]]) & (d[F["pos_y"]] < g.ymin + (ey + 1) * g.dy)).all()
    assert (iv[1] == 0).all()
    assert run.shape == nm.SHAPE_UGIMP
    lam, G = lame(1000e6, 0.33)
    cp = math.sqrt((lam + 2 * G) / 1e-3)
    np.testing.assert_allclose(mats[0].waveSpeed, cp, rtol=1e-14)
    np.testing.assert_allclose(run.timestep, 0.2 * 1.0 / cp, rtol=1e-14)
    assert run.strainFirst == 0.5 * run.timestep
    assert run.strainLast == run.timestep - run.strainFirst
    assert tmax == 1.0


def test_j2_block_3d_setup():
    prob = nm.Problem.from_xml(configs.j2_block_3d(cells=32, lo=4, hi=28))
    g, mats, run, n, _ = prob.describe()
    assert (g.nx, g.ny, g.nz) == (34, 34, 34)
    assert n == 24 ** 3 * 8
    d, iv = prob.particles()
    # 2x2x2 particles per element at the element's quarter points, lp = 1/2
    frac = np.mod(d[F["pos_x"]] - g.xmin, g.dx) / g.dx
    assert set(np.round(frac, 12)) == {0.25, 0.75}
    np.testing.assert_allclose(d[F["mass"]], 2.7e-3 * 8 * 0.25 ** 3, rtol=1e-15)
    # the two halves collide at +-50000 mm/s
    mid = 16.
    assert (d[F["vel_x"]][d[F["pos_x"]] < mid] == 50000.).all()
    assert (d[F["vel_x"]][d[F["pos_x"]] > mid] == -50000.).all()
    m = mats[0]
    assert m.type == nm.MAT_ISOPLASTICITY
    lam, G = lame(70000e6, 0.33)
    np.testing.assert_allclose(m.Gred, G / 2.7e-3, rtol=1e-13)
    np.testing.assert_allclose(m.Kred, (lam + 2 * G / 3) / 2.7e-3, rtol=1e-13)
    np.testing.assert_allclose(m.yldred, 300e6 / 2.7e-3, rtol=1e-14)
    np.testing.assert_allclose(run.timestep, 0.5 * 1.0 / math.sqrt((lam + 2 * G) / 2.7e-3), rtol=1e-14)


XML_UNITS = """<?xml version='1.0'?>
<!DOCTYPE JANFEAInput [
<!ENTITY len "4">
<!ENTITY rho "1.5">
]>
<JANFEAInput version='3'>
  <!-- comment with <tags> inside -->
  <Header><Analysis>12</Analysis></Header>
  <MPMHeader>
    <MPMMethod>2</MPMMethod>
    <MatlPtsPerElement>1</MatlPtsPerElement>
    <TimeStep units="ms">1e-6</TimeStep>
    <TimeFactor>0.3</TimeFactor>
    <MaxTime units="ms">2</MaxTime>
    <GIMP type="uGIMP"/>
  </MPMHeader>
  <Mesh>
    <Grid xmin="0" xmax="&len;" ymin="0" ymax="&len;" zmin="0" zmax="&len;">
      <Horiz cellsize="1"/><Vert cellsize="1"/><Depth cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body matname="gel" vx="2" vy="-1">
      <Box xmin="1" xmax="3" ymin="1" ymax="3" zmin="1" zmax="2"/>
    </Body>
  </MaterialPoints>
  <Material Type="28" Name="gel">
    <rho>&rho;</rho><G><![CDATA[2]]></G><K>10</K><UJOption>1</UJOption>
  </Material>
</JANFEAInput>
"""


def test_xml_entities_units_and_late_material():
    prob = nm.Problem.from_xml(XML_UNITS)
    g, mats, run, n, tmax = prob.describe()
    assert (g.nx, g.ny, g.nz) == (6, 6, 6)
    assert n == 2 * 2 * 1                          # one point per element
    assert tmax == 2e-3
    m = mats[0]
    assert m.type == nm.MAT_NEOHOOKEAN and m.ujOption == 1
    np.testing.assert_allclose(m.rho, 1.5e-3, rtol=1e-15)
    np.testing.assert_allclose(m.Gsp, 2e6 / 1.5e-3, rtol=1e-14)
    d, iv = prob.particles()
    assert (d[F["vel_x"]] == 2.).all() and (d[F["vel_y"]] == -1.).all()
    np.testing.assert_allclose(d[F["pos_z"]], 1.5)
    assert (d[F["hist0"]] == 1.).all() and (d[F["alt_xx"]] == 1.).all()
    # the user's step 1e-6 ms = 1e-9 s is below the CFL limit, so it is kept
    assert run.timestep == pytest.approx(1e-9, rel=1e-15)


@pytest.mark.parametrize("bad, msg", [
    ("<GridBCs><DisBC/></GridBCs>", "not on this path"),
    ("<CrackList><pt/></CrackList>", "not on this path"),
])
def test_xml_rejects_features_off_the_path(bad, msg):
    text = XML_UNITS.replace("</JANFEAInput>", bad + "\n</JANFEAInput>")
    with pytest.raises(nm.MPMError, match=msg):
        nm.Problem.from_xml(text)


def test_xml_errors_raise():
    with pytest.raises(nm.MPMError):
        nm.Problem.from_xml("<JANFEAInput><Header><Analysis>12</Analysis></Header>")   # unterminated
    with pytest.raises(nm.MPMError, match="Grid"):
        nm.Problem.from_xml(XML_UNITS.replace("<Grid", "<Grd").replace("</Grid>", "</Grd>"))
    with pytest.raises(nm.MPMError, match="Analysis"):
        nm.Problem.from_xml(XML_UNITS.replace("<Analysis>12<", "<Analysis>1<"))


def test_programmatic_problem_matches_xml():
    """The Problem API (set_grid/add_material/add_body) builds the same particles as the XML."""
    a = nm.Problem.from_xml(configs.j2_block_3d(cells=8, lo=2, hi=6))
    da, ia = a.particles()
    b = nm.Problem.create(nm.THREED)
    b.set_grid([0, 8, 0, 8, 0, 8], [1, 1, 1])
    b.set("CFL", 0.5)
    mid = b.add_material(nm.MAT_ISOPLASTICITY, "block", E=70000., nu=0.33, rho=2.7, **{"yield": 300., "Ep": 1000.})
    b.add_body(mid, "Box", [2, 4, 2, 6, 2, 6], vel=(50000., 0., 0.))
    b.add_body(mid, "Box", [4, 6, 2, 6, 2, 6], vel=(-50000., 0., 0.))
    db, ib = b.particles()
    assert da.shape == db.shape
    oa = np.lexsort(da[:3][::-1])
    ob = np.lexsort(db[:3][::-1])
    np.testing.assert_array_equal(da[:, oa], db[:, ob])
    np.testing.assert_array_equal(ia[:2, oa], ib[:2, ob])


@pytest.mark.parametrize("gimp, expect", [("uGIMP", 10 * 10 * 4), ("Classic", 12 * 12 * 4), ("lCPDI", 10 * 10 * 4)])
def test_gimp_border_layer_takes_no_particles(gimp, expect):
    """A body larger than the grid fills the one-cell padding layer only with classic shape
    functions: GIMP and CPDI mark the border elements as holes (Generators.cpp:1607-1619)."""
    xml = configs._header(10, 0.5, 1.0, "    <MatlPtsPerElement>4</MatlPtsPerElement>\n", shape=gimp) + """  <Mesh output="file">
    <Grid xmin="0" xmax="10" ymin="0" ymax="10">
      <Horiz cellsize="1"/>
      <Vert cellsize="1"/>
    </Grid>
  </Mesh>
  <MaterialPoints>
    <Body mat="1" vx="0" vy="0">
      <Rect xmin="-5" xmax="15" ymin="-5" ymax="15"/>
    </Body>
  </MaterialPoints>
  <Material Type="1" Name="e">
    <rho>1.0</rho>
    <E>1000</E>
    <nu>0.3</nu>
  </Material>
</JANFEAInput>
"""
    prob = nm.Problem.from_xml(xml)
    g, _, run, n, _ = prob.describe()
    assert (g.nx, g.ny) == (12, 12)
    assert n == expect
    d, _ = prob.particles()
    if gimp != "Classic":
        assert d[nm.F["pos_x"]].min() > 0. and d[nm.F["pos_x"]].max() < 10.
