"""Traction conditions (MatPtTractionBC): the <TractionBC> conditions of <ParticleBCs> as the
reference's reader builds them (Read_MPM/Generators.cpp:1105-1202, MatPtLoadBC.cpp:202-206), and
the oracle's restatement of AddMPFluxBC (MatPtTractionBC.cpp:64-227) with GetSurfaceInfo
(MatPoint2D.cpp:1138-1300, MatPoint3D.cpp:628-851) pinned by geometry and mechanics:

* Newton's second law: the internal forces cancel over the grid, so the total momentum of a free
  body grows each step by the net traction force (the stress times the face length and thickness,
  or area) times the time step, in the direction the face and dir name;
* hand-computed bilinear weights of the corners' elements for a single particle, including the
  share that is dropped at a node where the particle's material field has no points (classic
  particles) and kept where it has (uGIMP);
* face geometry from first principles: the four corners of each 3D face lie on that face and run
  around it, and the weighted normal of a deformed lCPDI face follows Nanson's formula
  (n da = J F^-T N dA), while uGIMP particles keep the undeformed corners and axis normal with
  the deformed face size."""
import math

import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def _one_particle(dim, pos, shape="uGIMP", thickness=None, cells=12, lp=0.5):
    """one particle of an elastic material on a grid of 1 mm cells"""
    pos = [float(v) for v in pos]
    thick = f' thickness="{thickness!r}"' if thickness is not None else ""
    if dim == 2:
        grid = (f'    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}"{thick}>\n'
                '      <Horiz cellsize="1"/>\n      <Vert cellsize="1"/>\n    </Grid>\n')
        mp = (f'    <mp matl="1" lpx="{lp!r}" lpy="{lp!r}"><pt x="{pos[0]!r}" y="{pos[1]!r}"/>'
              '<vel x="0.0" y="0.0"/></mp>\n')
    else:
        grid = (f'    <Grid xmin="0" xmax="{cells}" ymin="0" ymax="{cells}" zmin="0" zmax="{cells}">\n'
                '      <Horiz cellsize="1"/>\n      <Vert cellsize="1"/>\n      <Depth cellsize="1"/>\n'
                '    </Grid>\n')
        mp = (f'    <mp matl="1" lpx="{lp!r}" lpy="{lp!r}" lpz="{lp!r}"><pt x="{pos[0]!r}" y="{pos[1]!r}" '
              f'z="{pos[2]!r}"/><vel x="0.0" y="0.0" z="0.0"/></mp>\n')
    return (configs._header(10 if dim == 2 else 12, 0.2, 1.0, shape=shape) +
            '  <Mesh output="file">\n' + grid + '  </Mesh>\n  <MaterialPoints>\n    <PointList>\n' + mp +
            '    </PointList>\n  </MaterialPoints>\n  <Material Type="1" Name="m">\n    <rho>1.0</rho>\n'
            '    <E>1000.0</E>\n    <nu>0.3</nu>\n  </Material>\n</JANFEAInput>\n')


def _tractions_xml(xml, dim, pos, bcs):
    if dim == 2:
        shape = ("BCLine", {"x1": pos[0] - 0.1, "y1": pos[1], "x2": pos[0] + 0.1, "y2": pos[1]}, bcs)
    else:
        shape = ("BCBox", {"xmin": pos[0] - 0.1, "xmax": pos[0] + 0.1, "ymin": pos[1] - 0.1, "ymax": pos[1] + 0.1,
                           "zmin": pos[2] - 0.1, "zmax": pos[2] + 0.1}, bcs)
    return configs.traction_bcs(xml, [shape])


def _traction_forces(xml_with, xml_without):
    """ftot of field 0 after the post-forces task minus the same without the tractions, as a dict
    from node coordinates (the reader may pad the grid) to the force vector"""
    out = []
    for xml in (xml_with, xml_without):
        o = orc.Oracle.from_problem(nm.Problem.from_xml(xml))
        for t in range(6):
            o.run_task(t)
        out.append(o.grid_values()[2])
    f = out[0] - out[1]
    g = nm.Problem.from_xml(xml_with).describe()[0]
    nnx, nny = g.nx + 1, g.ny + 1
    three = "<Analysis>12</Analysis>" in xml_with
    nodes = {}
    for n in np.nonzero(np.any(f != 0., axis=0))[0]:
        i, j, k = n % nnx, (n // nnx) % nny, n // (nnx * nny)
        key = (round(g.xmin + i * g.dx, 9), round(g.ymin + j * g.dy, 9))
        nodes[key + ((round(g.zmin + k * g.dz, 9),) if three else ())] = f[:, n]
    return nodes


# ---------------------------------------------------------------------------------------------
# the reader

def test_traction_sets_units_and_order():
    base = configs.elastic_block_2d(cells=16, block=8, jitter=0., vmax=0.)
    base = base.replace('<Grid xmin="0" xmax="16" ymin="0" ymax="16">',
                        '<Grid xmin="0" xmax="16" ymin="0" ymax="16" thickness="2.5">')
    xml = configs.traction_bcs(base, [
        ("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75, "tolerance": 0.1},
         ["net", {"dir": 11, "face": 3, "stress": 2.0}, {"dir": 1, "face": 3, "style": 2, "stress": 0.5, "time": 1.}]),
        ("BCLine", {"x1": 4.25, "y1": 4, "x2": 4.25, "y2": 12, "tolerance": 0.1},
         [{"dir": 12, "face": 4, "style": 3, "stress": -1.5, "ftime": 2.}]),
    ])
    prob = nm.Problem.from_xml(xml)
    d, _ = prob.particles()
    tr = prob.particle_tractions()
    assert not prob.particle_loads()
    top = np.nonzero(np.abs(d[F["pos_y"]] - 11.75) <= 0.1)[0]
    left = np.nonzero(np.abs(d[F["pos_x"]] - 4.25) <= 0.1)[0]
    assert len(tr) == 2 * 16 + 16
    # list order: condition by condition, particles in mpm[] order; <net/> does not divide
    assert [t.particle for t in tr[:16]] == top.tolist() == [t.particle for t in tr[16:32]]
    assert [t.particle for t in tr[32:]] == left.tolist()
    # MPa -> Pa; linear: per ms -> per s and ms -> s; sine: 1/ms -> 1/s; the grid's thickness
    assert all(t.dir == 11 and t.face == 3 and t.style == 1 and t.value == 2e6 for t in tr[:16])
    assert all(t.dir == 1 and t.style == 2 and t.value == 0.5e6 * 1e3 and t.ftime == 1e-3 for t in tr[16:32])
    assert all(t.dir == 12 and t.face == 4 and t.style == 3 and t.value == -1.5e6 and t.ftime == 2000. for t in tr[32:])
    assert all(t.thickness == 2.5 for t in tr)


def test_traction_box_3d():
    xml = configs.traction_bcs(configs.j2_block_3d(cells=12, lo=3, hi=9), [
        ("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 8.5, "zmax": 9}, [{"dir": 11, "face": 6, "stress": 3.}]),
    ])
    prob = nm.Problem.from_xml(xml)
    d, _ = prob.particles()
    tr = prob.particle_tractions()
    inside = np.nonzero(d[F["pos_z"]] >= 8.5)[0]
    assert [t.particle for t in tr] == inside.tolist()
    assert all(t.value == 3e6 and t.face == 6 and t.dir == 11 for t in tr)


@pytest.mark.parametrize("dim, child, msg", [
    (2, {"dir": 3, "face": 1, "stress": 1.}, "must be 1, 2, 11, or 12 for 2D"),
    (2, {"dir": 13, "face": 1, "stress": 1.}, "must be 1, 2, 11, or 12 for 2D"),
    (2, {"dir": 1, "face": 5, "stress": 1.}, "must be 1 to 4 for 2D"),
    (2, {"dir": 1, "face": 1, "style": 5, "stress": 1.}, "cannot use silent style"),
    (2, {"dir": 1, "face": 1, "style": 6, "stress": 1.}, "styles other than 1 to 4"),
    (2, {"dir": 1, "face": 1, "function": "t"}, "function"),
    (3, {"dir": 12, "face": 1, "stress": 1.}, "must be 1-3 or 11 for 3D"),
    (3, {"dir": 1, "face": 7, "stress": 1.}, "must be 1 to 6 for 3D"),
])
def test_traction_refusals(dim, child, msg):
    if dim == 2:
        xml = configs.traction_bcs(configs.elastic_block_2d(cells=16, block=8),
                                   [("BCLine", {"x1": 4, "y1": 4, "x2": 12, "y2": 4}, [child])])
    else:
        xml = configs.traction_bcs(configs.j2_block_3d(cells=12, lo=3, hi=9),
                                   [("BCBox", {"xmin": 3, "xmax": 9, "ymin": 3, "ymax": 9, "zmin": 8.5, "zmax": 9}, [child])])
    with pytest.raises(nm.MPMError, match=msg):
        nm.Problem.from_xml(xml)


def test_exact_tractions_refused():
    xml = configs.elastic_block_2d(cells=16, block=8).replace("</MPMHeader>", "  <ExactTractions/>\n  </MPMHeader>")
    with pytest.raises(nm.MPMError, match="ExactTractions"):
        nm.Problem.from_xml(xml)


# ---------------------------------------------------------------------------------------------
# Newton's second law on a free block

def _momentum(o, dim):
    d, _ = o.particles()
    return np.array([np.sum(d[F["mass"]] * d[F["vel_" + a]]) for a in "xyz"[:dim]])


# (face, dir) -> unit direction of the force for a positive stress on an undeformed block
FACES_2D = {(1, 11): (0., -1.), (2, 11): (1., 0.), (3, 11): (0., 1.), (4, 11): (-1., 0.),
            (1, 12): (1., 0.), (2, 12): (0., 1.), (3, 12): (-1., 0.), (4, 12): (0., -1.),
            (3, 1): (1., 0.), (1, 2): (0., 1.)}


@pytest.mark.parametrize("face, dir", sorted(FACES_2D))
@pytest.mark.parametrize("shape", ["uGIMP", "lCPDI", "Classic"])
def test_momentum_follows_the_tractions_2d(face, dir, shape):
    # a block at rest, one row or column loaded on its outer face: the first step's momentum
    # change is stress * (8 mm) * thickness * dt along the face's direction
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0., jitter=0., shape=shape)
    base = base.replace('<Grid xmin="0" xmax="16" ymin="0" ymax="16">',
                        '<Grid xmin="0" xmax="16" ymin="0" ymax="16" thickness="3.0">')
    line = {1: (4, 4.25, 12, 4.25), 2: (11.75, 4, 11.75, 12), 3: (4, 11.75, 12, 11.75), 4: (4.25, 4, 4.25, 12)}[face]
    stress = 2.5
    xml = configs.traction_bcs(base, [("BCLine", dict(zip(("x1", "y1", "x2", "y2"), line)),
                                       [{"dir": dir, "face": face, "stress": stress}])])
    prob = nm.Problem.from_xml(xml)
    assert len(prob.particle_tractions()) == 16
    _, _, run, _, _ = prob.describe()
    o = orc.Oracle.from_problem(prob)
    o.step(1)
    p = _momentum(o, 2)
    expect = stress * 1e6 * 8. * 3.0 * run.timestep * np.array(FACES_2D[(face, dir)])
    assert p == pytest.approx(expect, rel=1e-12, abs=1e-12 * np.abs(expect).max())


FACES_3D = {1: (1, -1), 2: (0, 1), 3: (1, 1), 4: (0, -1), 5: (2, -1), 6: (2, 1)}   # axis, sign of the outward normal


@pytest.mark.parametrize("face", sorted(FACES_3D))
@pytest.mark.parametrize("shape", ["uGIMP", "lCPDI"])
def test_momentum_follows_the_tractions_3d(face, shape):
    axis, sign = FACES_3D[face]
    lo, hi = 3, 9
    box = {"xmin": lo, "xmax": hi, "ymin": lo, "ymax": hi, "zmin": lo, "zmax": hi}
    a = "xyz"[axis]
    box[a + "min"], box[a + "max"] = (hi - 0.5, hi) if sign > 0 else (lo, lo + 0.5)
    stress = 1.5
    xml = configs.traction_bcs(configs.j2_block_3d(cells=12, lo=lo, hi=hi, vx=0., shape=shape), [
        ("BCBox", box, [{"dir": 11, "face": face, "stress": stress}, {"dir": 3, "face": face, "stress": -0.5}])])
    prob = nm.Problem.from_xml(xml)
    assert len(prob.particle_tractions()) == 2 * 12 * 12
    _, _, run, _, _ = prob.describe()
    o = orc.Oracle.from_problem(prob)
    o.step(1)
    p = _momentum(o, 3)
    expect = np.zeros(3)
    expect[axis] += sign * stress
    expect[2] += -0.5
    expect *= 1e6 * 36. * run.timestep
    assert p == pytest.approx(expect, rel=1e-12, abs=1e-12 * np.abs(expect).max())


def test_momentum_follows_a_linear_traction():
    # BCValue at each step's start time: zero before ftime, then rate * (t - ftime)
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0., jitter=0.)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    dt = run.timestep
    ftime_ms = 2.5 * dt * 1e3
    xml = configs.traction_bcs(base, [("BCLine", {"x1": 4, "y1": 11.75, "x2": 12, "y2": 11.75},
                                       [{"dir": 2, "face": 3, "style": 2, "stress": 7., "ftime": ftime_ms}])])
    o = orc.Oracle.from_problem(nm.Problem.from_xml(xml))
    t, expect = 0., 0.
    for step in range(6):
        # the deformation of the first steps changes the face size by O(dt^2): rel 1e-9
        expect += (7e9 * (t - ftime_ms * 1e-3) if t >= ftime_ms * 1e-3 else 0.) * 8. * dt
        o.step(1)
        py = _momentum(o, 2)[1]
        assert py == pytest.approx(expect, rel=1e-9, abs=1e-6), step
        t += dt
    assert expect > 0.


# ---------------------------------------------------------------------------------------------
# the corners' elements and their linear weights

@pytest.mark.parametrize("shape, dropped", [("uGIMP", False), ("Classic", True)])
def test_single_particle_node_forces(shape, dropped):
    # particle at (5.9, 5.5), semi size 0.25, top face: corners (6.15, 5.75) in cell (6, 5) and
    # (5.65, 5.75) in cell (5, 5). Bilinear weights: x 0.85 / 0.15 on nodes 6 / 7 and 0.35 / 0.65
    # on nodes 5 / 6; y 0.25 / 0.75 on rows 5 / 6. A classic particle puts mass on nodes 5 and 6
    # only, so the share of node 7 is dropped (no points there); a uGIMP particle reaches node 7.
    pos = (5.9, 5.5)
    base = _one_particle(2, pos, shape=shape, thickness=2.0)
    xml = _tractions_xml(base, 2, pos, [{"dir": 2, "face": 3, "stress": 4.}])
    got = _traction_forces(xml, base)
    s = 4e6 * 0.25 * 2.0           # stress * half face length * thickness
    expect = {(5, 5): 0.35 * 0.25, (6, 5): (0.65 + 0.85) * 0.25, (5, 6): 0.35 * 0.75, (6, 6): (0.65 + 0.85) * 0.75}
    if not dropped:
        expect[(7, 5)] = 0.15 * 0.25
        expect[(7, 6)] = 0.15 * 0.75
    assert set(got) == set(expect)
    for k, w in expect.items():
        assert got[k][1] == pytest.approx(s * w, rel=1e-14), k
        assert got[k][0] == 0. and got[k][2] == 0.
    assert sum(v[1] for v in got.values()) == pytest.approx(s * (1.85 if dropped else 2.0), rel=1e-14)


def test_single_particle_3d_corner_on_a_node():
    # particle at (5.25, 5.25, 5.75), semi sizes 0.25: the +z face at z = 6 has corners
    # (5, 5, 6), (5.5, 5, 6), (5.5, 5.5, 6), (5, 5.5, 6): the first sits on node (5, 5, 6) and
    # gives it its whole share; node weights in x and y are 1 (x = 5) or 0.5 / 0.5 (x = 5.5)
    pos = (5.25, 5.25, 5.75)
    base = _one_particle(3, pos)
    xml = _tractions_xml(base, 3, pos, [{"dir": 3, "face": 6, "stress": 2.}])
    got = _traction_forces(xml, base)
    s = 2e6 * 0.25 * 0.25          # a quarter of the face area
    # per node the sum over the corners of the x weight times the y weight
    corners = [(1., 1.), (0.5, 1.), (0.5, 0.5), (1., 0.5)]        # weights of nodes x = 5, y = 5
    expect = {}
    for i in (5, 6):
        for j in (5, 6):
            w = sum((a if i == 5 else 1. - a) * (b if j == 5 else 1. - b) for a, b in corners)
            if w:
                expect[(i, j, 6)] = s * w
    assert expect[(5, 5, 6)] == s * 2.25 and expect[(6, 6, 6)] == s * 0.25
    assert set(got) == set(expect)
    for k, v in expect.items():
        assert got[k][2] == pytest.approx(v, rel=1e-14), k
        assert got[k][0] == 0. and got[k][1] == 0.
    assert sum(v[2] for v in got.values()) == pytest.approx(4 * s, rel=1e-14)


# ---------------------------------------------------------------------------------------------
# face geometry

def _oracle_with_F(dim, pos, Fm, shape):
    prob = nm.Problem.from_xml(_one_particle(dim, pos, shape=shape))
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    d[F["ep_xx"]] = Fm[0][0] - 1.
    d[F["ep_yy"]] = Fm[1][1] - 1.
    d[F["ep_zz"]] = Fm[2][2] - 1.
    d[F["ep_xy"]] = Fm[1][0] + Fm[0][1]
    d[F["wr_xy"]] = Fm[1][0] - Fm[0][1]
    if dim == 3:
        d[F["ep_xz"]] = Fm[2][0] + Fm[0][2]
        d[F["wr_xz"]] = Fm[2][0] - Fm[0][2]
        d[F["ep_yz"]] = Fm[2][1] + Fm[1][2]
        d[F["wr_yz"]] = Fm[2][1] - Fm[1][2]
    return orc.Oracle(g, mats, run, d, i)


NORMALS_3D = {1: (0, -1, 0), 2: (1, 0, 0), 3: (0, 1, 0), 4: (-1, 0, 0), 5: (0, 0, -1), 6: (0, 0, 1)}


@pytest.mark.parametrize("face", sorted(NORMALS_3D))
def test_undeformed_faces_3d(face):
    pos = np.array([5.5, 5.5, 5.5])
    o = _oracle_with_F(3, pos, np.eye(3), "uGIMP")
    n = np.array(NORMALS_3D[face], dtype=float)
    c, wn = o.surface_info(0, face, 11)
    assert c.shape == (4, 3)
    # the corners lie on the face, at the particle's semi size along the other two axes
    assert np.allclose((c - pos) @ n, 0.25, rtol=0, atol=1e-15)
    assert np.allclose(np.abs(c - pos), 0.25, rtol=0, atol=1e-15)
    # they run around the face: consecutive corners differ along one axis, opposite ones along two
    for k in range(4):
        assert np.count_nonzero(c[(k + 1) % 4] - c[k]) == 1
        assert np.count_nonzero(c[(k + 2) % 4] - c[k]) == 2
    # outward normal times a quarter of the face area (0.5 x 0.5)
    assert np.allclose(wn, 0.0625 * n, rtol=0, atol=1e-16)
    for dof, axis in ((1, 0), (2, 1), (3, 2)):
        _, w = o.surface_info(0, face, dof)
        e = np.zeros(3)
        e[axis] = 0.0625
        assert np.array_equal(w, e)


def _rot3(ax, ang):
    c, s = math.cos(ang), math.sin(ang)
    m = np.eye(3)
    i, j = [(1, 2), (0, 2), (0, 1)][ax]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


DEFORMATIONS = {
    "stretch": np.diag([1.2, 1.5, 0.8]),
    "shear": np.array([[1.0, 0.3, -0.1], [0.05, 1.1, 0.2], [0.0, -0.15, 0.95]]),
    "rotation": _rot3(2, 0.4) @ _rot3(0, -0.3),
}


@pytest.mark.parametrize("face", sorted(NORMALS_3D))
@pytest.mark.parametrize("defo", sorted(DEFORMATIONS))
def test_deformed_faces_3d_nanson(face, defo):
    Fm = DEFORMATIONS[defo]
    pos = np.array([5.5, 5.5, 5.5])
    N = np.array(NORMALS_3D[face], dtype=float)
    J = np.linalg.det(Fm)
    nda = J * np.linalg.solve(Fm.T, N) * 0.0625        # Nanson: n da = J F^-T N dA, dA / 4 = 0.0625
    # lCPDI: the deformed face: corners F (corner - pos) + pos, weighted normal n da
    o = _oracle_with_F(3, pos, Fm, "lCPDI")
    c, wn = o.surface_info(0, face, 11)
    u = _oracle_with_F(3, pos, np.eye(3), "lCPDI").surface_info(0, face, 11)[0]
    assert np.allclose(c, pos + (u - pos) @ Fm.T, rtol=0, atol=1e-14)
    assert np.allclose(wn, nda, rtol=0, atol=1e-14)
    _, wx = o.surface_info(0, face, 1)
    assert wx[0] == pytest.approx(np.linalg.norm(nda), rel=1e-13) and wx[1] == wx[2] == 0.
    # uGIMP: the undeformed corners and axis normal, the size of the deformed face
    o = _oracle_with_F(3, pos, Fm, "uGIMP")
    c, wn = o.surface_info(0, face, 11)
    assert np.array_equal(c, u)
    assert np.allclose(wn, np.linalg.norm(nda) * N, rtol=0, atol=1e-14)


@pytest.mark.parametrize("defo", ["stretch", "shear", "rotation"])
@pytest.mark.parametrize("face", [1, 2, 3, 4])
def test_deformed_faces_2d(face, defo):
    Fm = {"stretch": np.diag([1.3, 0.7, 1.0]),
          "shear": np.array([[1.1, 0.25, 0.], [-0.1, 0.9, 0.], [0., 0., 1.]]),
          "rotation": _rot3(2, 0.6)}[defo]
    F2 = Fm[:2, :2]
    pos = np.array([5.5, 5.5, 0.])
    N = {1: (0., -1.), 2: (1., 0.), 3: (0., 1.), 4: (-1., 0.)}[face]
    T = {1: (1., 0.), 2: (0., 1.), 3: (-1., 0.), 4: (0., -1.)}[face]       # edge direction c1 -> c2
    # deformed edge of half length |F T| * 0.25: normal n = its clockwise perpendicular
    t = F2 @ np.array(T)
    half = 0.25 * np.linalg.norm(t)
    n = np.array([t[1], -t[0]]) / np.linalg.norm(t)
    thick = 1.7
    o = _oracle_with_F(2, pos, Fm, "lCPDI")
    c, wn = o.surface_info(0, face, 11, thick)
    assert c.shape == (2, 3)
    assert np.allclose(wn[:2], half * thick * n, rtol=0, atol=1e-14) and wn[2] == 0.
    _, wt = o.surface_info(0, face, 12, thick)
    assert np.allclose(wt[:2], half * thick * t / np.linalg.norm(t), rtol=0, atol=1e-14)
    # uGIMP: axis normal and tangent of the undeformed face, deformed length
    o = _oracle_with_F(2, pos, Fm, "uGIMP")
    c, wn = o.surface_info(0, face, 11, thick)
    assert np.allclose(wn[:2], half * thick * np.array(N), rtol=0, atol=1e-14)
    _, wt = o.surface_info(0, face, 12, thick)
    assert np.allclose(wt[:2], half * thick * np.array(T), rtol=0, atol=1e-14)
    assert np.allclose(np.abs(c[:, :2] - pos[:2]), 0.25, rtol=0, atol=1e-15)
