"""Regression contact normals in the oracle (normals methods 5 and 6, the reference's default;
CrackVelocityFieldMulti::LinearRegressionNormal and LogisticRegressionNormal,
Nodes/MachineLearning.cpp:58-633, over the particles listed at each contact node by
MaterialContactNode::GetMaterialContactData, MaterialContactNode.cpp:253-350), pinned by what
fixes them independently of the oracle's code: the list is the set of particles whose uGIMP support
holds the node; the linear normal is the ridge least-squares fit of the +-1 material labels
(numpy's solver); the logistic normal is a stationary point of the penalised least-squares misfit
of tanh(beta.x/2) (scipy's optimiser); the separation is the gap between the two bodies' particle
surfaces; the deformed radius is where the ray from the particle's center leaves its deformed
domain; contact with these normals conserves momentum."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc

F = nm.F


def _blocks(normals=5, vy=3000., gap=1., cells=32, jitter=0.):
    xml = configs.contact_blocks_2d(normals=normals, vy=vy, gap=gap, cells=cells)
    prob = nm.Problem.from_xml(xml)
    g, mats, run, n, _ = prob.describe()
    d, i = prob.particles()
    if jitter:
        rng = np.random.default_rng(7)
        for a in ("pos_x", "pos_y"):
            d[F[a]] += rng.uniform(-jitter, jitter, n)
    return g, mats, run, d, i


def _node(g, i, j, k=0):
    return (k * (g.ny + 1) + j) * (g.nx + 1) + i


def _listed(g, d, node_xyz, lp=0.5, dim=2):
    """particles whose uGIMP support holds the node: |x_p - x_n| < dx (1 + lp/2) on every axis"""
    inside = np.ones(d.shape[1], dtype=bool)
    for a, (c, h) in enumerate(zip(("pos_x", "pos_y", "pos_z")[:dim], (g.dx, g.dy, g.dz)[:dim])):
        inside &= np.abs(d[F[c]] - node_xyz[a]) < h * (1. + lp / 2.)
    return np.nonzero(inside)[0]


def _interface_node(g):
    """a node between the two blocks (x = 16 mm, mid height)"""
    i = int(round((16. - g.xmin) / g.dx))
    j = int(round((16. - g.ymin) / g.dy))
    return i, j, (g.xmin + i * g.dx, g.ymin + j * g.dy)


@pytest.mark.parametrize("jitter", [0., 0.2])
def test_linear_normal_is_the_ridge_fit_of_the_labels(jitter):
    g, mats, run, d, iv = _blocks(normals=5, jitter=jitter, gap=0.)
    i, j, nd = _interface_node(g)
    ora = orc.Oracle(g, mats, run, d, iv)
    for field in (0, 1):
        norm, deln, count = ora.regression_normal(_node(g, i, j), field, logistic=False)
        listed = _listed(g, d, nd)
        assert count == len(listed) and count > 8
        x = d[F["pos_x"]][listed] - nd[0]
        y = d[F["pos_y"]][listed] - nd[1]
        fld = np.array([mats[m].field for m in iv[1][listed]])
        lab = np.where(fld == field, -1., 1.)
        X = np.stack([x, y, np.ones_like(x)], axis=1)
        beta = np.linalg.solve(X.T @ X + np.diag([0., 0., 0.1]), X.T @ lab)
        want = beta[:2] / np.linalg.norm(beta[:2])
        assert np.allclose(norm[:2], want, rtol=0, atol=1e-12), (norm, want)
        assert norm[2] == 0.
        # the normal points from the field's material to the others
        side = np.sign(want[0])
        assert side == (1. if field == 0 else -1.)


def test_linear_normal_3d_is_the_ridge_fit_of_the_labels():
    xml = configs.contact_cylinders_3d(cells=24, radius=5., length=8., gap=0.5, shape="uGIMP", normals=5)
    prob = nm.Problem.from_xml(xml)
    g, mats, run, n, _ = prob.describe()
    d, iv = prob.particles()
    rng = np.random.default_rng(11)
    for a in ("pos_x", "pos_y", "pos_z"):
        d[F[a]] += rng.uniform(-0.15, 0.15, n)
    i, j, k = [int(round((12. - lo) / h)) for lo, h in ((g.xmin, g.dx), (g.ymin, g.dy), (g.zmin, g.dz))]
    nd = (g.xmin + i * g.dx, g.ymin + (j + 1) * g.dy, g.zmin + k * g.dz)
    ora = orc.Oracle(g, mats, run, d, iv)
    norm, deln, count = ora.regression_normal(_node(g, i, j + 1, k), 0, logistic=False)
    listed = _listed(g, d, nd, dim=3)
    assert count == len(listed) and count > 20
    X = np.stack([d[F["pos_x"]][listed] - nd[0], d[F["pos_y"]][listed] - nd[1], d[F["pos_z"]][listed] - nd[2],
                  np.ones(len(listed))], axis=1)
    fld = np.array([mats[m].field for m in iv[1][listed]])
    lab = np.where(fld == 0, -1., 1.)
    beta = np.linalg.solve(X.T @ X + np.diag([0., 0., 0., 0.1]), X.T @ lab)
    want = beta[:3] / np.linalg.norm(beta[:3])
    assert np.allclose(norm, want, rtol=0, atol=1e-12), (norm, want)
    assert want[0] > 0.9                # the cylinders meet across x


def test_logistic_normal_is_a_stationary_point_of_the_penalised_misfit():
    from scipy.optimize import least_squares
    g, mats, run, d, iv = _blocks(normals=6, jitter=0.2, gap=0.)
    i, j, nd = _interface_node(g)
    ora = orc.Oracle(g, mats, run, d, iv)
    norm, deln, count = ora.regression_normal(_node(g, i, j), 0, logistic=True)
    listed = _listed(g, d, nd)
    x = d[F["pos_x"]][listed] - nd[0]
    y = d[F["pos_y"]][listed] - nd[1]
    fld = np.array([mats[m].field for m in iv[1][listed]])
    lab = np.where(fld == 0, -1., 1.)
    X = np.stack([x, y, np.ones_like(x)], axis=1)
    lam = np.array([1e-7 * g.dx ** 2, 1e-7 * g.dx ** 2, 0.])
    b0 = np.linalg.solve(X.T @ X + np.diag(lam), X.T @ lab) * 2.
    res = least_squares(lambda b: np.concatenate([lab - np.tanh(X @ b / 2.), np.sqrt(lam) * b]), b0,
                        xtol=1e-15, ftol=1e-15, gtol=1e-15)
    want = res.x[:2] / np.linalg.norm(res.x[:2])

    def angle(u):
        return np.degrees(np.arccos(np.clip(u[:2] @ want, -1., 1.)))
    # the reference stops once the normal turns by less than 1 - cos = 1e-5 (0.26 degrees) in one
    # Gauss-Newton iteration, so it lands near the optimum, much nearer than the linear fit it starts from
    nl, _, _ = ora.regression_normal(_node(g, i, j), 0, logistic=False)
    assert angle(norm) < 0.5 and angle(norm) < angle(nl) / 3., (angle(norm), angle(nl))


def test_logistic_normal_of_a_symmetric_cloud_is_the_symmetry_normal():
    g, mats, run, d, iv = _blocks(normals=6, gap=0.)
    i, j, nd = _interface_node(g)
    ora = orc.Oracle(g, mats, run, d, iv)
    for field, sign in ((0, 1.), (1, -1.)):
        norm, _, _ = ora.regression_normal(_node(g, i, j), field, logistic=True)
        assert np.allclose(norm, [sign, 0., 0.], atol=1e-12)


@pytest.mark.parametrize("gap", [0.5, 1.0])
def test_separation_is_the_gap_between_particle_surfaces(gap):
    g, mats, run, d, iv = _blocks(normals=5, gap=gap)
    i, j, nd = _interface_node(g)
    ora = orc.Oracle(g, mats, run, d, iv)
    norm, deln, count = ora.regression_normal(_node(g, i, j), 0, logistic=False)
    assert np.allclose(norm[:2], [1., 0.], atol=1e-12)
    listed = _listed(g, d, nd)
    fld = np.array([mats[m].field for m in iv[1][listed]])
    xs = d[F["pos_x"]][listed]
    r = 0.5 * 0.5 * g.dx                 # undeformed half size along x (lp = 0.5)
    want = (xs[fld == 1] - r).min() - (xs[fld == 0] + r).max()
    assert abs(deln - want) < 1e-12
    if gap == int(gap):
        assert abs(deln - gap) < 1e-12     # blocks filling whole cells: particle surfaces at their edges


def _set_F(d, p, Fm):
    d[F["ep_xx"]][p] = Fm[0][0] - 1.
    d[F["ep_yy"]][p] = Fm[1][1] - 1.
    d[F["ep_zz"]][p] = Fm[2][2] - 1.
    d[F["ep_xy"]][p] = Fm[1][0] + Fm[0][1]
    d[F["wr_xy"]][p] = Fm[1][0] - Fm[0][1]
    d[F["ep_xz"]][p] = Fm[2][0] + Fm[0][2]
    d[F["wr_xz"]][p] = Fm[2][0] - Fm[0][2]
    d[F["ep_yz"]][p] = Fm[2][1] + Fm[1][2]
    d[F["wr_yz"]][p] = Fm[2][1] - Fm[1][2]


def test_deformed_radius_is_where_the_ray_leaves_the_domain_3d():
    prob = nm.Problem.from_xml(configs.neohookean_block_3d(cells=8, lo=2, hi=6))
    g, mats, run, n, _ = prob.describe()
    d, iv = prob.particles()
    rng = np.random.default_rng(3)
    Fs = [np.diag([2., 1., 1.]), np.eye(3) + 0.3 * rng.standard_normal((3, 3))]
    for p, Fm in enumerate(Fs):
        _set_F(d, p, Fm)
    ora = orc.Oracle(g, mats, run, d, iv)
    psz = 0.5 * 0.5 * np.array([g.dx, g.dy, g.dz])
    assert abs(ora.deformed_radius(0, [1., 0., 0.]) - 2. * psz[0]) < 1e-15
    assert abs(ora.deformed_radius(0, [0., 1., 0.]) - psz[1]) < 1e-15
    for _ in range(5):
        nv = rng.standard_normal(3)
        nv /= np.linalg.norm(nv)
        R = ora.deformed_radius(1, nv)
        # the point R n of the deformed domain F(box) maps back onto the undeformed ellipsoid's surface
        back = np.linalg.solve(Fs[1], R * nv) / psz
        assert abs(np.linalg.norm(back) - 1.) < 1e-12


def test_deformed_radius_2d_symmetric_stretch():
    g, mats, run, d, iv = _blocks()
    Fm = np.array([[1.3, 0.2, 0.], [0.2, 0.9, 0.], [0., 0., 1.]])
    _set_F(d, 0, Fm)
    ora = orc.Oracle(g, mats, run, d, iv)
    psz = 0.5 * 0.5 * np.array([g.dx, g.dy])
    for ang in (0., 0.7, 2.1):
        nv = np.array([np.cos(ang), np.sin(ang), 0.])
        R = ora.deformed_radius(0, nv)
        back = np.linalg.solve(Fm[:2, :2], R * nv[:2]) / psz
        assert abs(np.linalg.norm(back) - 1.) < 1e-12


@pytest.mark.parametrize("normals", [5, 6])
def test_regression_contact_conserves_momentum_and_transfers_friction(normals):
    g, mats, run, d, iv = _blocks(normals=normals)
    ora = orc.Oracle(g, mats, run, d, iv)
    p0 = np.array([(d[F["mass"]] * d[F[c]]).sum() for c in ("vel_x", "vel_y")])
    ora.step(300)
    r, ri = ora.particles()
    p1 = np.array([(r[F["mass"]] * r[F[c]]).sum() for c in ("vel_x", "vel_y")])
    scale = (d[F["mass"]] * np.abs(d[F["vel_x"]])).sum()
    assert np.abs(p1 - p0).max() < 1e-12 * scale
    a = ri[1] == 0
    assert r[F["vel_x"]][a].mean() < 0.           # the left block bounced back
    assert r[F["vel_y"]][a].mean() > 100.         # friction dragged it along with the sliding block
    # the particle surfaces never interpenetrate by more than a small fraction of a cell
    assert (r[F["pos_x"]][~a].min() - 0.25) - (r[F["pos_x"]][a].max() + 0.25) > -0.1


def test_omitted_normals_default_to_logistic_regression():
    xml = configs.contact_blocks_2d(normals=2).replace(' Normals="2"', '')
    g, mats, run, n, _ = nm.Problem.from_xml(xml).describe()
    assert run.normalsMethod == 6


def test_specified_normal_from_polar_and_azimuth():
    xml = configs.contact_blocks_2d(normals=4).replace('Normals="4"', 'Normals="4" Azimuth="30"')
    _, _, run, _, _ = nm.Problem.from_xml(xml).describe()
    assert np.allclose(list(run.contactNormal), [np.cos(np.pi / 6), np.sin(np.pi / 6), 0.], atol=1e-15)
    xml3 = configs.contact_cylinders_3d(cells=24, radius=5., length=8., gap=1., shape="uGIMP", normals=4)
    xml3 = xml3.replace('Normals="4"', 'Normals="4" Polar="90" Azimuth="0"')
    _, _, run3, _, _ = nm.Problem.from_xml(xml3).describe()
    assert np.allclose(list(run3.contactNormal), [1., 0., 0.], atol=1e-15)
    with pytest.raises(nm.MPMError, match="0 to 6"):
        nm.Problem.from_xml(configs.contact_blocks_2d(normals=7))
