"""Pins the oracle's shape functions without the reference: uGIMP is the convolution of the grid
hat function with the particle's characteristic function (Bardenhagen & Kober 2004), computed
here from its antiderivative, independently of the oracle's piecewise closed form."""
import numpy as np
import pytest

import nairnmpm as nm
import oracle as orc


def hat_integral(u):
    """antiderivative of the hat max(0, 1-|u|/2) (natural coordinates, nodes 2 apart)"""
    u = np.asarray(u, dtype=np.float64)
    return np.where(u <= -2., 0., np.where(u <= 0., (u + 2.) ** 2 / 4.,
                                           np.where(u <= 2., 2. - (2. - u) ** 2 / 4., 2.)))


def hat(u):
    return np.maximum(0., 1. - np.abs(u) / 2.)


def gimp_1d(d, lp):
    return (hat_integral(d + lp) - hat_integral(d - lp)) / (2. * lp)


def dgimp_1d(d, lp):
    return (hat(d + lp) - hat(d - lp)) / (2. * lp)


def grid(analysis=nm.THREED, n=8, dx=1.0):
    g = nm.GridDesc()
    g.analysis = analysis
    g.nx = g.ny = n
    g.nz = n if analysis == nm.THREED else 0
    g.xmin = g.ymin = -1.0
    g.zmin = -1.0 if analysis == nm.THREED else 0.0
    g.dx = g.dy = g.dz = dx
    g.thickness = 1.0
    return g


def node_xyz(g, node):
    nnx, nny = g.nx + 1, g.ny + 1
    i = node % nnx
    j = (node // nnx) % nny
    k = node // (nnx * nny)
    return np.array([g.xmin + i * g.dx, g.ymin + j * g.dy, g.zmin + k * g.dz])


@pytest.mark.parametrize("lp", [0.5, 0.25, 1.0 / 3.0])
@pytest.mark.parametrize("dim", [2, 3])
def test_ugimp_is_hat_convolution(dim, lp):
    analysis = nm.THREED if dim == 3 else nm.PLANE_STRAIN
    g = grid(analysis)
    rng = np.random.default_rng(11)
    e_ijk = (3, 4, 3)
    elem = e_ijk[0] + g.nx * (e_ijk[1] + (g.ny * e_ijk[2] if dim == 3 else 0))
    for _ in range(40):
        pos = np.array([g.xmin + (e_ijk[0] + rng.uniform()) * g.dx, g.ymin + (e_ijk[1] + rng.uniform()) * g.dy,
                        g.zmin + (e_ijk[2] + rng.uniform()) * g.dz if dim == 3 else 0.0])
        lps = np.array([lp, lp, lp if dim == 3 else 1.0])
        nodes, fn, dx, dy, dz = orc.shape(g, nm.SHAPE_UGIMP, elem, pos, lps)
        found = {}
        for k in range(len(nodes)):
            found[int(nodes[k])] = (fn[k], dx[k], dy[k], dz[k])
        # every node within the support on all axes, and only those, with the convolution value
        cell_lo = np.array([g.xmin + e_ijk[0] * g.dx, g.ymin + e_ijk[1] * g.dy, g.zmin + e_ijk[2] * g.dz])
        xi = 2. * (pos - cell_lo) / g.dx - 1.
        expect = {}
        for oz in (range(-1, 3) if dim == 3 else [0]):
            for oy in range(-1, 3):
                for ox in range(-1, 3):
                    off = np.array([ox, oy, oz])
                    gnat = 2. * off - 1.
                    d = xi - gnat
                    axes = range(dim)
                    if any(abs(d[a]) >= 2. + lp for a in axes):
                        continue
                    w = np.prod([gimp_1d(d[a], lp) for a in axes])
                    grads = []
                    for a in range(3):
                        if a >= dim:
                            grads.append(0.)
                            continue
                        gr = dgimp_1d(d[a], lp) * 2. / g.dx
                        for b in axes:
                            if b != a:
                                gr *= gimp_1d(d[b], lp)
                        grads.append(gr)
                    ijk = np.array(e_ijk) + off
                    node = ijk[0] + (g.nx + 1) * (ijk[1] + ((g.ny + 1) * ijk[2] if dim == 3 else 0))
                    expect[int(node)] = (w, *grads)
        assert set(found) == set(expect)
        for node, vals in expect.items():
            np.testing.assert_allclose(found[node], vals, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_partition_of_unity_and_linear_reproduction(dim):
    analysis = nm.THREED if dim == 3 else nm.PLANE_STRAIN
    g = grid(analysis)
    rng = np.random.default_rng(5)
    e_ijk = (4, 3, 4)
    elem = e_ijk[0] + g.nx * (e_ijk[1] + (g.ny * e_ijk[2] if dim == 3 else 0))
    for shape, lp in ((nm.SHAPE_UGIMP, 0.5), (nm.SHAPE_LINEAR, 0.5)):
        for _ in range(30):
            pos = np.array([g.xmin + (e_ijk[0] + rng.uniform()) * g.dx, g.ymin + (e_ijk[1] + rng.uniform()) * g.dy,
                            g.zmin + (e_ijk[2] + rng.uniform()) * g.dz if dim == 3 else 0.0])
            nodes, fn, dx, dy, dz = orc.shape(g, shape, elem, pos, np.array([lp, lp, lp if dim == 3 else 1.]))
            assert abs(fn.sum() - 1.) < 1e-14
            assert abs(dx.sum()) < 1e-13 and abs(dy.sum()) < 1e-13 and abs(dz.sum()) < 1e-13
            xyz = np.array([node_xyz(g, n) for n in nodes])
            np.testing.assert_allclose(fn @ xyz[:, :dim], pos[:dim], rtol=0, atol=1e-13)
            # gradient of the linear field x reproduces the identity
            np.testing.assert_allclose(dx @ xyz[:, 0], 1., atol=1e-13)
            np.testing.assert_allclose(dy @ xyz[:, 1], 1., atol=1e-13)


def test_regular_site_node_counts():
    """lp = 1/2 particle at xi = +-1/2 sits on the support edge of the third node on every axis,
    and the strict test drops it: 8 nodes in 3D, 4 in 2D; off the site, up to 27 (9)"""
    for dim, full in ((3, 27), (2, 9)):
        analysis = nm.THREED if dim == 3 else nm.PLANE_STRAIN
        g = grid(analysis)
        elem = 3 + g.nx * (3 + (g.ny * 3 if dim == 3 else 0))
        lo = np.array([g.xmin + 3 * g.dx, g.ymin + 3 * g.dy, g.zmin + 3 * g.dz if dim == 3 else 0.])
        pos = lo + np.array([0.25, 0.75, 0.25 if dim == 3 else 0.])
        n, *_ = orc.shape(g, nm.SHAPE_UGIMP, elem, pos, np.array([0.5, 0.5, 0.5 if dim == 3 else 1.]))
        assert len(n) == 2 ** dim
        pos2 = lo + np.array([0.2, 0.8, 0.1 if dim == 3 else 0.])
        n2, *_ = orc.shape(g, nm.SHAPE_UGIMP, elem, pos2, np.array([0.5, 0.5, 0.5 if dim == 3 else 1.]))
        assert len(n2) == full


def test_3d_z_factor_uses_lp_y():
    """The reference's 3D uGIMP scales the z factor by 1/(4 lp.y) (EightNodeIsoparamBrick.cpp:303).
    With lp.y == lp.z the weights are the convolution and sum to one; with lp.y != lp.z the sum
    departs from one, so the quirk is present in the restatement."""
    g = grid(nm.THREED)
    elem = 3 + g.nx * (3 + g.ny * 3)
    pos = np.array([g.xmin + 3.3, g.ymin + 3.6, g.zmin + 3.05])   # z near a node: quadratic branch
    n, fn, *_ = orc.shape(g, nm.SHAPE_UGIMP, elem, pos, np.array([0.5, 0.4, 0.4]))
    assert abs(fn.sum() - 1.) < 1e-14
    n, fn, *_ = orc.shape(g, nm.SHAPE_UGIMP, elem, pos, np.array([0.5, 0.3, 0.4]))
    assert abs(fn.sum() - 1.) > 1e-4


def _cpdi_state(pos, lp, Fm):
    s = np.zeros(nm.NUM_FIELDS)
    F = nm.F
    s[F["pos_x"]], s[F["pos_y"]], s[F["pos_z"]] = pos
    s[F["lp_x"]], s[F["lp_y"]], s[F["lp_z"]] = lp
    s[F["ep_xx"]], s[F["ep_yy"]], s[F["ep_zz"]] = Fm[0, 0] - 1., Fm[1, 1] - 1., Fm[2, 2] - 1.
    s[F["ep_xy"]], s[F["wr_xy"]] = Fm[1, 0] + Fm[0, 1], Fm[1, 0] - Fm[0, 1]
    s[F["ep_xz"]], s[F["wr_xz"]] = Fm[2, 0] + Fm[0, 2], Fm[2, 0] - Fm[0, 2]
    s[F["ep_yz"]], s[F["wr_yz"]] = Fm[2, 1] + Fm[1, 2], Fm[2, 1] - Fm[1, 2]
    return s


@pytest.mark.parametrize("dim", [2, 3])
def test_lcpdi_reproduces_linear_fields(dim):
    """lCPDI (Sadeghirad et al. 2011) for an arbitrary deformed parallelepiped: sum N = 1,
    sum grad N = 0, sum N x_i = x_p and sum grad N_i (x) x_i = I. The last identity fixes every
    gradient weight wg of MatPoint3D.cpp:571-594 (a wrong sign in one term breaks it)."""
    analysis = nm.THREED if dim == 3 else nm.PLANE_STRAIN
    g = grid(analysis)
    rng = np.random.default_rng(21)
    for trial in range(25):
        e_ijk = (3, 4, 3)
        elem = e_ijk[0] + g.nx * (e_ijk[1] + (g.ny * e_ijk[2] if dim == 3 else 0))
        pos = np.array([g.xmin + (e_ijk[0] + rng.uniform(0.2, 0.8)) * g.dx,
                        g.ymin + (e_ijk[1] + rng.uniform(0.2, 0.8)) * g.dy,
                        g.zmin + (e_ijk[2] + rng.uniform(0.2, 0.8)) * g.dz if dim == 3 else 0.])
        Fm = np.eye(3) + 0.3 * rng.uniform(-1, 1, (3, 3))
        if dim == 2:
            Fm[2, :2] = Fm[:2, 2] = 0.
        lp = np.array([0.5, 0.5, 0.5 if dim == 3 else 1.])
        nodes, fn, dx, dy, dz = orc.cpdi_shape(g, elem, _cpdi_state(pos, lp, Fm))
        assert len(set(nodes.tolist())) == len(nodes)                  # merged, no duplicates
        xyz = np.array([node_xyz(g, n) for n in nodes])
        np.testing.assert_allclose(fn.sum(), 1., atol=1e-14)
        np.testing.assert_allclose([dx.sum(), dy.sum(), dz.sum()], 0., atol=1e-13)
        np.testing.assert_allclose(fn @ xyz[:, :dim], pos[:dim], atol=1e-13)
        grads = np.stack([dx, dy, dz])[:dim]
        np.testing.assert_allclose(grads @ xyz[:, :dim], np.eye(dim), atol=1e-12)


def test_lcpdi_undeformed_is_trilinear_average():
    """With F = I the particle domain is the box pos +- lp/2 cells and lCPDI's function of a node
    is the average of the node's trilinear function over the 8 corners (ws = 1/8)."""
    g = grid(nm.THREED)
    elem = 3 + g.nx * (3 + g.ny * 3)
    pos = np.array([g.xmin + 3.6, g.ymin + 3.3, g.zmin + 3.45])
    nodes, fn, *_ = orc.cpdi_shape(g, elem, _cpdi_state(pos, [0.5, 0.5, 0.5], np.eye(3)))
    r = 0.25
    expect = {}
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                c = pos + r * np.array([sx, sy, sz])
                ijk = np.floor((c - [g.xmin, g.ymin, g.zmin]) / g.dx).astype(int)
                loc = (c - [g.xmin, g.ymin, g.zmin]) / g.dx - ijk
                for a in (0, 1):
                    for b in (0, 1):
                        for d in (0, 1):
                            wgt = (loc[0] if a else 1 - loc[0]) * (loc[1] if b else 1 - loc[1]) * (loc[2] if d else 1 - loc[2])
                            n = (ijk[0] + a) + (g.nx + 1) * ((ijk[1] + b) + (g.ny + 1) * (ijk[2] + d))
                            expect[n] = expect.get(n, 0.) + wgt / 8.
    got = dict(zip(nodes.tolist(), fn))
    for n, w in expect.items():
        if w > 1e-14:
            np.testing.assert_allclose(got[n], w, atol=1e-15)
