"""Pins the oracle's constitutive updates (and the setup's material constants) against closed
forms: Hooke's law in specific units, first-order stress rotation, J2 radial-return consistency,
and neo-Hookean hydrostatic and simple-shear states."""
import numpy as np
import pytest

import nairnmpm as nm
import oracle as orc

F = nm.F


def material(analysis, mtype, **props):
    prob = nm.Problem.create(analysis)
    dim3 = analysis == nm.THREED
    prob.set_grid([0, 4, 0, 4, 0, 4 if dim3 else 0], [1, 1, 1 if dim3 else -1])
    mid = prob.add_material(mtype, "m", **props)
    prob.add_point(mid, [1.5, 1.5, 1.5 if dim3 else 0.])
    _, mats, _, _, _ = prob.describe()
    return mats[0]


def state0(mtype):
    s = np.zeros(nm.NUM_FIELDS)
    if mtype == nm.MAT_NEOHOOKEAN:
        s[F["alt_xx"]] = s[F["alt_yy"]] = s[F["alt_zz"]] = 1.
        s[F["hist0"]] = s[F["hist1"]] = 1.
    return s


def lame(E, nu):
    G = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return lam, G


def test_isotropic_constants_and_wave_speed():
    E, nu, rho = 70000., 0.33, 2.7
    m = material(nm.THREED, nm.MAT_ISOTROPIC, E=E, nu=nu, rho=rho)
    lam, G = lame(E * 1e6, nu)
    r = rho * 1e-3
    np.testing.assert_allclose(m.C[0][0], (lam + 2 * G) / r, rtol=1e-13)
    np.testing.assert_allclose(m.C[0][1], lam / r, rtol=1e-13)
    np.testing.assert_allclose(m.C[5][5], G / r, rtol=1e-13)
    np.testing.assert_allclose(m.waveSpeed, np.sqrt((lam + 2 * G) / r), rtol=1e-13)
    m2 = material(nm.PLANE_STRAIN, nm.MAT_ISOTROPIC, E=E, nu=nu, rho=rho)
    np.testing.assert_allclose(m2.C[1][1], (lam + 2 * G) / r, rtol=1e-13)
    np.testing.assert_allclose(m2.C[4][1], lam / r, rtol=1e-13)


@pytest.mark.parametrize("analysis", [nm.THREED, nm.PLANE_STRAIN])
def test_isotropic_uniaxial_strain_is_hooke(analysis):
    E, nu, rho = 1000., 0.3, 1.0
    m = material(analysis, nm.MAT_ISOTROPIC, E=E, nu=nu, rho=rho)
    lam, G = lame(E * 1e6, nu)
    r = rho * 1e-3
    eps = 1e-5
    du = np.zeros((3, 3))
    du[0, 0] = eps
    s = orc.constitutive(analysis, m, du, state0(nm.MAT_ISOTROPIC))
    np.testing.assert_allclose(s[F["sp_xx"]], (lam + 2 * G) * eps / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["sp_yy"]], lam * eps / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["sp_zz"]], lam * eps / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["ep_xx"]], eps, rtol=1e-10)   # F = I + du for the first increment
    # engineering shear: sigma_xy = G gamma
    du = np.zeros((3, 3))
    du[0, 1] = du[1, 0] = 0.5e-5
    s = orc.constitutive(analysis, m, du, state0(nm.MAT_ISOTROPIC))
    np.testing.assert_allclose(s[F["sp_xy"]], G * 1e-5 / r, rtol=1e-12)


@pytest.mark.parametrize("analysis", [nm.THREED, nm.PLANE_STRAIN])
def test_spin_rotates_stress_to_first_order(analysis):
    """A pure spin W dt = theta*(e_y e_x - e_x e_y) rotates the stress: sigma' = R sigma R^T to
    first order in theta (hypoelastic update, MaterialBaseMPM.cpp:1012-1047). A sign error in
    the rotation terms would leave an O(theta) mismatch."""
    m = material(analysis, nm.MAT_ISOTROPIC, E=1000., nu=0.3, rho=1.0)
    s0 = state0(nm.MAT_ISOTROPIC)
    sig = np.array([[3., 1.5, 0.7], [1.5, -2., 0.4], [0.7, 0.4, 1.]]) * 1e9
    if analysis == nm.PLANE_STRAIN:
        sig[0, 2] = sig[2, 0] = sig[1, 2] = sig[2, 1] = 0.
    for name, (i, j) in (("sp_xx", (0, 0)), ("sp_yy", (1, 1)), ("sp_zz", (2, 2)), ("sp_xy", (0, 1)),
                         ("sp_xz", (0, 2)), ("sp_yz", (1, 2))):
        s0[F[name]] = sig[i, j]
    for theta in (1e-4, 1e-6):
        du = np.zeros((3, 3))
        du[1, 0], du[0, 1] = theta, -theta           # counterclockwise spin about z
        s = orc.constitutive(analysis, m, du, s0)
        c, sn = np.cos(theta), np.sin(theta)
        R = np.array([[c, -sn, 0.], [sn, c, 0.], [0., 0., 1.]])
        rot = R @ sig @ R.T
        got = np.array([[s[F["sp_xx"]], s[F["sp_xy"]], s[F["sp_xz"]]],
                        [s[F["sp_xy"]], s[F["sp_yy"]], s[F["sp_yz"]]],
                        [s[F["sp_xz"]], s[F["sp_yz"]], s[F["sp_zz"]]]])
        err = np.abs(got - rot).max() / np.abs(sig).max()
        assert err < 10 * theta ** 2, (theta, err)


def j2(analysis=nm.THREED):
    return material(analysis, nm.MAT_ISOPLASTICITY, E=70000., nu=0.33, rho=2.7, **{"yield": 300., "Ep": 1000.})


@pytest.mark.parametrize("analysis", [nm.THREED, nm.PLANE_STRAIN])
def test_j2_elastic_below_yield(analysis):
    m = j2(analysis)
    E, nu, r = 70000e6, 0.33, 2.7e-3
    lam, G = lame(E, nu)
    K = lam + 2 * G / 3
    eps = 1e-4                                       # well below yield 300 MPa / E
    du = np.zeros((3, 3))
    du[0, 0] = eps
    s = orc.constitutive(analysis, m, du, state0(nm.MAT_ISOPLASTICITY))
    assert s[F["hist0"]] == 0.
    np.testing.assert_allclose(s[F["pressure"]], -K * eps / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["sp_xx"]], 2 * G * (eps - eps / 3) / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["sp_yy"]], 2 * G * (-eps / 3) / r, rtol=1e-12)


@pytest.mark.parametrize("analysis", [nm.THREED, nm.PLANE_STRAIN])
def test_j2_radial_return_consistency(analysis):
    """After a plastic increment the deviatoric stress lies on the hardened yield surface,
    ||s|| = sqrt(2/3)(sigma_y + Ep alpha)/rho, is parallel to the trial deviator, and the
    plastic strain increment is normal to the surface (its norm is sqrt(3/2) alpha)."""
    m = j2(analysis)
    E, nu, r = 70000e6, 0.33, 2.7e-3
    lam, G = lame(E, nu)
    sy, Ep = 300e6, 1000e6
    s = state0(nm.MAT_ISOPLASTICITY)
    for k in range(4):
        du = np.zeros((3, 3))
        du[0, 1] = du[1, 0] = 6e-3                  # shear beyond yield
        du[0, 0] = 2e-3
        s = orc.constitutive(analysis, m, du, s)
        alpha = s[F["hist0"]]
        assert alpha > 0
        dev = np.array([s[F["sp_xx"]], s[F["sp_yy"]], s[F["sp_zz"]]])
        shear = np.array([s[F["sp_xy"]]] + ([s[F["sp_xz"]], s[F["sp_yz"]]] if analysis == nm.THREED else []))
        smag = np.sqrt((dev ** 2).sum() + 2 * (shear ** 2).sum())
        np.testing.assert_allclose(smag, np.sqrt(2. / 3.) * (sy + Ep * alpha) / r, rtol=1e-10)
        assert abs(dev.sum()) < 1e-6 * smag          # deviatoric
        ep = np.array([s[F["alt_xx"]], s[F["alt_yy"]], s[F["alt_zz"]]])
        eps_sh = np.array([s[F["alt_xy"]]] + ([s[F["alt_xz"]], s[F["alt_yz"]]] if analysis == nm.THREED else [])) / 2.
        if k == 0:
            # first plastic step from a virgin state: ||dep|| = sqrt(3/2) dalpha
            np.testing.assert_allclose(np.sqrt((ep ** 2).sum() + 2 * (eps_sh ** 2).sum()),
                                       np.sqrt(1.5) * alpha, rtol=1e-10)


def test_j2_hydrostatic_never_yields():
    m = j2()
    s = state0(nm.MAT_ISOPLASTICITY)
    for _ in range(10):
        s = orc.constitutive(nm.THREED, m, -2e-3 * np.eye(3), s)
    assert s[F["hist0"]] == 0.
    assert max(abs(s[F[c]]) for c in ("sp_xx", "sp_yy", "sp_zz", "sp_xy")) < 1e-3 * abs(s[F["pressure"]])


def mooney(G1=8., G2=2., K=50., rho=1.1, analysis=nm.THREED, **kw):
    return material(analysis, nm.MAT_MOONEY, G1=G1, G2=G2, K=K, rho=rho, **kw)


def test_mooney_constants():
    m = mooney()
    np.testing.assert_allclose(m.G1sp, 8e6 / 1.1e-3, rtol=1e-14)
    np.testing.assert_allclose(m.G2sp, 2e6 / 1.1e-3, rtol=1e-14)
    np.testing.assert_allclose(m.Ksp, 50e6 / 1.1e-3, rtol=1e-14)
    np.testing.assert_allclose(m.waveSpeed, np.sqrt((50e6 + 4 * 10e6 / 3) / 1.1e-3), rtol=1e-14)
    # E and nu input: G1 + G2 = E / (2 (1 + nu))
    m2 = material(nm.THREED, nm.MAT_MOONEY, E=30., nu=0.4, G2=3., rho=1.)
    np.testing.assert_allclose(m2.G1sp + m2.G2sp, 30e6 / 2.8 / 1e-3, rtol=1e-13)


@pytest.mark.parametrize("analysis", [nm.THREED, nm.PLANE_STRAIN])
def test_mooney_simple_shear(analysis):
    """Mooney-Rivlin simple shear: sigma_xy = (G1 + G2) gamma at J = 1 (Kirchhoff/rho0), and the
    normal stress differences N1 = (G1 + G2) gamma^2, N2 = -G2 gamma^2 (Rivlin's universal results)."""
    G1, G2, r = 8e6, 2e6, 1.1e-3
    m = mooney(analysis=analysis)
    gam = 0.05
    du = np.zeros((3, 3))
    du[0, 1] = gam
    s = orc.constitutive(analysis, m, du, state0(nm.MAT_NEOHOOKEAN))
    np.testing.assert_allclose(s[F["hist0"]], 1.0, rtol=1e-15)
    np.testing.assert_allclose(s[F["sp_xy"]], (G1 + G2) * gam / r, rtol=1e-13)
    np.testing.assert_allclose(s[F["sp_xx"]] - s[F["sp_yy"]], (G1 + G2) * gam ** 2 / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["sp_yy"]] - s[F["sp_zz"]], -G2 * gam ** 2 / r, rtol=1e-12)
    assert abs(s[F["sp_xx"]] + s[F["sp_yy"]] + s[F["sp_zz"]]) < 1e-9 * abs(s[F["sp_xy"]])


@pytest.mark.parametrize("uj", [0, 1, 2])
def test_mooney_hydrostatic(uj):
    """Uniform stretch: deviatoric stress vanishes and P = -J U'(J) with the bulk modulus."""
    K, r = 50e6, 1.1e-3
    m = mooney(UJOption=uj)
    e = 0.02
    s = orc.constitutive(nm.THREED, m, e * np.eye(3), state0(nm.MAT_NEOHOOKEAN))
    J = (1 + e) ** 3
    dU = {0: 0.5 * K * (J - 1 / J), 1: K * (J - 1), 2: K * np.log(J) / J}[uj]
    np.testing.assert_allclose(s[F["pressure"]], -J * dU / r, rtol=1e-12)
    assert max(abs(s[F[c]]) for c in ("sp_xx", "sp_yy", "sp_zz", "sp_xy")) < 1e-9 * abs(s[F["pressure"]])


def test_mooney_with_g2_zero_is_neohookean_deviator():
    """G2 = 0 reduces the Mooney deviator to G1 dev(B) / J^(2/3) (neo-Hookean, Kirchhoff/rho0)."""
    m = mooney(G2=0.)
    rng = np.random.default_rng(3)
    du = 0.02 * rng.standard_normal((3, 3))
    s = orc.constitutive(nm.THREED, m, du, state0(nm.MAT_NEOHOOKEAN))
    Fm = np.eye(3) + du
    B = Fm @ Fm.T
    J = np.linalg.det(Fm)
    dev = (B - np.trace(B) / 3 * np.eye(3)) * 8e6 / 1.1e-3 / J ** (2. / 3.)
    got = np.array([[s[F["sp_xx"]], s[F["sp_xy"]], s[F["sp_xz"]]],
                    [s[F["sp_xy"]], s[F["sp_yy"]], s[F["sp_yz"]]],
                    [s[F["sp_xz"]], s[F["sp_yz"]], s[F["sp_zz"]]]])
    np.testing.assert_allclose(got, dev, rtol=1e-11, atol=1e-11 * np.abs(dev).max())


def nh():
    return material(nm.THREED, nm.MAT_NEOHOOKEAN, G=10., K=50., rho=1.1)


def test_neohookean_hydrostatic_pressure():
    m = nh()
    G, K, r = 10e6, 50e6, 1.1e-3
    lam = K - 2 * G / 3
    e = 0.01
    s = orc.constitutive(nm.THREED, m, e * np.eye(3), state0(nm.MAT_NEOHOOKEAN))
    J = (1 + e) ** 3
    b = (1 + e) ** 2
    np.testing.assert_allclose(s[F["hist0"]], J, rtol=1e-14)
    # Kirchhoff pressure / rho0 for U(J) = (lam/2)((J^2-1)/2 - ln J) plus the G(I1/3 - 1) term
    P = -(J * 0.5 * lam * (J - 1 / J) + G * (b - 1)) / r
    np.testing.assert_allclose(s[F["pressure"]], P, rtol=1e-12)
    assert max(abs(s[F[c]]) for c in ("sp_xx", "sp_yy", "sp_zz", "sp_xy")) < 1e-8 * abs(P)


def test_neohookean_simple_shear():
    m = nh()
    G, r = 10e6, 1.1e-3
    gam = 0.05
    du = np.zeros((3, 3))
    du[0, 1] = gam                                    # F = [[1, gam, 0], [0, 1, 0], [0, 0, 1]]
    s = orc.constitutive(nm.THREED, m, du, state0(nm.MAT_NEOHOOKEAN))
    B = np.array([[1 + gam ** 2, gam, 0.], [gam, 1., 0.], [0., 0., 1.]])
    dev = B - np.trace(B) / 3 * np.eye(3)
    np.testing.assert_allclose(s[F["sp_xy"]], G * gam / r, rtol=1e-13)
    np.testing.assert_allclose(s[F["sp_xx"]], G * dev[0, 0] / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["sp_zz"]], G * dev[2, 2] / r, rtol=1e-12)
    np.testing.assert_allclose(s[F["hist0"]], 1.0, rtol=1e-15)   # isochoric
