"""Closed-form pins of the oracle where conservation checks say nothing: the J2 rotation terms
(3D, and the 2D plastic-strain quirk), the half velocity change in the FLIP position update,
grid against particle damping, the Coulomb friction magnitude and static branch, the normal
picked by each normals method, the displacement correction at the momentum update, and the
ReturnToElement bisection. Each test states the mechanics it checks; none restates the oracle's
own arithmetic."""
import numpy as np
import pytest

import nairnmpm as nm
from nairnmpm import configs
import oracle as orc
from test_oracle_materials import j2, lame

F = nm.F
COMPS = ("xx", "yy", "zz", "yz", "xz", "xy")


def _skew(w):
    return np.array([[0., -w[2], w[1]], [w[2], 0., -w[0]], [-w[1], w[0], 0.]])


def _rotation(axis, theta):
    """Rodrigues: rotation by theta about the unit vector axis."""
    k = _skew(axis)
    return np.eye(3) + np.sin(theta) * k + (1. - np.cos(theta)) * (k @ k)


def _get(s, prefix, engineering):
    """3x3 tensor of state rows prefix_xx ... prefix_xy (engineering shears halved)."""
    f = 0.5 if engineering else 1.
    xx, yy, zz, yz, xz, xy = (s[F[prefix + c]] for c in COMPS)
    return np.array([[xx, f * xy, f * xz], [f * xy, yy, f * yz], [f * xz, f * yz, zz]])


def _put(s, prefix, t, engineering):
    f = 2. if engineering else 1.
    s[F[prefix + "xx"]], s[F[prefix + "yy"]], s[F[prefix + "zz"]] = t[0, 0], t[1, 1], t[2, 2]
    s[F[prefix + "yz"]], s[F[prefix + "xz"]], s[F[prefix + "xy"]] = f * t[1, 2], f * t[0, 2], f * t[0, 1]


def _dev(a):
    return a - np.trace(a) / 3. * np.eye(3)


def _yielded_j2_state(rng, dim):
    """A J2 particle that yielded earlier and unloaded: alpha > 0, plastic strain deviatoric and
    non-zero, deviatoric stress inside the hardened surface, some pressure."""
    E, nu, r = 70000e6, 0.33, 2.7e-3
    sy = 300e6
    a = rng.normal(size=(3, 3))
    s = _dev(a + a.T)
    b = rng.normal(size=(3, 3))
    ep = _dev(b + b.T)
    if dim == 2:
        for t in (s, ep):
            t[0, 2] = t[2, 0] = t[1, 2] = t[2, 1] = 0.
    s *= 0.5 * np.sqrt(2. / 3.) * sy / r / np.linalg.norm(s)
    ep *= 2e-3 / np.linalg.norm(ep)
    st = np.zeros(nm.NUM_FIELDS)
    _put(st, "sp_", s, False)
    _put(st, "alt_", ep, True)
    st[F["hist0"]] = 3e-3
    st[F["pressure"]] = 1e8 / r
    return st, s, ep


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_j2_3d_spin_rotates_stress_and_plastic_strain(seed):
    """Under a pure spin (du = theta [w]x, no strain) a hypoelastic J2 particle inside its yield
    surface only rotates: s -> R s R^T, eplast -> R eplast R^T, pressure and alpha unchanged
    (IsoPlasticity.cpp:224-277). A sign error in any rotation term leaves an O(theta) error."""
    rng = np.random.default_rng(seed)
    m = j2(nm.THREED)
    st, s, ep = _yielded_j2_state(rng, 3)
    w = rng.normal(size=3)
    w /= np.linalg.norm(w)
    for theta in (1e-4, 1e-5):
        out = orc.constitutive(nm.THREED, m, theta * _skew(w), st)
        R = _rotation(w, theta)
        err_s = np.abs(_get(out, "sp_", False) - R @ s @ R.T).max() / np.abs(s).max()
        err_e = np.abs(_get(out, "alt_", True) - R @ ep @ R.T).max() / np.abs(ep).max()
        assert err_s < 10 * theta ** 2, (theta, err_s)
        assert err_e < 10 * theta ** 2, (theta, err_e)
        assert out[F["pressure"]] == st[F["pressure"]]
        assert out[F["hist0"]] == st[F["hist0"]]


def test_j2_2d_spin_rotates_shear_stress_with_plastic_strain():
    """In 2D the reference rotates the trial shear stress with the plastic strain difference,
    st0.xy += dwrot/2 (ep.xx - ep.yy), where the stress rotation would use sp.xx - sp.yy
    (IsoPlasticity.cpp:252, against the 3D branch :270-275). With dwrot = 2 theta: the normal
    stresses rotate physically, the shear stress moves by theta (ep.xx - ep.yy), and the plastic
    strain (engineering shear) rotates physically."""
    rng = np.random.default_rng(7)
    m = j2(nm.PLANE_STRAIN)
    st, s, ep = _yielded_j2_state(rng, 2)
    theta = 1e-5
    du = np.zeros((3, 3))
    du[1, 0], du[0, 1] = theta, -theta
    out = orc.constitutive(nm.PLANE_STRAIN, m, du, st)
    dsp = s[0, 0] - s[1, 1]
    dep = ep[0, 0] - ep[1, 1]
    scale = np.abs(s).max()
    assert abs(out[F["sp_xy"]] - (s[0, 1] + theta * dep)) < 1e-13 * scale
    # the physical rotation would have moved it by theta * (sp.xx - sp.yy): far from the result
    assert abs(out[F["sp_xy"]] - (s[0, 1] + theta * dsp)) > 1e-3 * theta * abs(dsp)
    assert abs(out[F["sp_xx"]] - (s[0, 0] - 2 * theta * s[0, 1])) < 1e-13 * scale
    assert abs(out[F["sp_yy"]] - (s[1, 1] + 2 * theta * s[0, 1])) < 1e-13 * scale
    # plastic strain: eps' = R eps R^T to first order (engineering shear gamma = 2 eps_xy)
    R = _rotation(np.array([0., 0., 1.]), theta)
    err = np.abs(_get(out, "alt_", True) - R @ ep @ R.T).max() / np.abs(ep).max()
    assert err < 10 * theta ** 2


def _gravity(xml, g):
    body = "  <Gravity>\n" + "".join(f"    <Body{a}Force>{v!r}</Body{a}Force>\n" for a, v in zip("XYZ", g))
    return xml.replace("</JANFEAInput>", body + "  </Gravity>\n</JANFEAInput>", 1)


def _damped(xml, ga, pa):
    return xml.replace("  </MPMHeader>", f"    <Damping>{ga!r}</Damping>\n    <PDamping>{pa!r}</PDamping>\n"
                       "  </MPMHeader>", 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_free_fall_positions_follow_half_g_t_squared(dim):
    """Free fall from rest: FLIP moves each particle by (v_start + dv/2) dt with dv = g dt, so
    after k steps the displacement is g (k dt)^2 / 2 exactly (the trapezoid rule is exact for
    constant acceleration). Moving by (v_start + dv) dt would add g k dt^2 / 2."""
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0., jitter=0.) if dim == 2 else \
        configs.j2_block_3d(cells=12, lo=3, hi=9, vx=0.)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    dt = run.timestep
    # accelerations that move the block about half a millimetre in 12 steps, so that the
    # displacements stand well above the rounding of the positions
    a0 = 1.0 / (144. * dt * dt)
    g = (0.3 * a0, -a0, 0.) if dim == 2 else (0., 0.2 * a0, -a0)
    prob = nm.Problem.from_xml(_gravity(base, g))
    o = orc.Oracle.from_problem(prob)
    d0, _ = o.particles()
    for k in range(1, 13):
        o.step(1)
        d, _ = o.particles()
        for a in range(dim):
            if g[a] == 0.:
                continue
            ax = "xyz"[a]
            disp = d[F["pos_" + ax]] - d0[F["pos_" + ax]]
            np.testing.assert_allclose(disp, 0.5 * g[a] * (k * dt) ** 2, rtol=1e-9)


@pytest.mark.parametrize("ga, pa", [(2.0e5, 0.), (0., 3.0e5), (1.0e5, 2.0e5)])
def test_damped_free_fall_positions(ga, pa):
    """Damped free fall of a rigid block (particle velocity = grid velocity = v): each step
    dv = (g - (alpha_p + alpha_g) v) dt and dx = (v + dv/2) dt (MatPoint3D.cpp:104-160), summed
    step by step."""
    base = configs.elastic_block_2d(cells=16, block=8, vmax=0., jitter=0.)
    _, _, run, _, _ = nm.Problem.from_xml(base).describe()
    dt = run.timestep
    g = (0., -1.0 / (144. * dt * dt), 0.)
    prob = nm.Problem.from_xml(_damped(_gravity(base, g), ga, pa))
    o = orc.Oracle.from_problem(prob)
    d0, _ = o.particles()
    v = x = 0.
    for k in range(1, 13):
        dv = (g[1] - (pa + ga) * v) * dt
        x += (v + 0.5 * dv) * dt
        v += dv
        o.step(1)
        d, _ = o.particles()
        np.testing.assert_allclose(d[F["vel_y"]], v, rtol=1e-10)
        np.testing.assert_allclose(d[F["pos_y"]] - d0[F["pos_y"]], x, rtol=1e-9)


def test_grid_and_particle_damping_act_on_different_velocities():
    """Grid damping scales the grid velocity at the particle, vm = S v - S a dt, and particle
    damping the particle's own velocity v (Adamp0 = alpha_p v + alpha_g vm,
    MatPoint3D.cpp:117-120). Damping does not change the forces of the step, so against the
    undamped step (v1, x1): v' = v1 - alpha dt u and x' = x1 - alpha dt^2 u / 2 with u = vm for
    grid damping and u = v for particle damping, where vm follows from the undamped step as
    (x1 - x0)/dt - (v1 - v0)/2. The test block has particle velocities that differ from the
    grid's (jittered positions and a sine profile), so swapping the two would fail."""
    base = configs.elastic_block_2d(cells=16, block=8)
    prob = nm.Problem.from_xml(base)
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    rng = np.random.default_rng(3)
    d[F["vel_x"]] += rng.uniform(-2000., 2000., d.shape[1])
    d[F["vel_y"]] += rng.uniform(-2000., 2000., d.shape[1])
    dt = run.timestep
    alpha = 0.02 / dt
    res = {}
    for key, (ga, pa) in {"none": (0., 0.), "grid": (alpha, 0.), "particle": (0., alpha)}.items():
        r = nm.RunParams.from_buffer_copy(run)
        r.gridDamping, r.particleDamping = ga, pa
        o = orc.Oracle(g, mats, r, d, i)
        o.step(1)
        res[key] = o.particles()[0]
    for ax in "xy":
        v0, x0 = d[F["vel_" + ax]], d[F["pos_" + ax]]
        v1, x1 = res["none"][F["vel_" + ax]], res["none"][F["pos_" + ax]]
        vm = (x1 - x0) / dt - 0.5 * (v1 - v0)
        assert np.abs(vm - v0).max() > 100.            # the grid velocity differs from the particle's
        for key, u in (("grid", vm), ("particle", v0)):
            vd, xd = res[key][F["vel_" + ax]], res[key][F["pos_" + ax]]
            np.testing.assert_allclose(vd, v1 - alpha * dt * u, rtol=0, atol=1e-9 * np.abs(v0).max())
            np.testing.assert_allclose(xd - x0, (x1 - x0) - 0.5 * alpha * dt * dt * u, rtol=0,
                                       atol=1e-9 * np.abs(x1 - x0).max())


# ---- contact at one node: two material velocity fields ----------------------------------

def _law(mu, mu_static=0., law_type=nm.MAT_COULOMB):
    m = nm.Material()
    m.type = law_type
    m.friction = mu
    m.frictionStatic = mu_static
    return m


def _pair(v1, v2, m1=2.0, m2=3.0, d1=(0., 0., 0.), d2=(-1e-3, 0., 0.)):
    """Node state of body 1 (left) and body 2 (right) meeting across the plane x = const: volume
    gradients of the extrapolation sum V grad N point from each body toward the other (+x for
    body 1, -x for body 2), so every normals method finds n = +x. Displacements d1, d2 give the
    separation (d2 - d1).n < 0 (the bodies overlap) by default."""
    mass = np.array([m1, m2])
    pk = np.array([m1 * np.asarray(v1, float), m2 * np.asarray(v2, float)])
    vol = np.array([1.0, 1.0])
    grad = np.array([[0.8, 0., 0.], [-0.8, 0., 0.]])
    disp = np.array([m1 * np.asarray(d1, float), m2 * np.asarray(d2, float)])
    return mass, vol, disp, grad, pk


def _contact(law, v1, v2, normals=2, call=0, dt=1e-6, ftot=None, **kw):
    mass, vol, disp, grad, pk = _pair(v1, v2, **kw)
    f = np.zeros((2, 3)) if ftot is None else np.asarray(ftot, float)
    p, fo, pc = orc.contact_node(law, normals, call, dt, mass, [4, 4], vol, disp, grad, pk, f, pk)
    return mass, pk, p, fo, pc


def test_frictionless_contact_stops_normal_approach_only():
    mass, pk0, pk, _, pkc = _contact(_law(0.), [3., 1., -0.5], [-2., -1., 0.25])
    v = pk / mass[:, None]
    assert abs(v[0, 0] - v[1, 0]) < 1e-15 * 3.                  # normal velocities equal
    np.testing.assert_allclose(v[:, 1:], (pk0 / mass[:, None])[:, 1:], rtol=1e-15)   # tangential kept
    np.testing.assert_allclose(pk.sum(0), pk0.sum(0), rtol=1e-15, atol=1e-15)
    np.testing.assert_array_equal(pkc, pk)                        # mass-momentum call keeps the copy


@pytest.mark.parametrize("mu", [0.1, 0.3])
def test_coulomb_sliding_tangential_impulse_is_mu_times_normal(mu):
    """Sliding contact (large tangential relative velocity): the normal approach stops, the
    tangential impulse on body 1 is mu times its normal impulse (Coulomb), opposes the relative
    sliding and does not reverse it (CoulombFriction.cpp:146-262)."""
    v1, v2 = np.array([3., 4., -2.]), np.array([-2., -1., 1.])
    mass, pk0, pk, _, _ = _contact(_law(mu), v1, v2)
    dp = pk[0] - pk0[0]
    n = np.array([1., 0., 0.])
    dpn = dp @ n
    dpt = dp - dpn * n
    assert dpn < 0.
    np.testing.assert_allclose(np.linalg.norm(dpt), mu * abs(dpn), rtol=1e-14)
    rel0 = (v1 - v2)[1:]
    v = pk / mass[:, None]
    rel1 = (v[0] - v[1])[1:]
    assert dpt[1:] @ rel0 < 0.                                     # opposes sliding
    np.testing.assert_allclose(rel1 / np.linalg.norm(rel1), rel0 / np.linalg.norm(rel0), rtol=1e-12)
    assert np.linalg.norm(rel1) < np.linalg.norm(rel0)
    assert abs(v[0, 0] - v[1, 0]) < 1e-14 * 3.
    np.testing.assert_allclose(pk.sum(0), pk0.sum(0), rtol=1e-14, atol=1e-14)


def test_coulomb_sticks_below_mu_n():
    """When the tangential impulse needed to stick is below mu N the fields stick: both leave
    with the center-of-mass velocity."""
    v1, v2 = np.array([3., 0.2, 0.]), np.array([-2., 0., 0.1])
    mass, pk0, pk, _, _ = _contact(_law(0.3), v1, v2)
    vc = pk0.sum(0) / mass.sum()
    np.testing.assert_allclose(pk / mass[:, None], np.array([vc, vc]), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("vt, sticks", [(2.0, True), (3.0, False)])
def test_static_friction_branch(vt, sticks):
    """With a static coefficient mu_s > mu the node sticks while the stick impulse is at most
    mu_s N and otherwise slides with the kinetic mu (CoulombFriction::GetSslideAcDt,
    CoulombFriction.cpp:266-282). The normal relative speed is 5 and the tangential one vt, so
    with the reduced mass the stick impulses are N = 5 mr and T = vt mr: vt = 2 has
    mu N < T <= mu_s N and sticks (it would slide without the static branch), vt = 3 has
    T > mu_s N and slides with mu = 0.3."""
    law = _law(0.3, 0.5)
    v1, v2 = np.array([3., vt, 0.]), np.array([-2., 0., 0.])
    mass, pk0, pk, _, _ = _contact(law, v1, v2)
    dp = pk[0] - pk0[0]
    if sticks:
        vc = pk0.sum(0) / mass.sum()
        np.testing.assert_allclose(pk / mass[:, None], np.array([vc, vc]), rtol=1e-14, atol=1e-15)
    else:
        np.testing.assert_allclose(np.linalg.norm(dp[1:]), 0.3 * abs(dp[0]), rtol=1e-14)


def test_stick_law_and_separating_node():
    # the stick contact law (type 60) always takes the center-of-mass velocity
    v1, v2 = np.array([3., 4., -2.]), np.array([-2., -1., 1.])
    mass, pk0, pk, _, _ = _contact(_law(0., law_type=nm.MAT_CONTACT), v1, v2)
    vc = pk0.sum(0) / mass.sum()
    np.testing.assert_allclose(pk / mass[:, None], np.array([vc, vc]), rtol=1e-14, atol=1e-15)
    # separating velocities, or a positive separation: no exchange under Coulomb
    for a, b, kw in (([-3., 1., 0.], [2., 0., 0.], {}), ([3., 1., 0.], [-2., 0., 0.], {"d2": (1e-3, 0., 0.)})):
        _, pk0, pk, _, _ = _contact(_law(0.3), a, b, **kw)
        np.testing.assert_array_equal(pk, pk0)


def test_momentum_update_contact_uses_mean_relative_velocity_over_the_step():
    """At the momentum-update call the separation is advanced over the step before the contact
    test: gap_end = gap + dt (u_start + u_end)/2, with u the relative normal velocity of body 2
    with respect to body 1 at the step's start (momenta minus f dt) and end (the trapezoid rule
    for a constant force, CoulombFriction.cpp:160-167). Gaps a little either side of zero
    decide contact; using u_start + u_end instead of their mean would flip one of them."""
    dt = 1e-3
    m1, m2 = 2.0, 3.0
    v1, v2 = np.array([3., 0., 0.]), np.array([-2., 0., 0.])
    f = np.array([[-4000., 0., 0.], [1500., 0., 0.]])           # decelerating approach
    u_end = (v2 - v1)[0]
    u_start = ((m2 * v2 - f[1] * dt) / m2 - (m1 * v1 - f[0] * dt) / m1)[0]
    shift = dt * 0.5 * (u_start + u_end)
    eps = 0.05 * abs(0.5 * dt * (u_end - u_start))
    for gap0, contact in ((-shift + eps, False), (-shift - eps, True)):
        _, pk0, pk, fo, _ = _contact(_law(0.), v1, v2, call=1, dt=dt, ftot=f, m1=m1, m2=m2,
                                     d1=(0., 0., 0.), d2=(gap0, 0., 0.))
        assert (not np.array_equal(pk, pk0)) == contact, (gap0, contact)
        if contact:
            # the momentum change is applied as a force too: f += dp/dt
            np.testing.assert_allclose(fo[0] - f[0], (pk[0] - pk0[0]) / dt, rtol=1e-12)


def _asym_pair():
    """Body 1's gradient along +x (magnitude 3), body 2's along -(cos 30, sin 30) (magnitude 1),
    volumes 1.0 and 2.0."""
    mass = np.array([2.0, 3.0])
    c, s = np.cos(np.pi / 6), np.sin(np.pi / 6)
    grad = np.array([[3., 0., 0.], [-c, -s, 0.]])
    vol = np.array([1.0, 2.0])
    v1, v2 = np.array([3., 2., 0.]), np.array([-2., -2., 0.])
    pk = np.array([mass[0] * v1, mass[1] * v2])
    disp = np.array([[0., 0., 0.], [-3e-3 * mass[1], -3e-3 * mass[1], 0.]])
    return mass, vol, disp, grad, pk


@pytest.mark.parametrize("normals, expect", [
    (0, "grad1"),       # MAXG: the larger volume gradient (body 1's, magnitude 3)
    (1, "grad2"),       # MAXV: the gradient of the larger volume (body 2's, turned toward body 2)
    (2, "avg"),         # AVGG: grad1 - grad2, normalised
])
def test_normals_method_picks(normals, expect):
    """Frictionless contact moves body 1 along the contact normal only, so the direction of its
    momentum change names the normal (CrackVelocityFieldMulti::GetNormalVector, :960-1069)."""
    mass, vol, disp, grad, pk = _asym_pair()
    p, _, _ = orc.contact_node(_law(0.), normals, 0, 1e-6, mass, [4, 4], vol, disp, grad, pk, np.zeros((2, 3)), pk)
    dp = p[0] - pk[0]
    want = {"grad1": grad[0], "grad2": -grad[1], "avg": grad[0] - grad[1]}[expect]
    want = want / np.linalg.norm(want)
    np.testing.assert_allclose(abs(dp @ want), np.linalg.norm(dp), rtol=1e-13)
    assert np.linalg.norm(dp) > 0.


def test_own_gradient_normals_treat_each_field_separately():
    """OWNG (3): each material takes its own gradient as the normal and its own momentum change
    toward the center of mass; the two changes are not equal and opposite pairs."""
    mass, vol, disp, grad, pk = _asym_pair()
    p, _, _ = orc.contact_node(_law(0.), 3, 0, 1e-6, mass, [4, 4], vol, disp, grad, pk, np.zeros((2, 3)), pk)
    for k in range(2):
        dp = p[k] - pk[k]
        nk = grad[k] / np.linalg.norm(grad[k])
        if np.linalg.norm(dp) > 0.:
            np.testing.assert_allclose(abs(dp @ nk), np.linalg.norm(dp), rtol=1e-13)
    assert np.linalg.norm(p[0] - pk[0]) > 0.


# ---- ResetElementsTask: a particle that leaves the grid -----------------------------------

def _leaver(frac, speed_cells=1.5, lp=0.5):
    """One particle in cell [15, 16) of a 16 mm square grid (padded to [-1, 17]) moving +x so
    that in one step it crosses x = 16 into the pad cell, which counts as leaving the grid."""
    prob = nm.Problem.create(nm.PLANE_STRAIN)
    prob.set_grid([0, 16, 0, 16, 0, 0], [1, 1, -1])
    mid = prob.add_material(nm.MAT_ISOTROPIC, "m", E=1000., nu=0.3, rho=1.0)
    prob.add_point(mid, [15.5, 8.5, 0.], vel=[0., 0., 0.], lp=(lp, lp, lp))
    prob.set("LeaveLimit", 1000)        # push back (one particle alone would be deleted by default)
    g, mats, run, _, _ = prob.describe()
    d, i = prob.particles()
    dt = run.timestep
    L = speed_cells                     # mm moved in one step
    x0 = 16. - frac * L
    d[F["pos_x"]] = x0
    d[F["orig_x"]] = x0
    d[F["vel_x"]] = L / dt
    return g, mats, run, d, i


@pytest.mark.parametrize("k", [547, 601, 513])
def test_return_to_element_bisects_ten_times(k):
    """ReturnToElement (ResetElementsTask.cpp:232-265) bisects 10 times between the retraced
    start (pos - dt v) and the position outside, keeping the last midpoint inside the element
    (half-open cells): the particle lands on the last multiple of 1/1024 of the segment before
    the cell edge. The edge sits at (k + 1/2)/1024 of the segment with k odd, where 9 passes
    would stop one 1/1024 short."""
    g, mats, run, d, i = _leaver((k + 0.5) / 1024.)
    o = orc.Oracle(g, mats, run, d, i)
    o.step(1)
    r, ri = o.particles()
    assert o.counters()["leftGrid"] == 1 and ri[nm.I["left"]][0] == 1
    dt = run.timestep
    v = r[F["vel_x"]][0]
    outside = d[F["pos_x"]][0] + v * dt
    inside = outside - dt * v
    seg = outside - inside
    s = (16. - inside) / seg
    assert abs(s * 1024 - round(s * 1024)) > 0.1        # edge well between two grid points
    want = inside + np.floor(s * 1024) / 1024 * seg
    assert abs(r[F["pos_x"]][0] - want) < 1e-9 * seg
    assert r[F["pos_x"]][0] < 16.
    assert ri[nm.I["elem"]][0] == i[nm.I["elem"]][0]     # kept its element
