"""Pins the FP64 oracle (oracle/mpm_oracle.cpp) against what the mathematics fixes.

The reference holds no golden vectors, fixtures or tests for its MPM path (its repository has no tests
at all), so every check here is a closed form, an invariant, a brute-force computation or an
independent library routine (numpy SVD), each chosen so that a plausible transcription mistake (dropped
term, wrong sign or index, transposed operand) fails at least one of them.
"""
import numpy as np
import pytest

from mpm_amd import abi


def unit_params(cells=64, dt=1e-4, E=1e3, nu=0.2, rho=1.0):
    p = abi.MpmParams()
    p.grid_w = p.grid_h = cells + 4
    p.halo = 2
    p.wall = 2
    p.dx = 1.0 / cells
    p.inv_dx = float(cells)
    p.sim_w = p.sim_h = 1.0
    p.dt = dt
    p.gravity_x = 0.0
    p.gravity_y = -9.8
    p.vol = p.dx * p.dx / 4
    p.mass = rho * p.vol
    p.tait_k = 400.0
    p.tait_gamma = 1.0
    p.p_min = 0.0
    p.p_max = 4000.0
    p.mu = E / (2 * (1 + nu))
    p.lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    p.theta_c = 2.5e-2
    p.theta_s = 7.5e-3
    p.hardening = 10.0
    p.j_min = 0.1
    p.j_max = 10.0
    p.num_scenes = 1
    return p


def node_pos(p, W, H):
    """Physical position of every node: (index - halo) * dx."""
    xs = (np.arange(W) - p.halo) * p.dx
    ys = (np.arange(H) - p.halo) * p.dx
    return np.meshgrid(xs, ys)  # [H, W] each


# ---------------------------------------------------------------------------------------------- weights

def test_bspline_moments(oracle_mod):
    """Quadratic B-spline (MPM.cu:97-110): partition of unity, zero first moment, second moment 1/4 —
    the identity behind the 4/dx^2 factor of C (MPM.cu:252)."""
    p = unit_params()
    rng = np.random.default_rng(1)
    for x in rng.uniform(0.05, 0.95, 200):
        base, f, w = oracle_mod.weights(p, x)
        assert 0.5 <= f < 1.5
        d = np.arange(3) - f
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs((w * d).sum()) < 1e-14
        assert abs((w * d * d).sum() - 0.25) < 1e-14
        assert np.all(w >= 0)
        # base node: the node left of the particle, minus one
        g = x * p.inv_dx + p.halo
        assert base == int(np.floor(g - 0.5))


# ---------------------------------------------------------------------------------------------- P2G

def random_state(oracle_mod, n, seed, mat=abi.MAT_JELLY, lo=0.3, hi=0.7, spread_F=0.05):
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, (n, 2))
    v = rng.uniform(-1, 1, (n, 2))
    C = rng.uniform(-3, 3, (n, 4))
    F = np.tile([1.0, 0.0, 0.0, 1.0], (n, 1)) + rng.uniform(-spread_F, spread_F, (n, 4))
    J = rng.uniform(0.9, 1.1, n)
    m = np.full(n, mat, np.uint8) if np.isscalar(mat) else mat
    return oracle_mod.State(x, v, C, F, J, m)


@pytest.mark.parametrize("mat", [abi.MAT_FLUID, abi.MAT_JELLY, abi.MAT_SNOW])
def test_p2g_conserves_mass_momentum_angular_momentum(oracle_mod, mat):
    """P2G (MPM.cu:53-129): grid mass = N m; grid momentum = sum m v (the affine term has zero first
    moment); grid angular momentum = sum x_p x m v_p + dx^2/4 (Q10 - Q01) — a transposed Q or a sign
    error in the affine term breaks the last identity."""
    p = unit_params()
    s = random_state(oracle_mod, 300, 2, mat)
    g = oracle_mod.p2g(p, s)
    assert abs(g[..., 2].sum() - s.n * p.mass) < 1e-12 * s.n
    assert np.allclose(g[..., 0].sum(), p.mass * s.v[:, 0].sum(), rtol=1e-12, atol=1e-15)
    assert np.allclose(g[..., 1].sum(), p.mass * s.v[:, 1].sum(), rtol=1e-12, atol=1e-15)
    X, Y = node_pos(p, p.grid_w, p.grid_h)
    L_grid = (X * g[..., 1] - Y * g[..., 0]).sum()
    tau = oracle_mod.tau(p, s.material, s.F, s.J)
    D = p.dt * p.vol * 4 * p.inv_dx ** 2
    Q = p.mass * s.C - D * tau
    L_part = (s.x[:, 0] * p.mass * s.v[:, 1] - s.x[:, 1] * p.mass * s.v[:, 0]).sum() + \
        (p.dx ** 2 / 4 * (Q[:, 2] - Q[:, 1])).sum()
    assert abs(L_grid - L_part) < 1e-12 * max(1.0, abs(L_part))


def test_p2g_single_particle_by_hand(oracle_mod):
    """One particle, C = 0, F = I, J = 1: every node gets w * (m v, m) with w the tensor B-spline."""
    p = unit_params(cells=32)
    s = oracle_mod.State(np.array([[0.37, 0.61]]), np.array([[0.3, -0.2]]), material=np.array([abi.MAT_JELLY]))
    g = oracle_mod.p2g(p, s)
    bx, fx, wx = oracle_mod.weights(p, 0.37)
    by, fy, wy = oracle_mod.weights(p, 0.61)
    exp = np.zeros_like(g)
    for a in range(3):
        for b in range(3):
            w = wx[a] * wy[b]
            exp[by + b, bx + a] = [w * p.mass * 0.3, w * p.mass * -0.2, w * p.mass]
    assert np.allclose(g, exp, rtol=0, atol=1e-18)


# ---------------------------------------------------------------------------------------------- grid

def test_grid_update_walls_and_gravity(oracle_mod):
    """gridKernel (MPM.cu:150-172): v = mom / m + g dt; outward component zeroed within `wall` nodes of
    each side; massless nodes get v = 0."""
    p = unit_params(cells=16)
    W, H = p.grid_w, p.grid_h
    rng = np.random.default_rng(3)
    g = np.zeros((H, W, 3))
    g[..., 2] = rng.uniform(0.5, 2.0, (H, W))
    g[..., 0] = rng.uniform(-1, 1, (H, W))
    g[..., 1] = rng.uniform(-1, 1, (H, W))
    g[5, 5, 2] = 0.0
    out = oracle_mod.grid_update(p, g)
    m = np.where(g[..., 2] > 0, g[..., 2], 1.0)
    vx = g[..., 0] / m + p.gravity_x * p.dt
    vy = g[..., 1] / m + p.gravity_y * p.dt
    xs = np.arange(W)[None, :].repeat(H, 0)
    ys = np.arange(H)[:, None].repeat(W, 1)
    vx = np.where((xs <= p.wall) & (vx < 0), 0, vx)
    vx = np.where((xs >= W - 1 - p.wall) & (vx > 0), 0, vx)
    vy = np.where((ys <= p.wall) & (vy < 0), 0, vy)
    vy = np.where((ys >= H - 1 - p.wall) & (vy > 0), 0, vy)
    vx[5, 5] = 0
    vy[5, 5] = 0
    assert np.allclose(out[..., 0], vx, rtol=1e-14, atol=1e-14)
    assert np.allclose(out[..., 1], vy, rtol=1e-14, atol=1e-14)
    assert np.array_equal(out[..., 2], g[..., 2])


def test_grid_update_wall_rule_hand_built_cases(oracle_mod):
    """The wall rule of gridKernel (MPM.cu:159-168) on hand-built nodes, with the answers written out rather
    than computed: with wall = 2 on a 20-node grid, nodes 0..2 and 17..19 are wall nodes along each axis,
    node 3 and node 16 are not.  At a wall node only the outward component is zeroed (v_x < 0 at the left,
    v_x > 0 at the right; the same for y); an inward component passes unchanged; a non-wall node keeps
    both signs.  Gravity is off so every expected value is exact: mom = (+-0.5, +-0.25), mass 2."""
    p = unit_params(cells=16)
    p.gravity_x = p.gravity_y = 0.0
    W, H = p.grid_w, p.grid_h
    assert (W, H, p.wall) == (20, 20, 2)
    g = np.zeros((H, W, 3))
    # (x node, y node, sign of mom_x, sign of mom_y) -> expected (v_x, v_y)
    cases = {
        (2, 10, -1, +1): (0.0, 0.125),    # left wall, outward x zeroed
        (2, 11, +1, -1): (0.25, -0.125),  # left wall, inward x kept
        (3, 10, -1, +1): (-0.25, 0.125),  # first interior node: kept
        (17, 10, +1, -1): (0.0, -0.125),  # right wall (x = W - 3), outward x zeroed
        (17, 11, -1, +1): (-0.25, 0.125),  # right wall, inward kept
        (16, 10, +1, -1): (0.25, -0.125),  # x = W - 4: interior, kept
        (10, 2, +1, -1): (0.25, 0.0),     # bottom wall, outward y zeroed
        (11, 2, -1, +1): (-0.25, 0.125),  # bottom wall, inward kept
        (10, 3, +1, -1): (0.25, -0.125),  # y = 3: interior
        (10, 17, -1, +1): (-0.25, 0.0),   # top wall (y = H - 3), outward y zeroed
        (11, 17, +1, -1): (0.25, -0.125),  # top wall, inward kept
        (10, 16, -1, +1): (-0.25, 0.125),  # y = H - 4: interior
        (0, 0, -1, -1): (0.0, 0.0),       # corner: both outward, both zeroed
        (19, 19, -1, -1): (-0.25, -0.125),  # opposite corner: both inward, both kept
    }
    for (x, y, sx, sy) in cases:
        g[y, x] = [0.5 * sx, 0.25 * sy, 2.0]
    out = oracle_mod.grid_update(p, g)
    for (x, y, sx, sy), (vx, vy) in cases.items():
        assert (out[y, x, 0], out[y, x, 1]) == (vx, vy), (x, y)
        assert out[y, x, 2] == 2.0


# ---------------------------------------------------------------------------------------------- G2P

def test_g2p_reproduces_linear_field(oracle_mod):
    """APIC gather (MPM.cu:227-255) of a linear velocity field v = a + A x returns v_p = a + A x_p and
    C = A exactly (B-spline reproduces linear functions; D_p = dx^2/4 I)."""
    p = unit_params(cells=32, dt=1e-4)
    p.gravity_y = 0.0
    X, Y = node_pos(p, p.grid_w, p.grid_h)
    a = np.array([0.2, -0.1])
    A = np.array([[0.3, -0.7], [0.5, 0.1]])
    grid = np.zeros((p.grid_h, p.grid_w, 3))
    grid[..., 0] = a[0] + A[0, 0] * X + A[0, 1] * Y
    grid[..., 1] = a[1] + A[1, 0] * X + A[1, 1] * Y
    grid[..., 2] = 1.0
    rng = np.random.default_rng(4)
    xp = rng.uniform(0.3, 0.7, (50, 2))
    s = oracle_mod.State(xp, material=np.full(50, abi.MAT_JELLY, np.uint8))
    oracle_mod.g2p(p, s, grid)
    vexp = a + xp @ A.T
    assert np.allclose(s.v, vexp, rtol=0, atol=1e-13)
    assert np.allclose(s.C, np.tile(A.reshape(4), (50, 1)), rtol=0, atol=1e-11)
    # advection and F update: x += v dt, F = (I + dt C) F, J *= det(I + dt C)
    assert np.allclose(s.x, xp + vexp * p.dt, rtol=0, atol=1e-15)
    Aup = np.eye(2) + p.dt * A
    assert np.allclose(s.F, np.tile(Aup.reshape(4), (50, 1)), rtol=0, atol=1e-14)
    assert np.allclose(s.J, np.linalg.det(Aup), rtol=1e-14)


def test_oob_and_j_clamp(oracle_mod):
    """handleOOB (MPM.cu:267-301) and the J clamp (MPM.cu:265) on forced states."""
    p = unit_params(cells=32, dt=1.0)
    p.gravity_y = 0.0
    grid = np.zeros((p.grid_h, p.grid_w, 3))
    grid[..., 0] = -2.0  # every node moves left by 2 per unit time
    grid[..., 1] = 3.0
    s = oracle_mod.State(np.array([[0.5, 0.5], [0.9, 0.2]]), material=np.full(2, abi.MAT_JELLY, np.uint8))
    oracle_mod.g2p(p, s, grid)
    # x: 0.5 - 2 < 0 -> 0 and v_x = max(v_x, 0) = 0; y: 0.5 + 3 > 1 -> 1 and v_y = min(v_y, 0) = 0
    assert np.allclose(s.x, [[0.0, 1.0], [0.0, 1.0]])
    assert np.allclose(s.v, 0.0)
    # NaN velocity field -> position reset to the centre, v = 0 on that axis
    grid[..., 0] = np.nan
    s = oracle_mod.State(np.array([[0.25, 0.5]]), material=np.full(1, abi.MAT_JELLY, np.uint8))
    grid[..., 1] = 0.0
    oracle_mod.g2p(p, s, grid)
    assert s.x[0, 0] == 0.5 and s.v[0, 0] == 0.0
    # J clamp: huge expansion clamps J to j_max, huge compression to j_min
    p2 = unit_params(cells=32, dt=1.0)
    p2.gravity_y = 0.0
    X, Y = node_pos(p2, p2.grid_w, p2.grid_h)
    g2 = np.zeros((p2.grid_h, p2.grid_w, 3))
    g2[..., 0] = 20.0 * (X - 0.5)
    g2[..., 1] = 20.0 * (Y - 0.5)
    s = oracle_mod.State(np.array([[0.5, 0.5]]), material=np.full(1, abi.MAT_FLUID, np.uint8))
    oracle_mod.g2p(p2, s, g2)
    assert s.J[0] == pytest.approx(10.0)
    g2[..., 0] = -0.95 * (X - 0.5)
    g2[..., 1] = -0.95 * (Y - 0.5)
    s = oracle_mod.State(np.array([[0.5, 0.5]]), material=np.full(1, abi.MAT_FLUID, np.uint8))
    oracle_mod.g2p(p2, s, g2)
    assert s.J[0] == pytest.approx(0.1)


# ---------------------------------------------------------------------------------------------- stress

def rand_F(rng, n, lo=0.5, hi=1.8, allow_inverted=False):
    th = rng.uniform(-np.pi, np.pi, n)
    ph = rng.uniform(-np.pi, np.pi, n)
    s = rng.uniform(lo, hi, (n, 2))
    if allow_inverted:
        s[: n // 4, 1] *= -1
    U = np.stack([np.cos(th), -np.sin(th), np.sin(th), np.cos(th)], 1).reshape(n, 2, 2)
    V = np.stack([np.cos(ph), -np.sin(ph), np.sin(ph), np.cos(ph)], 1).reshape(n, 2, 2)
    F = U @ (s[:, :, None] * np.transpose(V, (0, 2, 1)))
    return F.reshape(n, 4)


def test_polar_rotation_against_numpy_svd_and_brute_force(oracle_mod):
    """R is a rotation (R^T R = I, det R = +1) with S = R^T F symmetric; for det F > 0 it equals
    U V^T from numpy's SVD; for inverted F it maximises tr(R^T F) over all rotations (brute force)."""
    rng = np.random.default_rng(5)
    F = rand_F(rng, 400, allow_inverted=True)
    R = oracle_mod.polar(F).reshape(-1, 2, 2)
    Fm = F.reshape(-1, 2, 2)
    assert np.allclose(np.transpose(R, (0, 2, 1)) @ R, np.eye(2), atol=1e-13)
    assert np.allclose(np.linalg.det(R), 1.0, atol=1e-13)
    S = np.transpose(R, (0, 2, 1)) @ Fm
    assert np.allclose(S, np.transpose(S, (0, 2, 1)), atol=1e-12)
    pos = np.linalg.det(Fm) > 0
    U, _, Vt = np.linalg.svd(Fm[pos])
    assert np.allclose(R[pos], U @ Vt, atol=1e-12)
    th = np.linspace(-np.pi, np.pi, 20001)
    for k in np.nonzero(~pos)[0][:20]:
        f = Fm[k]
        tr = np.cos(th) * (f[0, 0] + f[1, 1]) + np.sin(th) * (f[1, 0] - f[0, 1])
        best = th[np.argmax(tr)]
        ang = np.arctan2(R[k, 1, 0], R[k, 0, 0])
        assert abs(np.angle(np.exp(1j * (best - ang)))) < 1e-3


def test_svd_reconstructs_and_matches_numpy(oracle_mod):
    rng = np.random.default_rng(6)
    F = rand_F(rng, 300, allow_inverted=True)
    U, s, V = oracle_mod.svd(F)
    U = U.reshape(-1, 2, 2)
    V = V.reshape(-1, 2, 2)
    rec = U @ (s[:, :, None] * np.transpose(V, (0, 2, 1)))
    assert np.allclose(rec.reshape(-1, 4), F, atol=1e-12)
    assert np.allclose(np.linalg.det(U), 1.0) and np.allclose(np.linalg.det(V), 1.0)
    ref = np.linalg.svd(F.reshape(-1, 2, 2), compute_uv=False)
    assert np.allclose(np.sort(np.abs(s), 1), np.sort(ref, 1), atol=1e-12)
    assert np.all(np.sign(s.prod(1)) == np.sign(np.linalg.det(F.reshape(-1, 2, 2))))


def test_corotated_closed_forms(oracle_mod):
    """Fixed-corotated: tau(I) = 0; tau(sI) = [2 mu (s-1) s + lambda (s^2-1) s^2] I; rotation
    covariance tau(QF) = Q tau(F) Q^T; symmetric."""
    p = unit_params(E=1e4, nu=0.3)
    mu, lam = p.mu, p.lam
    n = 5
    s = np.linspace(0.7, 1.4, n)
    F = np.zeros((n, 4))
    F[:, 0] = s
    F[:, 3] = s
    tau = oracle_mod.tau(p, np.full(n, abi.MAT_JELLY, np.uint8), F, np.ones(n))
    exp = 2 * mu * (s - 1) * s + lam * (s * s - 1) * s * s
    assert np.allclose(tau[:, 0], exp, rtol=1e-12) and np.allclose(tau[:, 3], exp, rtol=1e-12)
    assert np.allclose(tau[:, 1], 0, atol=1e-9) and np.allclose(tau[:, 2], 0, atol=1e-9)
    rng = np.random.default_rng(7)
    F = rand_F(rng, 100)
    t0 = oracle_mod.tau(p, np.full(100, abi.MAT_JELLY, np.uint8), F, np.ones(100)).reshape(-1, 2, 2)
    assert np.allclose(t0, np.transpose(t0, (0, 2, 1)), atol=1e-9)
    th = 0.7
    Q = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    FQ = (Q @ F.reshape(-1, 2, 2)).reshape(-1, 4)
    t1 = oracle_mod.tau(p, np.full(100, abi.MAT_JELLY, np.uint8), FQ, np.ones(100)).reshape(-1, 2, 2)
    assert np.allclose(t1, Q @ t0 @ Q.T, atol=1e-8)


def test_fluid_tait_pressure(oracle_mod):
    """Tait (MPM.cu:68-74): tau = -J clamp(k((1/J)^gamma - 1), p_min, p_max) I."""
    p = unit_params()
    p.tait_k = 8.2e4
    p.tait_gamma = 3.2
    p.p_min = 0.0
    p.p_max = 8.2e5
    J = np.array([0.5, 0.9, 1.0, 1.2, 0.2])
    tau = oracle_mod.tau(p, np.zeros(5, np.uint8), np.tile([1.0, 0, 0, 1.0], (5, 1)), J)
    ph = np.float64(np.float32(8.2e4)) * ((1 / J) ** np.float64(np.float32(3.2)) - 1)
    pr = np.clip(ph, 0.0, np.float64(np.float32(8.2e5)))
    assert np.allclose(tau[:, 0], -J * pr, rtol=1e-14)
    assert tau[3, 0] == 0.0  # expansion: no tension
    assert tau[4, 0] == pytest.approx(-0.2 * np.float64(np.float32(8.2e5)))  # clamp at p_max


def test_snow_projection(oracle_mod):
    """Snow return mapping: clamped singular values, Jp det(F_E) conserved, rotations preserved."""
    p = unit_params()
    rng = np.random.default_rng(8)
    F = rand_F(rng, 200, lo=0.9, hi=1.1)
    Jp = rng.uniform(0.8, 1.2, 200)
    F2, Jp2 = oracle_mod.snow_project(p, F, Jp)
    sv = np.linalg.svd(F2.reshape(-1, 2, 2), compute_uv=False)
    lo, hi = 1 - np.float64(np.float32(2.5e-2)), 1 + np.float64(np.float32(7.5e-3))
    assert np.all(sv >= lo - 1e-12) and np.all(sv <= hi + 1e-12)
    det0 = np.linalg.det(F.reshape(-1, 2, 2))
    det1 = np.linalg.det(F2.reshape(-1, 2, 2))
    assert np.allclose(Jp2 * det1, Jp * det0, rtol=1e-12)
    inside = np.all((np.linalg.svd(F.reshape(-1, 2, 2), compute_uv=False) >= lo) &
                    (np.linalg.svd(F.reshape(-1, 2, 2), compute_uv=False) <= hi), axis=1)
    assert np.allclose(F2[inside], F[inside], atol=1e-13)
    assert np.allclose(Jp2[inside], Jp[inside], rtol=1e-13)
    # polar rotation unchanged by the projection
    assert np.allclose(oracle_mod.polar(F2), oracle_mod.polar(F), atol=1e-12)


# ---------------------------------------------------------------------------------------------- full substeps

def test_free_fall(oracle_mod):
    """Isolated fluid particle at J = 1 (Tait pressure 0), C = 0: v_n = n g dt, x follows the discrete sum."""
    p = unit_params(cells=64, dt=1e-3)
    s = oracle_mod.State(np.array([[0.5, 0.8]]), material=np.array([abi.MAT_FLUID], np.uint8))
    n = 50
    oracle_mod.run(p, s, n)
    g = p.gravity_y
    assert s.v[0, 1] == pytest.approx(n * g * p.dt, rel=1e-12)
    assert s.v[0, 0] == 0.0
    y = 0.8 + p.dt * p.dt * g * n * (n + 1) / 2
    assert s.x[0, 1] == pytest.approx(y, rel=1e-12)
    assert np.allclose(s.C, 0, atol=1e-9)


def test_rigid_translation_keeps_F(oracle_mod):
    """A jelly block moving uniformly without gravity keeps F = I, J = 1, C = 0, v constant."""
    p = unit_params(cells=64, dt=1e-4)
    p.gravity_y = 0.0
    rng = np.random.default_rng(9)
    x = rng.uniform(0.4, 0.5, (400, 2))
    v = np.tile([0.3, -0.2], (400, 1))
    s = oracle_mod.State(x, v, material=np.full(400, abi.MAT_JELLY, np.uint8))
    oracle_mod.run(p, s, 20)
    assert np.allclose(s.v, v, atol=1e-12)
    assert np.allclose(s.C, 0, atol=1e-9)
    assert np.allclose(s.F, np.tile([1.0, 0, 0, 1.0], (400, 1)), atol=1e-12)
    assert np.allclose(s.x, x + 20 * p.dt * v, atol=1e-13)


def test_openmp_matches_serial(oracle_mod):
    p = unit_params(cells=64, dt=1e-4)
    s1 = random_state(oracle_mod, 2000, 10, np.array([0, 1, 2] * 666 + [1, 1], np.uint8), spread_F=0.01)
    s2 = s1.copy()
    oracle_mod.run(p, s1, 5, threads=1)
    g = oracle_mod.run(p, s2, 5, threads=4, want_grid=True)
    assert np.allclose(s1.x, s2.x, rtol=0, atol=1e-13)
    assert np.allclose(s1.F, s2.F, rtol=0, atol=1e-11)
    assert abs(g[0].sum() - 2000 * p.mass) < 1e-12


# ---------------------------------------------------------------------------------------------- binning

def test_bin_keys_and_stable_order(oracle_mod):
    """The FP32 bin-key restatement follows floorf(fmaf(x, inv_dx, halo) - 0.5) and the order is a
    stable sort (ties keep ascending index) — equal to numpy's stable argsort."""
    p = unit_params(cells=128)
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, (5000, 2)).astype(np.float32)
    x[:10] = 1.0  # domain edge
    x[10:20] = 0.0
    keys = oracle_mod.bin_keys(p, x, 16, 16)
    nbx = (p.grid_w + 15) // 16
    bx = np.clip(np.floor(x[:, 0].astype(np.float64) * 128 + 2 - 0.5), 0, p.grid_w - 3).astype(int)
    by = np.clip(np.floor(x[:, 1].astype(np.float64) * 128 + 2 - 0.5), 0, p.grid_h - 3).astype(int)
    assert np.array_equal(keys, (by // 16) * nbx + bx // 16)
    perm = oracle_mod.stable_order(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))
    scene = np.repeat(np.arange(5), 1000).astype(np.int32)
    k2 = oracle_mod.bin_keys(p, x, 16, 16, scene)
    assert np.array_equal(k2, keys + scene.astype(np.uint32) * nbx * ((p.grid_h + 15) // 16))
    nodes = oracle_mod.bin_nodes(p, x, 16, 8)
    assert np.array_equal(nodes, (by % 8) * 16 + bx % 16)


def test_checker_slots(oracle_mod):
    """Checkerboard slots: a hand-worked 4 x 2 bin, then on every compiled bin shape a bijection whose first
    half holds the nodes with lx + ly even, no two of them 4-neighbours, each half in node order."""
    # 4 x 2 bin: even nodes 0, 2, 5, 7 -> slots 0..3; odd nodes 1, 3, 4, 6 -> slots 4..7
    assert oracle_mod.checker_slots(np.arange(8), 4, 2).tolist() == [0, 4, 1, 5, 6, 2, 7, 3]
    for bx, by in [(8, 8), (16, 8), (16, 16), (32, 16)]:
        nb = bx * by
        nodes = np.arange(nb, dtype=np.uint32)
        slots = oracle_mod.checker_slots(nodes, bx, by).astype(np.int64)
        assert np.array_equal(np.sort(slots), np.arange(nb))
        lx, ly = nodes % bx, nodes // bx
        even = (lx + ly) % 2 == 0
        assert np.all(slots[even] < nb // 2) and np.all(slots[~even] >= nb // 2)
        assert np.all(np.diff(slots[even]) > 0) and np.all(np.diff(slots[~even]) > 0)
        first = np.zeros((by, bx), bool)
        first[ly[even], lx[even]] = True
        assert not np.any(first[:, 1:] & first[:, :-1]) and not np.any(first[1:, :] & first[:-1, :])


def test_layered_order(oracle_mod):
    """Hand-worked cases of the layered in-bin order (rank r of a particle among those of its node, layers
    min(r, L-1) in order, nodes in order inside a layer, the last layer in (node, rank) order), then its
    properties on random input against a brute-force restatement."""
    k = np.zeros(6, np.uint32)
    assert oracle_mod.layered_order(k, np.array([5, 3, 5, 5, 3, 1], np.uint32)).tolist() == [5, 1, 0, 4, 2, 3]
    n2 = np.array([7, 7, 7, 2, 2, 2], np.uint32)
    assert oracle_mod.layered_order(k, n2).tolist() == [3, 0, 4, 1, 5, 2]
    assert oracle_mod.layered_order(k, n2, layers=2).tolist() == [3, 0, 4, 5, 1, 2]
    rng = np.random.default_rng(12)
    keys = rng.integers(0, 7, 3000).astype(np.uint32)
    nodes = np.minimum(rng.geometric(0.08, 3000) - 1, 127).astype(np.uint32)  # some nodes above 24 particles
    perm = oracle_mod.layered_order(keys, nodes)
    assert np.array_equal(np.sort(perm), np.arange(3000))
    assert np.all(np.diff(keys[perm].astype(np.int64)) >= 0)  # bins contiguous and ascending
    rank = np.zeros(3000, np.int64)
    seen = {}
    for i in range(3000):  # rank in input order
        rank[i] = seen.get((keys[i], nodes[i]), 0)
        seen[(keys[i], nodes[i])] = rank[i] + 1
    last = oracle_mod.LAYERS - 1
    assert rank.max() > last  # the last layer holds more than one rank
    layer = np.minimum(rank, last)
    brute = sorted(range(3000), key=lambda i: (keys[i], layer[i], nodes[i], rank[i]))
    assert perm.tolist() == brute
    for b in range(7):  # every layer below the last holds distinct nodes
        for l in range(last):
            sel = (keys == b) & (layer == l)
            assert len(np.unique(nodes[sel])) == sel.sum()


# ---------------------------------------------------------------------------------------------- MLP

def test_mlp_gradient_matches_finite_differences(oracle_mod):
    rng = np.random.default_rng(12)
    params = {k: rng.normal(0, 0.4, s) for k, s in oracle_mod.MLP.SHAPES}
    net = oracle_mod.MLP(params)
    x = rng.normal(0, 1, (16, 4))
    t = rng.normal(0, 1, (16, 3))
    loss, g = net.loss_grad(x, t, (1.0, 2.0, 0.5))
    eps = 1e-6
    for k in ("W1", "b1", "W2", "b3", "W3"):
        flat = net.p[k].reshape(-1)
        for idx in rng.choice(flat.size, min(5, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = net.loss_grad(x, t, (1.0, 2.0, 0.5))
            flat[idx] = old - eps
            lm, _ = net.loss_grad(x, t, (1.0, 2.0, 0.5))
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[k].reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd))
    # forward agrees with a plain numpy restatement
    def silu(z):
        return z / (1 + np.exp(-z))
    h1 = silu(x @ params["W1"].T + params["b1"])
    h2 = silu(h1 @ params["W2"].T + params["b2"])
    y = h2 @ params["W3"].T + params["b3"]
    assert np.allclose(net.forward(x), y, atol=1e-12)
    assert loss == pytest.approx(np.mean(((y - t) ** 2 * np.array([1.0, 2.0, 0.5])).sum(1)), rel=1e-12)


def test_adam_first_step_closed_form(oracle_mod):
    """Adam (Kingma & Ba 2015): after one step from zero moments the update is lr * g / (|g| + eps)."""
    rng = np.random.default_rng(13)
    w = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50)
    m = np.zeros(50)
    v = np.zeros(50)
    w0 = w.copy()
    oracle_mod.adam(w, g, m, v, 1e-3, 0.9, 0.999, 1e-8, 1)
    assert np.allclose(w, w0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
    g2 = rng.normal(0, 1, 50)
    w1 = w.copy()
    oracle_mod.adam(w, g2, m, v, 1e-3, 0.9, 0.999, 1e-8, 2)
    m2 = 0.9 * 0.1 * g + 0.1 * g2
    v2 = 0.999 * 0.001 * g * g + 0.001 * g2 * g2
    mh = m2 / (1 - 0.9 ** 2)
    vh = v2 / (1 - 0.999 ** 2)
    assert np.allclose(w, w1 - 1e-3 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-12, atol=1e-15)


def test_philox_known_answer_vectors(oracle_mod):
    """The Philox4x32-10 restatement against the published known-answer vectors of the Random123 library
    (Salmon et al. 2011): zero counter and key, all-ones counter and key, and the pi-digits counter/key."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        got = oracle_mod.philox4x32_10(np.array([ctr], np.uint32), key)[0]
        assert [int(g) for g in got] == list(want)


def test_hose_pipe_jitter_is_standard_normal(oracle_mod):
    """Box-Muller on the Philox outputs of a million emitted particles: each axis passes a Kolmogorov-Smirnov
    test against N(0, jitter^2), mean and standard deviation within 5 standard errors, axes uncorrelated."""
    from scipy import stats
    lattice, off = oracle_mod.hose_pipe_positions((0.0, 0.0), (1000.0, 0.0), 1000, 1000, 1.0, 0.5, 7)
    z = off / 0.5
    n = z.shape[0]
    for axis in (0, 1):
        assert stats.kstest(z[:, axis], "norm").pvalue > 1e-3
    assert np.all(np.abs(z.mean(axis=0)) < 5 / np.sqrt(n))
    assert np.all(np.abs(z.std(axis=0) - 1) < 5 / np.sqrt(2 * n))
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 5 / np.sqrt(n)
    # the lattice: 1000 columns across the nozzle normal (+y), rows trailing along -x at half-spacing offsets
    assert np.allclose(lattice[:1000, 1], np.arange(1000) - 499.5) and np.allclose(lattice[::1000, 0], -(np.arange(1000) + 0.5))

