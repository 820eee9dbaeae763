"""Oracle pinning: closed forms and invariants of every per-particle and
per-texel term of the coupled step.

What each check would catch:
  poly6 quadrature        a wrong normalisation (315/(64 pi h^9)) or power
  spiky / viscosity       the gradient / laplacian constants (45/(pi h^6))
                          against finite differences of the published,
                          quadrature-normalised kernels; force direction
  isolated particle       gravity, torque (cross product order), drag against
                          WaveVelocity, wave normal force, symplectic Euler,
                          velocity damping, in closed form
  boundary / foam / clamp every branch of integrate_comp.glsl
  wave                    zero / constant fields, the Fourier recurrence of
                          the leapfrog (pins lambda per axis, beta, atten),
                          mirror symmetry, the wake line
  sampler                 affine fields reproduced, clamp to edge
  InitWave                closed form at the centres and on the radius
  flux                    the particle -> wave source in closed form
"""
import math

import numpy as np
import pytest


def _free_params(cwa, h=0.05, band=1e-6):
    """Box [0,1]^3, grid cell h, no wave coupling for particles away from
    the surface (band tiny), reference torque/crest switched off."""
    p = cwa.reference_params()
    p.h = h
    p.mass = 0.01
    p.rest_rho = 1e-3
    p.gas_const = 10.0
    p.particle_radius = 0.01
    p.visc = 0.0
    p.spiky_coeff = 45.0
    p.visc_coeff = 45.0
    p.gravity[:] = [0.0, 0.0, 0.0]
    p.torque_coeff = 0.0
    p.crest_threshold = 1e9
    p.vel_damping = 0.0
    p.foam_speed = 1e9
    p.upper[:] = [1.0, 1.0, 1.0]
    p.lower[:] = [0.0, 0.0, 0.0]
    p.wave_uv_scale[:] = [1.0, 1.0]
    p.couple_band = band
    p.couple_strength = 0.0
    p.cell_size = h
    for a in range(3):
        p.grid_min[a] = 0.0
        p.grid_dims[a] = int(math.ceil(1.0 / h - 1e-4))
    return p


def _run(oracle, p, pos, vel=None, force=None, wave=(8, 8), h0=None, dt=1e-3, steps=1,
         double=True, mode="grid"):
    pos = np.asarray(pos, np.float64).reshape(-1, 4)
    s = oracle.OracleScene(p, pos, vel, force, wave, h0, None, double=double, mode=mode)
    s.step(dt, steps)
    return s.state()


def _pts(*xyz):
    return np.array([[x, y, z, 1.0] for (x, y, z) in xyz], np.float64)


# ---------------------------------------------------------------- density
def test_poly6_normalisation_by_quadrature(oracle, cwa):
    """On a fine lattice (spacing h/5) the density sum is m/s^3 * int W dV;
    W = 315/(64 pi h^9)(h^2-r^2)^3 integrates to 1."""
    h = 0.05
    p = _free_params(cwa, h)
    s = h / 5
    g = np.arange(-7, 8) * s
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel() + 0.5, Y.ravel() + 0.5, Z.ravel() + 0.5,
                    np.ones(X.size)], 1)
    st = _run(oracle, p, pos, mode="all_pairs")
    centre = np.argmin(np.abs(pos[:, :3] - 0.5).sum(1))
    expect = p.mass / s ** 3
    assert abs(st["rho"][centre] / expect - 1.0) < 2e-3


def test_isolated_particle_density(oracle, cwa):
    """rho = max(rho0, m*315/(64 pi h^3)); the reference's constants give
    0.02*315/(64*pi*1e-6) = 31 334 (rho_pres_comp.glsl:66,79)."""
    p = cwa.reference_params()
    p.couple_band = 1e-6
    st = _run(oracle, p, _pts((0.2, 0.5, 0.2)), wave=(64, 64))
    w0 = 315.0 / (64.0 * 3.141592741 * p.h ** 3)   # p.h, p.mass: the fp32 values
    assert st["rho"][0] == pytest.approx(max(1000.0, p.mass * w0), rel=1e-12)
    assert st["rho"][0] == pytest.approx(31334.0, abs=1.0)
    assert st["prs"][0] == pytest.approx(4000.0 * (p.mass * w0 - 1000.0), rel=1e-12)


def test_two_particle_density_and_eos(oracle, cwa):
    p = _free_params(cwa)
    p.rest_rho = 500.0
    h, m, r = p.h, p.mass, 0.03
    st = _run(oracle, p, _pts((0.5, 0.5, 0.5), (0.5 + r, 0.5, 0.5)))
    c = 315.0 / (64.0 * 3.141592741 * h ** 9)
    rho = m * c * (h ** 6 + (h * h - r * r) ** 3)
    assert st["rho"][0] == pytest.approx(max(500.0, rho), rel=1e-12)
    assert st["prs"][0] == pytest.approx(max(10.0 * (rho - 500.0), 0.0), rel=1e-12, abs=1e-12)
    # beyond h: no contribution, the rest-density clamp applies
    st = _run(oracle, p, _pts((0.5, 0.5, 0.5), (0.5 + h * 1.0001, 0.5, 0.5)))
    assert st["rho"][0] == pytest.approx(max(500.0, m * c * h ** 6), rel=1e-12)


def test_wave_term_in_density(oracle, cwa):
    """Inside the band: p += H rho, rho += H rho/(k r_p)  (:75-77)."""
    p = _free_params(cwa, band=0.05)
    H = 0.3
    h0 = np.full((8, 8), H)
    st = _run(oracle, p, _pts((0.5, 0.31, 0.5)), h0=h0)
    w0 = p.mass * 315.0 / (64.0 * 3.141592741 * p.h ** 3)
    pr = max(p.gas_const * (w0 - p.rest_rho), 0.0) + H * w0
    rho = w0 + H * w0 / (p.gas_const * p.particle_radius)
    assert st["rho"][0] == pytest.approx(rho, rel=1e-12)
    assert st["prs"][0] == pytest.approx(pr, rel=1e-12)


# ---------------------------------------------------------------- forces
def _quad_norm(f, h, singular=False):
    r = np.linspace(0.0, h, 200001)[1:] if singular else np.linspace(0.0, h, 200001)
    y = 4.0 * np.pi * r * r * f(r)
    return 1.0 / np.trapezoid(y, r)


def test_spiky_gradient_constant(oracle, cwa):
    """Pressure force of a pair = m (p_i+p_j)/(2 rho_j) |dW/dr| away from j,
    W = A (h-r)^3 with A from quadrature; dW/dr by central differences."""
    p = _free_params(cwa)
    h, m, r = p.h, p.mass, 0.021
    st = _run(oracle, p, _pts((0.5, 0.5, 0.5), (0.5, 0.5 + r, 0.5)))
    A = _quad_norm(lambda x: (h - x) ** 3, h)
    eps = 1e-7
    dW = (A * (h - (r + eps)) ** 3 - A * (h - (r - eps)) ** 3) / (2 * eps)
    rho, prs = st["rho"], st["prs"]
    mag = m * (prs[0] + prs[1]) / (2.0 * rho[1]) * abs(dW)
    F = st["force"][0, :3]
    assert F[1] < 0 and abs(F[0]) < 1e-9 * mag and abs(F[2]) < 1e-9 * mag
    assert -F[1] == pytest.approx(mag, rel=1e-6)
    assert st["force"][1, 1] == pytest.approx(mag, rel=1e-6)


def test_viscosity_laplacian_constant(oracle, cwa):
    """Viscosity force of a pair = mu m (v_j - v_i)/rho_j lap W, W the
    viscosity kernel of Mueller et al. 2003, lap W = W'' + 2W'/r by finite
    differences after normalising W by quadrature."""
    p = _free_params(cwa)
    p.gas_const = 0.0
    p.visc = 0.7
    h, m, r = p.h, p.mass, 0.02
    vel = np.array([[0.0, 0.0, 0.0, 0.0], [0.3, -0.2, 0.1, 0.0]])
    st = _run(oracle, p, _pts((0.5, 0.5, 0.5), (0.5 + r, 0.5, 0.5)), vel=vel, dt=1e-9)

    def f(x):
        return -x ** 3 / (2 * h ** 3) + x ** 2 / h ** 2 + h / (2 * x) - 1.0

    A = _quad_norm(f, h, singular=True)
    e = 1e-5
    W = lambda x: A * f(x)  # noqa: E731
    d1 = (W(r + e) - W(r - e)) / (2 * e)
    d2 = (W(r + e) - 2 * W(r) + W(r - e)) / (e * e)
    lap = d2 + 2.0 * d1 / r
    expect = 0.7 * m * (vel[1, :3] - vel[0, :3]) / st["rho"][1] * lap
    np.testing.assert_allclose(st["force"][0, :3], expect, rtol=2e-4)


def test_isolated_particle_closed_form(oracle, cwa):
    """One particle in the coupling band over the affine heightfield
    H(i,j) = a + b i + c j: gravity, torque (crest branch on), drag against
    WaveVelocity, wave normal force, then symplectic Euler with damping."""
    p = _free_params(cwa, band=0.5)
    p.gravity[:] = [0.1, -9.8, 0.2]
    p.torque_coeff = 0.25
    p.crest_threshold = 0.01
    p.vel_damping = 0.3
    W = H = 16
    a, b, c = 0.02, 0.003, -0.002
    jj, ii = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    h0 = a + b * ii + c * jj
    x = np.array([0.4, 0.2, 0.55])
    v = np.array([0.3, -0.1, 0.2])
    F0 = np.array([1.0, -2.0, 0.5])
    dt = 1e-3
    st = _run(oracle, p, _pts(tuple(x)), vel=np.array([[*v, 0.0]]),
              force=np.array([[*F0, 0.0]]), wave=(W, H), h0=h0, dt=dt)
    u, vv = x[0], x[2]
    tx, ty = u * W - 0.5, vv * H - 0.5
    Hs = a + b * tx + c * ty
    du = p.wave_vel_du
    wv = np.array([b * du * W / dt, c * du * H / dt, (b * du * W - c * du * H) / du])
    nrm = np.array([b * (W - 1 - tx), c * (H - 1 - ty), -1.0])
    w0 = p.mass * 315.0 / (64.0 * 3.141592741 * p.h ** 3)
    rho = max(p.rest_rho, w0 + Hs * w0 / (p.gas_const * p.particle_radius))
    torque = 0.25 * np.cross(x, F0 / 0.25)
    F = rho * np.array(p.gravity) + torque - 0.25 * (v - wv) - Hs * nrm * 0.5
    np.testing.assert_allclose(st["force"][0, :3], F, rtol=1e-9)
    nv = v + dt * F / rho
    nx = x + dt * nv
    nv = nv * (1 - p.vel_damping * dt)
    np.testing.assert_allclose(st["pos"][0, :3], nx, rtol=1e-12)
    np.testing.assert_allclose(st["vel"][0, :3], nv, rtol=1e-9)
    assert st["rho"][0] == pytest.approx(rho, rel=1e-12)


# ---------------------------------------------------------------- integrate
@pytest.mark.parametrize("axis,sign", [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)])
def test_boundary_faces(oracle, cwa, axis, sign):
    p = _free_params(cwa)
    x = np.array([0.5, 0.5, 0.5])
    x[axis] = 0.0005 if sign < 0 else 0.9995
    v = np.zeros(3)
    v[axis] = sign * 2.0
    v[(axis + 1) % 3] = 0.4
    st = _run(oracle, p, _pts(tuple(x)), vel=np.array([[*v, 0.0]]), dt=1e-3)
    bound = 0.0 if sign < 0 else 1.0
    assert st["pos"][0, axis] == bound
    assert st["vel"][0, axis] == pytest.approx(-p.bounce * v[axis], rel=1e-12)
    o = (axis + 1) % 3
    assert st["vel"][0, o] == pytest.approx(0.4, rel=1e-12)


@pytest.mark.parametrize("speed,foam", [(24.9, False), (25.1, True)])
def test_foam_branch(oracle, cwa, speed, foam):
    p = _free_params(cwa)
    p.foam_speed = 25.0
    v = np.array([0.0, 0.0, speed])
    F0 = np.array([0.0, 0.0, 0.0, 2.0])
    st = _run(oracle, p, _pts((0.5, 0.5, 0.2)), vel=np.array([[*v, 0.0]]),
              force=F0[None], dt=1e-6)
    w0 = p.mass * 315.0 / (64.0 * 3.141592741 * p.h ** 3)
    pr = p.gas_const * (w0 - p.rest_rho)
    if foam:
        assert st["vel"][0, 2] == pytest.approx(p.foam_vel * speed, rel=1e-9)
        assert st["rho"][0] == pytest.approx(p.foam_rho * w0, rel=1e-12)
        assert st["prs"][0] == pytest.approx(p.foam_pressure * pr, rel=1e-12)
        assert st["force"][0, 3] == pytest.approx(1.0, rel=1e-12)
    else:
        assert st["vel"][0, 2] == pytest.approx(speed, rel=1e-9)
        assert st["rho"][0] == pytest.approx(w0, rel=1e-12)
        assert st["force"][0, 3] == pytest.approx(2.0, rel=1e-12)


@pytest.mark.parametrize("wtype,extra,y0,vy,y_end,vy_factor", [
    (0.5, -0.2, 0.85, 2.0, 0.8, -0.3),     # wake mode: ceiling at upper.y - 0.2
    (1.0, -0.2, 0.85, 2.0, None, 1.0),     # splash mode: no ceiling
    (0.5, 0.1, 0.9995, 2.0, 1.0, -0.3),    # the reference's +0.1: box clamp only
    (0.5, -0.2, 0.9995, 2.0, 0.8, 0.09),   # box clamp, then the ceiling: two bounces
])
def test_wake_ceiling(oracle, cwa, wtype, extra, y0, vy, y_end, vy_factor):
    """CheckBoundary's wake-mode ceiling (integrate_comp.glsl:170-177): for
    0 < wave type < 1, after the box clamps, y > upper.y + extra is clamped
    to upper.y + extra with v.y *= -bounce.  With the reference's extra of
    +0.1 it can never fire (the box already clamped y to upper.y); with a
    negative extra it does, and a particle leaving through the box top is
    reflected twice (v.y *= (-0.3)^2), in the shader's order."""
    p = _free_params(cwa)
    p.wave_type = wtype
    p.ceiling_extra = extra
    dt = 1e-3
    st = _run(oracle, p, _pts((0.5, y0, 0.5)), vel=np.array([[0.0, vy, 0.0, 0.0]]), dt=dt)
    if y_end is None:
        assert st["pos"][0, 1] == pytest.approx(y0 + dt * vy, rel=1e-12)
        assert st["vel"][0, 1] == pytest.approx(vy, rel=1e-12)
    else:
        assert st["pos"][0, 1] == pytest.approx(y_end, rel=1e-7)
        assert st["vel"][0, 1] == pytest.approx(vy_factor * vy, rel=1e-7)


def test_surface_clamp(oracle, cwa):
    """y < H -> y = H - particle_radius (integrate_comp.glsl:79-83)."""
    p = _free_params(cwa)
    Hs = 0.5
    st = _run(oracle, p, _pts((0.3, 0.45, 0.6)), h0=np.full((8, 8), Hs))
    assert st["pos"][0, 1] == pytest.approx(Hs - p.particle_radius, rel=1e-12)
    st = _run(oracle, p, _pts((0.3, 0.55, 0.6)), h0=np.full((8, 8), Hs))
    assert st["pos"][0, 1] == pytest.approx(0.55, rel=1e-12)


def test_coincident_pair_is_finite(oracle, cwa):
    """Two particles on one point (as boundary clamping can produce) have no
    pressure direction; the step stays finite."""
    p = _free_params(cwa)
    st = _run(oracle, p, _pts((0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), double=False)
    assert np.isfinite(st["force"]).all() and np.isfinite(st["pos"]).all()


# ---------------------------------------------------------------- flux
def test_particle_flux_into_wave(oracle, cwa):
    """One particle inside the band pushes (m/rho0) v_y (1-q^2)^2 into its
    column; with the wave texels on the column centres the new height there
    is kappa dt flux / cell^2 (zero initial height)."""
    p = _free_params(cwa, h=0.1, band=0.08)
    p.couple_strength = 2.5
    p.rest_rho = 800.0
    x = np.array([0.34, 0.05, 0.77])
    vy = -0.6
    dt = 1e-3
    st = _run(oracle, p, _pts(tuple(x)), vel=np.array([[0.0, vy, 0.0, 0.0]]), wave=(10, 10),
              dt=dt)
    q = 0.05 / p.couple_band
    flux = p.mass / p.rest_rho * vy * (1 - q * q) ** 2
    assert st["flux"][0] == pytest.approx(flux, rel=1e-12)
    hnew = st["height0"]
    expect = np.zeros((10, 10))
    expect[7, 3] = 2.5 * dt * flux / p.cell_size ** 2
    # texel centres sit on column centres up to the fp32 rounding of the cell
    # size, so bilinear weights leak at most ~1e-7 of the peak to neighbours
    np.testing.assert_allclose(hnew, expect, rtol=0, atol=1e-6 * abs(expect[7, 3]))
    assert hnew.sum() == pytest.approx(expect[7, 3], rel=1e-6)


# ---------------------------------------------------------------- wave
def _wave_params(cwa, lx=0.11, lz=0.07, atten=0.97, beta=0.003, wtype=1.0):
    p = cwa.reference_params()
    p.lambda_x, p.lambda_z, p.atten, p.beta, p.wave_type = lx, lz, atten, beta, wtype
    return p


def test_wave_zero_and_constant(oracle, cwa):
    p = _wave_params(cwa)
    z = np.zeros((20, 30))
    assert np.all(oracle.wave_evolve(p, z, z, double=True) == 0.0)
    c = np.full((20, 30), 0.37)
    out = oracle.wave_evolve(p, c, c, double=True)
    np.testing.assert_allclose(out, p.atten * 0.37, rtol=1e-14)


def test_wave_fourier_recurrence(oracle, cwa):
    """For h = A cos(kx i + kz j), interior texels obey
    h+ = atten [(2 - beta - 4 lx sin^2(kx/2) - 4 lz sin^2(kz/2)) h - (1-beta) h-],
    the leapfrog characteristic polynomial; lx != lz pins the axes."""
    p = _wave_params(cwa)
    W, H = 40, 30
    kx, kz = 0.7, 0.3
    jj, ii = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    mode = np.cos(kx * ii + kz * jj)
    A, B = 0.8, 1.3
    out = oracle.wave_evolve(p, B * mode, A * mode, double=True)
    g = (2 - p.beta - 4 * p.lambda_x * math.sin(kx / 2) ** 2
         - 4 * p.lambda_z * math.sin(kz / 2) ** 2)
    expect = p.atten * (g * B - (1 - p.beta) * A) * mode
    np.testing.assert_allclose(out[1:-1, 1:-1], expect[1:-1, 1:-1], atol=1e-13)
    swapped = _wave_params(cwa, lx=0.07, lz=0.11)
    out2 = oracle.wave_evolve(swapped, B * mode, A * mode, double=True)
    assert np.abs(out2[1:-1, 1:-1] - expect[1:-1, 1:-1]).max() > 1e-3


def test_wave_mirror_symmetry_and_wake(oracle, cwa):
    rng = np.random.default_rng(3)
    W, H = 24, 17
    f = rng.normal(size=(H, W))
    f = f + f[:, ::-1]
    g = rng.normal(size=(H, W))
    g = g + g[:, ::-1]
    out = oracle.wave_evolve(_wave_params(cwa), f, g, double=True)
    np.testing.assert_allclose(out, out[:, ::-1], atol=1e-13)
    # wake mode: +wake_amount on the column i == W/2 where the value > 1e-4
    pw = _wave_params(cwa, wtype=0.5)
    base = oracle.wave_evolve(_wave_params(cwa), f, g, double=True)
    wk = oracle.wave_evolve(pw, f, g, double=True)
    diff = wk - base
    col = W // 2
    assert np.all(diff[:, np.arange(W) != col] == 0)
    np.testing.assert_allclose(diff[:, col],
                               np.where(base[:, col] > pw.wake_threshold, pw.wake_amount, 0.0),
                               atol=1e-12)


def test_sampler_affine_and_clamp(oracle):
    W, H = 13, 9
    jj, ii = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    img = (0.25 + 0.125 * ii - 0.0625 * jj).astype(np.float32)
    rng = np.random.default_rng(0)
    for _ in range(200):
        u = rng.uniform(0.5 / W, 1 - 0.5 / W)
        v = rng.uniform(0.5 / H, 1 - 0.5 / H)
        x, y = u * W - 0.5, v * H - 0.5
        assert oracle.tex_sample(img, u, v) == pytest.approx(0.25 + 0.125 * x - 0.0625 * y,
                                                             abs=2e-6)
    # clamp to edge beyond the borders
    assert oracle.tex_sample(img, -3.0, -1.0) == pytest.approx(img[0, 0])
    assert oracle.tex_sample(img, 5.0, 7.0) == pytest.approx(img[-1, -1])
    assert oracle.tex_sample(img, 1.7, 0.5 + 0.5 / H) == pytest.approx(
        0.5 * (img[H // 2, -1] + img[H // 2 + 1, -1]), abs=1e-6)


def test_init_wave_closed_form(oracle):
    W = H = 64
    s = oracle.init_wave(1.0, W, H, double=True)          # splash
    assert s[16, 16] == 0.5 and s[48, 48] == 0.5
    assert s[16, 16 + 5] == 0.0 and s[0, 0] == 0.0
    d = math.hypot(1.0, 2.0)      # texel (i=17, j=18) from the centre (16, 16)
    tt = (d - 5.0) / -5.0         # smoothstep(5, 0, d) with reversed edges
    assert s[16 + 2, 16 + 1] == pytest.approx(0.5 * tt * tt * (3 - 2 * tt))
    w = oracle.init_wave(0.0, W, H, double=True)          # wave: centres on row H
    assert w[H - 1, 16] == pytest.approx(1.0 * 0.8 * 0.8 * (3 - 1.6))
    assert w[H - 1, 32] == w[H - 1, 16] == w[H - 1, 48]
    k = oracle.init_wave(0.5, W, H, double=True)          # wake
    assert k[6, 32] == pytest.approx(0.1) and k.max() == pytest.approx(0.1)
