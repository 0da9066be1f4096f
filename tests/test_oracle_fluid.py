"""Pins the oracle's fluid operators by dense-matrix brute force and invariants.

The dense operators are built from Kronecker products of 1D periodic difference matrices
on a small non-cubic grid with unequal spacings, so a transposed axis, a wrong neighbour or
a wrong staggering fails."""
import numpy as np
import pytest

import oracle_lib as O
from fluam_amd._abi import (make_params, SCHEME_INCOMPRESSIBLE, SCHEME_STOKES_LIMIT,
                            SCHEME_QUASI_NEUTRALLY_BUOYANT)

CELLS = (4, 6, 3)
BOX = (4.0, 9.0, 3.6)  # dx=1, dy=1.5, dz=1.2


def _ops(cells=CELLS, box=BOX):
    mx, my, mz = cells
    h = [box[0] / mx, box[1] / my, box[2] / mz]

    def fwd(m, hh):  # (f[i+1]-f[i])/h
        return (np.roll(np.eye(m), 1, axis=1) - np.eye(m)) / hh

    def bwd(m, hh):  # (f[i]-f[i-1])/h
        return (np.eye(m) - np.roll(np.eye(m), -1, axis=1)) / hh

    Ix, Iy, Iz = np.eye(mx), np.eye(my), np.eye(mz)

    def along(a, M):  # cell index i = x + mx (y + my z): kron order (z, y, x)
        mats = [Iz, Iy, Ix]
        mats[2 - a] = M
        return np.kron(mats[0], np.kron(mats[1], mats[2]))

    F = [along(0, fwd(mx, h[0])), along(1, fwd(my, h[1])), along(2, fwd(mz, h[2]))]
    B = [along(0, bwd(mx, h[0])), along(1, bwd(my, h[1])), along(2, bwd(mz, h[2]))]
    N = mx * my * mz
    # divergence of a staggered field (v_a on +a faces): sum_a B_a v_a ; gradient: F_a p
    D = np.hstack(B)
    G = np.vstack(F)
    Ls = sum(F[a] @ B[a] for a in range(3))
    L = np.kron(np.eye(3), Ls)
    return dict(F=F, B=B, D=D, G=G, L=L, Ls=Ls, N=N, h=h)


def _params(scheme, **kw):
    return make_params(scheme, cells=CELLS, box=BOX, **kw)


def test_laplacian_term_matches_dense():
    ops = _ops()
    p = _params(SCHEME_INCOMPRESSIBLE, dt=0.3, shearviscosity=1.7, densfluid=1.3)
    o = O.Oracle(p)
    v = np.random.default_rng(0).standard_normal(3 * ops["N"])
    got = O.laplacian_term(o, v)
    ref = 0.5 * p.dt * p.shearviscosity / p.densfluid * ops["L"] @ v
    np.testing.assert_allclose(got, ref, atol=1e-12)


def test_incompressible_solve_matches_dense():
    ops = _ops()
    p = _params(SCHEME_INCOMPRESSIBLE, dt=0.3, shearviscosity=1.7, densfluid=1.3)
    o = O.Oracle(p)
    W = np.random.default_rng(1).standard_normal(3 * ops["N"])
    got = o.solve_incompressible(W)
    D, G, L = ops["D"], ops["G"], ops["L"]
    P = np.eye(3 * ops["N"]) - G @ np.linalg.pinv(D @ G) @ D
    A = np.eye(3 * ops["N"]) - 0.5 * p.dt * p.shearviscosity / p.densfluid * L
    ref = np.linalg.solve(A, P @ W)
    np.testing.assert_allclose(got, ref, atol=1e-11)
    np.testing.assert_allclose(D @ got, 0, atol=1e-11)


def test_stokes_solve_matches_dense_saddle_point():
    ops = _ops()
    p = _params(SCHEME_STOKES_LIMIT, shearviscosity=0.7)
    o = O.Oracle(p)
    W = np.random.default_rng(2).standard_normal(3 * ops["N"])
    got = o.solve_stokes(W)
    D, G, L, N = ops["D"], ops["G"], ops["L"], ops["N"]
    # -eta L v + G pi = W - mean(W), D v = 0, mean(v) = 0  (least squares on the singular system)
    Wm = W.reshape(3, N) - W.reshape(3, N).mean(axis=1, keepdims=True)
    K = np.block([[-p.shearviscosity * L, G], [D, np.zeros((N, N))]])
    C = np.zeros((3, 3 * N + N))
    for a in range(3):
        C[a, a * N:(a + 1) * N] = 1.0
    Kfull = np.vstack([K, C])
    rhs = np.concatenate([Wm.ravel(), np.zeros(N), np.zeros(3)])
    sol = np.linalg.lstsq(Kfull, rhs, rcond=None)[0]
    np.testing.assert_allclose(got, sol[:3 * N], atol=1e-10)
    np.testing.assert_allclose(got.reshape(3, N).mean(axis=1), 0, atol=1e-13)


def test_noise_divergence_matches_dense_and_fdt():
    ops = _ops()
    p = _params(SCHEME_INCOMPRESSIBLE)
    o = O.Oracle(p)
    N = ops["N"]
    F, B = ops["F"], ops["B"]
    f1, f4 = 1.3, 0.7
    # columns: the six stress components XX, XY, XZ, YY, YZ, ZZ (diag at cells, off-diag at edges)
    for trace_free in (True, False):
        M = np.zeros((3 * N, 6 * N))
        T = np.eye(N) if not trace_free else None
        blocks = {}
        for (a, comp) in ((0, 0), (1, 3), (2, 5)):
            blocks[(a, comp)] = f1 * F[a]
        blocks.update({(0, 1): f4 * B[1], (0, 2): f4 * B[2], (1, 1): f4 * B[0], (1, 4): f4 * B[2],
                       (2, 2): f4 * B[0], (2, 4): f4 * B[1]})
        for (a, comp), Mb in blocks.items():
            M[a * N:(a + 1) * N, comp * N:(comp + 1) * N] += Mb
        if trace_free:
            for (a, comp) in ((0, 0), (1, 3), (2, 5)):
                for c2 in (0, 3, 5):
                    M[a * N:(a + 1) * N, c2 * N:(c2 + 1) * N] -= f1 * F[a] / 3.0
        Z = np.random.default_rng(3).standard_normal(6 * N)
        got = O.noise_divergence(o, Z, f1, f4, trace_free)
        np.testing.assert_allclose(got, M @ Z, atol=1e-12)
    # fluctuation-dissipation balance on the divergence-free subspace: with fact1^2 = 2 fact4^2
    # (fluam's 4 kT eta / 2 kT eta), P M M^T P = -fact4^2 P L P for the trace-free stress.
    f4 = 0.9
    f1 = np.sqrt(2.0) * f4
    cols = []
    for k in range(6 * N):
        e = np.zeros(6 * N)
        e[k] = 1.0
        cols.append(O.noise_divergence(o, e, f1, f4, True))
    M = np.array(cols).T
    D, G, L = ops["D"], ops["G"], ops["L"]
    P = np.eye(3 * N) - G @ np.linalg.pinv(D @ G) @ D
    np.testing.assert_allclose(P @ M @ M.T @ P, -f4 ** 2 * P @ L @ P, atol=1e-10)


def test_advection_term_properties():
    # uniform velocity: zero advection; momentum conservation: sum_i A_a[i] = 0
    p = _params(SCHEME_INCOMPRESSIBLE, dt=0.2)
    o = O.Oracle(p)
    N = o.N
    v = np.concatenate([np.full(N, 0.3), np.full(N, -1.1), np.full(N, 0.7)])
    np.testing.assert_allclose(O.advection_term(o, v), 0, atol=1e-14)
    v = np.random.default_rng(4).standard_normal(3 * N)
    A = O.advection_term(o, v).reshape(3, N)
    np.testing.assert_allclose(A.sum(axis=1), 0, atol=1e-12)


def test_advection_term_continuum_limit():
    # smooth divergence-free field: A -> dt d_b (u_a u_b) at second order
    errs = []
    for m in (16, 32):
        L = 2 * np.pi
        p = make_params(SCHEME_INCOMPRESSIBLE, cells=(m, m, m), box=(L, L, L), dt=1.0)
        o = O.Oracle(p)
        h = L / m
        c = (np.arange(m) + 0.5) * h - L / 2
        f = (np.arange(m) + 1.0) * h - L / 2
        Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
        # u = (sin y, sin z, sin x): divergence free; ux on x-faces etc.
        Zx, Yx, Xx = np.meshgrid(c, c, f, indexing="ij")
        Zy, Yy, Xy = np.meshgrid(c, f, c, indexing="ij")
        Zz, Yz, Xz = np.meshgrid(f, c, c, indexing="ij")
        ux = np.sin(Yx)
        uy = np.sin(Zy)
        uz = np.sin(Xz)
        v = np.concatenate([ux.ravel(), uy.ravel(), uz.ravel()])
        A = O.advection_term(o, v).reshape(3, m, m, m)
        # d_b(u_x u_b) = u . grad u_x = uy * d_y sin y = sin z cos y  at x-faces
        ex = np.sin(Zx) * np.cos(Yx)
        ey = np.sin(Xy) * np.cos(Zy)
        ez = np.sin(Yz) * np.cos(Xz)
        err = max(np.abs(A[0] - ex).max(), np.abs(A[1] - ey).max(), np.abs(A[2] - ez).max())
        errs.append(err)
    assert errs[1] < 0.3 * errs[0]
    assert errs[1] < 1e-2  # second order: error ~ 0.17 h^2




def test_incompressible_step_divergence_and_momentum():
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(8, 12, 6), box=(8.0, 12.0, 9.0), temperature=0.1,
                    dt=0.1, background=(0.01, -0.02, 0.03))
    o = O.Oracle(p)
    o.step(5)
    v = o.velocity()
    div = o.divergence(v)
    assert np.abs(div).max() < 1e-12
    N = o.N
    np.testing.assert_allclose(v.reshape(3, N).mean(axis=1), [0.01, -0.02, 0.03], atol=1e-13)


def test_incompressible_step_linear_part_matches_dense():
    """kT=0-free check: with zero initial velocity the first step is linear in the noise:
    v1 = A^-1 P M Z with the second substep's advection quadratic in the noise amplitude; at
    kT = 1e-24 the advection (relative size ~ |v| dt/dx ~ 1e-12) is negligible and the step equals the dense linear map."""
    ops = _ops()
    kT = 1e-24
    p = _params(SCHEME_INCOMPRESSIBLE, temperature=kT, dt=0.3, shearviscosity=1.7, densfluid=1.3)
    o = O.Oracle(p)
    o.step(1)
    v1 = o.velocity()
    N = ops["N"]
    Z = O.fluid_noise(p.seed, 0, 0, N)
    dV = np.prod(ops["h"])
    f1 = np.sqrt(4 * kT * p.shearviscosity * p.dt / (dV * p.densfluid ** 2))
    f4 = np.sqrt(2 * kT * p.shearviscosity * p.dt / (dV * p.densfluid ** 2))
    noise = O.noise_divergence(o, Z, f1, f4, True)
    D, G, L = ops["D"], ops["G"], ops["L"]
    P = np.eye(3 * N) - G @ np.linalg.pinv(D @ G) @ D
    A = np.eye(3 * N) - 0.5 * p.dt * p.shearviscosity / p.densfluid * L
    ref = np.linalg.solve(A, P @ noise)
    np.testing.assert_allclose(v1, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_incompressible_equal_time_structure_factor():
    """S(k) = <|v_hat_a(k)|^2>/N averaged over modes matches kT/(rho dV) P(k): the trace of
    P(k) is 2 for k != 0, so sum_a S_a(k) = 2 kT/(rho dV).  Equilibrium holds at any dt for the
    linear (Stokes) part; kT small makes advection negligible."""
    m = 8
    kT, rho = 1e-3, 1.0
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(m, m, m), box=(m * 1.0, m * 1.0, m * 1.0),
                    temperature=kT, dt=0.5, seed=5)
    o = O.Oracle(p)
    o.step(20)
    N = m ** 3
    acc = np.zeros(N)
    ns = 0
    for s in range(400):
        o.step(2)
        v = o.velocity().reshape(3, m, m, m)
        for a in range(3):
            acc += (np.abs(np.fft.fftn(v[a])) ** 2).ravel() / N
        ns += 1
    S = acc / ns
    S[0] = np.nan
    ratio = np.nanmean(S) / (2 * kT / (rho * 1.0))
    assert abs(ratio - 1.0) < 0.01


def test_stokes_step_runs_and_moves_particles():
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=30, kernel_points=4, cutoff=2 ** (1 / 6),
                    wca_sigma=1.0, temperature=1e-3, dt=0.05, seed=3)
    o = O.Oracle(p)
    q0 = o.particles()
    o.step(3)
    q1 = o.particles()
    d = np.abs(q1 - q0)
    assert np.isfinite(q1).all()
    assert 0 < d.max() < 1.0
    v = o.velocity()
    assert np.abs(o.divergence(v)).max() < 1e-12


def test_qnb_without_particles_equals_incompressible_when_linear():
    """QNB with np=0: the AB2 history only changes the (negligible) advection at tiny kT, so the
    trajectory equals the incompressible scheme's to the advection's order."""
    kw = dict(cells=(8, 8, 8), temperature=1e-24, dt=0.2, seed=9)
    a = O.Oracle(make_params(SCHEME_INCOMPRESSIBLE, **kw))
    b = O.Oracle(make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, **kw))
    a.step(4)
    b.step(4)
    va, vb = a.velocity(), b.velocity()
    np.testing.assert_allclose(vb, va, atol=1e-10 * np.abs(va).max())


def test_qnb_particle_velocity_is_interpolated_fluid():
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(16, 16, 16), np=20, cutoff=2 ** (1 / 6),
                    wca_sigma=1.0, temperature=1e-3, dt=0.05, seed=4)
    o = O.Oracle(p)
    o.step(3)
    q = o.particles()
    assert np.isfinite(q).all()
    u = o.particle_velocities()
    assert np.isfinite(u).all() and np.abs(u).max() > 0



def test_oracle_restart_reads_fluid_file_and_step(tmp_path):
    """The oracle's restart path: velocity and density from the file, noise counter from
    initialStep (a restarted oracle continues the uninterrupted run's noise)."""
    import fluam_amd as F
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(8, 8, 8), temperature=1e-3, dt=0.05, seed=4)
    a = O.Oracle(p)
    a.step(3)
    v3 = a.velocity()
    f = tmp_path / "r.fluidFinalConfiguration"
    F.write_fluid_file(f, (8, 8, 8), (p.lx, p.ly, p.lz), v3, densfluid=p.densfluid)
    q = p.copy()
    q.fluid = str(f).encode()
    q.initial_step = 3
    b = O.Oracle(q)
    assert np.array_equal(b.velocity(), v3)
    a.step(2)
    b.step(2)
    # the restarted run projects its loaded field once more (as a fresh run does), so it agrees
    # to round-off rather than bit for bit
    assert np.abs(a.velocity() - b.velocity()).max() < 1e-12 * max(1.0, np.abs(v3).max())


def test_stokes_midpoint_step_composes_pinned_blocks(tmp_path):
    """A Stokes-limit step of one particle in a harmonic trap at kT = 0 (no noise, no RFD drift)
    is the midpoint rule over the separately pinned blocks (src/runSchemeStokesLimit.cu:163-276):
    u(q) = J(q) (eta L)^-1 P S(q) F(q), q_half = q + dt/2 u(q), q_new = q + dt u(q_half), with F the
    trap force, S the spread (1/dV included), the Stokes solve and J the interpolation."""
    bonds = tmp_path / "trap.dat"
    bonds.write_text("0\n1\n0 0.7 0.0 3.0 -2.0 1.5\n")
    for kernel in (3, 4):
        p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 12), box=(16.0, 12.0, 9.0), np=1, kernel_points=kernel,
                        nonbonded=0, temperature=0.0, dt=0.2, seed=4, initial_positions=0,
                        bonded_forces=str(bonds))
        o = O.Oracle(p)
        q0 = np.array([-4.3, 2.6, -1.2])
        o.set_particles(q0)

        def u(q):
            W = o.spread(q, o.bonded(q))
            return o.interp(q, o.solve_stokes(W))

        qh = q0 + 0.5 * p.dt * u(q0)
        q1 = q0 + p.dt * u(qh)
        o.step(1)
        got = o.particles()
        assert np.abs(got - q1).max() < 1e-12 * np.abs(q1).max()
        assert np.linalg.norm(q1 - np.array([3.0, -2.0, 1.5])) < np.linalg.norm(q0 - np.array([3.0, -2.0, 1.5]))


def test_stokes_stochastic_step_composes_pinned_blocks(tmp_path):
    """The full Stokes-limit step at kT > 0 (src/runSchemeStokesLimit.cu:163-276): solve 1 takes
    W = S F(q) + sqrt(2) D.Sigma_1; solve 2 takes W = S F(q_half) + (kT/delta)[S(q + delta/2 w) w
    - S(q - delta/2 w) w] + sqrt(1/2) D.(Sigma_1 + Sigma_2), with the RFD direction w and the
    stress fields drawn from the step's counters.  Rebuilt here from the separately pinned
    stress fields, noise divergence, spread, Stokes solve and interpolation; it equals the
    oracle's step."""
    bonds = tmp_path / "traps.dat"
    bonds.write_text("0\n3\n0 0.7 0.0 3.0 -2.0 1.5\n1 0.4 0.5 -1.0 1.0 0.0\n2 1.1 0.0 0.0 0.0 -2.5\n")
    kT, eta, dt = 1e-3, 1.0, 0.2
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 12), box=(16.0, 12.0, 9.0), np=3, kernel_points=4,
                    nonbonded=0, temperature=kT, shearviscosity=eta, dt=dt, seed=11, initial_positions=0,
                    bonded_forces=str(bonds))
    o = O.Oracle(p)
    q0 = np.array([-4.3, 2.6, 1.0, 2.6, -1.2, 0.4, -1.2, 3.0, -2.0])  # [x0 x1 x2 y0 y1 y2 z0 z1 z2]
    o.set_particles(q0)
    N = 16 * 16 * 12
    dV = 1.0 * 0.75 * 0.75
    f1 = np.sqrt(4 * kT * eta / (dV * dt))
    f4 = np.sqrt(2 * kT * eta / (dV * dt))
    # the noise sources as the solve keeps them, drawn in Fourier space (pinned as real,
    # divergence-free, projection-invariant fields with the law of the projected divergence of
    # the pinned stress, tests/test_oracle_kspace_noise.py)
    U1, _ = O.kspace_noise(o, 0, 1, np.sqrt(2) * f4)
    U12, _ = O.kspace_noise(o, 0, 2, np.sqrt(0.5) * f4)
    w = np.array([O.normals4(p.seed, 0, 3, 0, i)[:3] for i in range(3)]).T.ravel()  # RFD directions
    delta = p.rfd_delta

    def vel(q, W):
        return o.interp(q, o.solve_stokes(W))

    W1 = o.spread(q0, o.bonded(q0)) + U1
    qh = q0 + 0.5 * dt * vel(q0, W1)
    rfd = o.spread(q0 + 0.5 * delta * w, kT / delta * w) - o.spread(q0 - 0.5 * delta * w, kT / delta * w)
    W2 = o.spread(qh, o.bonded(qh)) + rfd + U12
    q1 = q0 + dt * vel(qh, W2)
    o.step(1)
    got = o.particles()
    assert np.abs(got - q1).max() < 1e-11 * np.abs(q1).max()
    assert np.abs(q1 - q0).max() > 1e-4  # the particles moved


def test_qnb_steps_compose_pinned_blocks(tmp_path):
    """Three QNB (mass 0) steps of trapped particles at kT = 0 (no noise, no thermal drift), from
    rest: the bootstrap predictor-corrector (src/firstStepQuasiNeutrallyBuoyant.cu:35-1152), then
    Adams-Bashforth steps (src/runSchemeQuasiNeutrallyBuoyant.cu:178-677), rebuilt from the
    separately pinned Laplacian and advection terms, spread, Crank-Nicolson solve and
    interpolation: q_half = q + dt/2 J(q) v, b = v + lap(v) + dt S F(q_half) / rho; first step
    v~ = solve(b - A(v_half)) with v_half from a predictor solve(b - A(v)); later steps
    v~ = solve(b - 3/2 A(v) + 1/2 A(v_prev)); q_new = q + dt J(q_half) (v + v~) / 2."""
    bonds = tmp_path / "traps.dat"
    bonds.write_text("0\n2\n0 0.9 0.0 2.0 -1.0 1.5\n1 0.6 0.3 -2.0 2.0 -1.0\n")
    rho, dt = 1.3, 0.1
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(16, 12, 12), box=(16.0, 12.0, 9.0), np=2,
                    kernel_points=3, nonbonded=0, temperature=0.0, densfluid=rho, dt=dt, seed=13,
                    initial_positions=0, bonded_forces=str(bonds))
    o = O.Oracle(p)
    q = np.array([-3.1, 3.4, 1.7, -0.6, 0.9, 2.2])
    o.set_particles(q)
    v = np.zeros(3 * o.N)
    a_prev = None
    for n in range(3):
        qh = q + 0.5 * dt * o.interp(q, v)
        b = v + O.laplacian_term(o, v) + dt * o.spread(qh, o.bonded(qh)) / rho
        A = O.advection_term(o, v)
        if n == 0:
            vp = o.solve_incompressible(b - A)
            vt = o.solve_incompressible(b - O.advection_term(o, 0.5 * (v + vp)))
        else:
            vt = o.solve_incompressible(b - 1.5 * A + 0.5 * a_prev)
        a_prev = A
        q = q + dt * o.interp(qh, 0.5 * (v + vt))
        v = vt
        o.step(1)
        assert np.abs(o.particles() - q).max() < 1e-11 * np.abs(q).max(), n
        assert np.abs(o.velocity() - v).max() < 1e-11 * max(np.abs(v).max(), 1e-300), n
    assert np.abs(v).max() > 1e-4  # the trapped particles drove the fluid
