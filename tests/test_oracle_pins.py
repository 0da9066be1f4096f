"""Pins the oracle terms that enter only at kT > 0 or with optional keywords: the 4-point kernel
derivative, the inertial-coupling thermal drift, the extra-mobility particle terms and the
reference-variant random-finite-difference (RFD) spread.  Each is tied to an independently
pinned block (finite differences of the pinned spread, the pinned kernels, the pinned noise
slots) and then to the oracle's full step, so a wrong sign, component, factor or node choice
fails here rather than being replicated on the GPU."""
import numpy as np
import pytest

import oracle_lib as O
from fluam_amd._abi import (make_params, SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT,
                            RFD_REFERENCE)


@pytest.mark.parametrize("r", [-1.8, -1.3, -0.7, -0.2, 0.1, 0.45, 0.8, 1.2, 1.45, 1.9])
def test_ddelta4_is_derivative(r):
    eps = 1e-6
    fd = (O.delta4(r + eps) - O.delta4(r - eps)) / (2 * eps)
    assert abs(O.ddelta4(r) - fd) < 1e-8


def test_ddelta4_moments():
    """d/dx of the moment conditions: sum_j phi'(x - j) = 0 and sum_j (x - j) phi'(x - j) = -1."""
    for x in np.linspace(0.01, 0.99, 17):
        d = [O.ddelta4(x - j) for j in range(-2, 3)]
        assert abs(sum(d)) < 1e-12
        assert abs(sum((x - j) * dj for j, dj in zip(range(-2, 3), d)) + 1.0) < 1e-12


def _fd_spread_gradient(o, q, eps):
    """sum_c [S(q + eps e_c) e_c - S(q - eps e_c) e_c] / (2 eps): the divergence in q of the pinned
    spread of unit vectors, i.e. field_c = d/dq_c [phi phi phi] / dV summed over the particles."""
    npart = o.p.np
    out = np.zeros(3 * o.N)
    for c in range(3):
        F = np.zeros(3 * npart)
        F[c * npart:(c + 1) * npart] = 1.0
        dq = np.zeros(3 * npart)
        dq[c * npart:(c + 1) * npart] = eps
        out += (o.spread(q + dq, F) - o.spread(q - dq, F)) / (2 * eps)
    return out


@pytest.mark.parametrize("kernel", [3, 4])
def test_thermal_drift_spread_is_gradient_of_pinned_spread(kernel):
    """The QNB thermal-drift spread (src/quasiNeutrallyBuoyantFunctions.cu:6281-6912: slot +=
    -phi'(d)/dx phi phi kT m_f/(m_e+m_f) with functionDeltaDerived = -phi',
    src/nonBondedForceExtraPressure.cu:21-33, and the gather subtracting the slots) is
    +kT m_f/(m_e+m_f) d/dq_c [phi phi phi] on component c.  Here it must equal the central
    difference of the pinned spread; a flipped sign, a swapped component or a missing 1/h fails."""
    rng = np.random.default_rng(7 + kernel)
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(12, 10, 8), box=(12.0, 12.5, 6.4), np=4,
                    kernel_points=kernel, nonbonded=0)
    o = O.Oracle(p)
    q = np.concatenate([rng.uniform(-0.5, 0.5, 4) * L for L in (12.0, 12.5, 6.4)])
    dV = 1.0 * 1.25 * 0.8
    got = o.spread_drift(q, 1.0 / dV)
    want = _fd_spread_gradient(o, q, 1e-6)
    assert np.abs(got - want).max() < 1e-6 * np.abs(want).max()
    assert np.abs(want).max() > 0.01
    # each component's drift integrates to zero over the grid (a gradient of a unit-mass kernel)
    assert np.abs(got.reshape(3, -1).sum(axis=1)).max() < 1e-12 * np.abs(got).sum()


def test_qnb_step_at_finite_temperature_composes_pinned_blocks(tmp_path):
    """Two QNB (mass 0) steps at kT > 0, trapped particles, no WCA: the bootstrap step then an
    Adams-Bashforth step (src/runSchemeQuasiNeutrallyBuoyant.cu:178-677,
    src/firstStepQuasiNeutrallyBuoyant.cu:35-1152), rebuilt from the pinned Laplacian, advection,
    trace-free noise divergence of the step's noise slots, spread, Crank-Nicolson solve and
    interpolation, with the thermal drift taken as kT m_f/(m_e+m_f) times the finite-difference
    gradient of the pinned spread (m_e = 0: ratio 1).  Dropping the drift, the noise or the
    trace removal fails it."""
    bonds = tmp_path / "traps.dat"
    bonds.write_text("0\n2\n0 0.9 0.0 2.0 -1.0 1.5\n1 0.6 0.3 -2.0 2.0 -1.0\n")
    rho, dt, kT, eta = 1.3, 0.1, 2e-3, 0.8
    cells, box = (16, 12, 12), (16.0, 12.0, 9.0)
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=cells, box=box, np=2, kernel_points=3, nonbonded=0,
                    temperature=kT, shearviscosity=eta, densfluid=rho, dt=dt, seed=23, initial_positions=0,
                    bonded_forces=str(bonds))
    o = O.Oracle(p)
    q = np.array([-3.1, 3.4, 1.7, -0.6, 0.9, 2.2])
    o.set_particles(q)
    N = o.N
    dV = 1.0 * 1.0 * 0.75
    f1 = np.sqrt(4 * kT * eta * dt / (dV * rho * rho))
    f4 = np.sqrt(2 * kT * eta * dt / (dV * rho * rho))
    v = np.zeros(3 * N)
    a_prev = None
    drift_seen = 0.0
    for n in range(2):
        Z = O.fluid_noise(p.seed, n, 0, N)
        qh = q + 0.5 * dt * o.interp(q, v)
        drift = kT * 1.0 * _fd_spread_gradient(o, qh, 1e-6)
        drift_seen = max(drift_seen, np.abs(drift).max())
        f = o.spread(qh, o.bonded(qh)) + drift
        b = v + O.laplacian_term(o, v) + O.noise_divergence(o, Z, f1, f4, True) + dt * f / rho
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
        # the drift is taken by finite differences here: agreement to its truncation error
        assert np.abs(o.particles() - q).max() < 1e-9 * np.abs(q).max(), n
        assert np.abs(o.velocity() - v).max() < 1e-7 * np.abs(v).max(), n
    assert drift_seen > 1e-4


def _stokes_rebuild(o, p, q0, mu_e, quirk):
    """One Stokes-limit midpoint step rebuilt from pinned blocks (src/runSchemeStokesLimit.cu:
    163-276, src/stokesLimitFunctions.cu:2479-2892,3513-3894): solve 1 W = S F(q) + sqrt(2) D.Sigma_1,
    q_half = q + dt/2 J v1 [+ dt/2 mu_e F(q) + sqrt(kT mu_e dt) Z_p1]; solve 2 W = S F(q_half) + RFD
    + sqrt(1/2) D.(Sigma_1 + Sigma_2), q_new = q + dt J(q_half) v2 [+ dt mu_e F(q_half) +
    sqrt(kT mu_e dt) (Z_p1 + Z_p2)] (src/stokesLimitFunctions.cu:3676-3681,3887-3892,
    src/createCellsStokesLimitGPU.cu:61).  The RFD spread is the displaced one, or with `quirk`
    the reference's (nodes from q +- delta/2 w, weights at q)."""
    npart, N = p.np, o.N
    kT, eta, dt = p.temperature, p.shearviscosity, p.dt
    dV = (p.lx / p.mx) * (p.ly / p.my) * (p.lz / p.mz)
    f1 = np.sqrt(4 * kT * eta / (dV * dt))
    f4 = np.sqrt(2 * kT * eta / (dV * dt))
    U1, _ = O.kspace_noise(o, 0, 1, np.sqrt(2) * f4)  # the kept noise sources (drawn in Fourier space)
    U12, _ = O.kspace_noise(o, 0, 2, np.sqrt(0.5) * f4)

    def draws(stream):
        return np.array([O.normals4(p.seed, 0, stream, 0, i)[:3] for i in range(npart)]).T.ravel()

    w, zp1, zp2 = draws(3), draws(4), draws(5)
    delta = p.rfd_delta
    fact2 = np.sqrt(kT * mu_e * dt)

    def vel(q, W):
        return o.interp(q, o.solve_stokes(W))

    F0 = o.bonded(q0)
    W1 = o.spread(q0, F0) + U1
    qh = q0 + 0.5 * dt * vel(q0, W1) + 0.5 * dt * mu_e * F0 + fact2 * zp1
    if quirk:
        rfd = (o.spread_quirk(q0, q0 + 0.5 * delta * w, w, kT / delta / dV)
               - o.spread_quirk(q0, q0 - 0.5 * delta * w, w, kT / delta / dV))
    else:
        rfd = o.spread(q0 + 0.5 * delta * w, kT / delta * w) - o.spread(q0 - 0.5 * delta * w, kT / delta * w)
    Fh = o.bonded(qh)
    W2 = o.spread(qh, Fh) + rfd + U12
    return q0 + dt * vel(qh, W2) + dt * mu_e * Fh + fact2 * (zp1 + zp2)


def test_stokes_extra_mobility_step_composes_pinned_blocks(tmp_path):
    """extraMobility adds mu_e F and sqrt(kT mu_e dt) times the particle noise slots to both
    midpoint updates, with F^n in the half step and F^{n+1/2} in the full step.  The oracle's step
    equals the rebuild; the same step without those terms differs by more than the tolerance."""
    bonds = tmp_path / "traps.dat"
    bonds.write_text("0\n3\n0 0.7 0.0 3.0 -2.0 1.5\n1 0.4 0.5 -1.0 1.0 0.0\n2 1.1 0.0 0.0 0.0 -2.5\n")
    mu_e = 0.35
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 12), box=(16.0, 12.0, 9.0), np=3, kernel_points=4,
                    nonbonded=0, temperature=1e-3, dt=0.2, seed=17, initial_positions=0, extra_mobility=mu_e,
                    bonded_forces=str(bonds))
    o = O.Oracle(p)
    q0 = np.array([-4.3, 2.6, 1.0, 2.6, -1.2, 0.4, -1.2, 3.0, -2.0])
    o.set_particles(q0)
    want = _stokes_rebuild(o, p, q0, mu_e, quirk=False)
    without = _stokes_rebuild(o, p, q0, 0.0, quirk=False)
    o.step(1)
    got = o.particles()
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
    assert np.abs(want - without).max() > 1e-3


def _quirk_by_definition(p, q, qd, F, scale):
    """field += scale * F_c * prod_a phi((q_a - x_n)/h_a) over the nodes n with |qd_a - x_n| below
    the kernel's half-width (2 h for 4 points, 1.5 h for 3): the stencil chosen by the displaced
    position, the weights evaluated at the undisplaced one.  Face nodes of component c along axis
    c sit at (n + 1) h - L/2, the others at (n + 1/2) h - L/2.  Positions are kept away from the
    box faces, so no periodic image is needed."""
    cells = (p.mx, p.my, p.mz)
    box = (p.lx, p.ly, p.lz)
    phi, hw = (O.delta4, 2.0) if p.kernel_points == 4 else (O.delta3, 1.5)
    N = p.mx * p.my * p.mz
    npart = p.np
    out = np.zeros(3 * N)
    for i in range(npart):
        for c in range(3):
            rows = []
            for a in range(3):
                h = box[a] / cells[a]
                off = 1.0 if a == c else 0.5
                xs = qd[a * npart + i]
                x0 = q[a * npart + i]
                nodes = [n for n in range(-3, cells[a] + 3) if abs(xs - ((n + off) * h - box[a] / 2)) < hw * h]
                rows.append([(n % cells[a], phi((x0 - ((n + off) * h - box[a] / 2)) / h)) for n in nodes])
            for nz, wz in rows[2]:
                for ny, wy in rows[1]:
                    for nx, wx in rows[0]:
                        out[c * N + nx + p.mx * (ny + p.my * nz)] += scale * wx * wy * wz * F[c * npart + i]
    return out


@pytest.mark.parametrize("kernel", [3, 4])
def test_rfd_reference_spread_by_definition(kernel):
    """The reference's RFD spread (src/stokesLimitFunctions.cu:1741-1799,1872-1906): cell indices
    from the displaced position, kernel weights from the undisplaced one.  Particle 0 is placed so
    that +delta/2 and -delta/2 select different stencils along x; the others do not straddle."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 12, 12), box=(16.0, 12.0, 9.0), np=3, kernel_points=kernel,
                    nonbonded=0)
    o = O.Oracle(p)
    # 4-point stencils switch where the position crosses a node (x = n h + off h - L/2), 3-point ones
    # half-way between nodes
    edge = -8.0 + 3.0 * 1.0 + (0.5 if kernel == 4 else 1.0)  # a node (4-pt) or a half-node (3-pt) of x-centred rows
    q = np.array([edge + 0.02, 1.37, -2.1, 0.31, -1.9, 2.4, 0.52, 1.17, -0.8])
    w = np.array([1.0, 0.6, -0.3, 0.2, 0.9, 0.1, -0.4, 0.5, 0.7])
    F = np.array([0.9, -1.3, 0.4, 1.1, 0.2, -0.6, -0.5, 0.8, 1.7])
    for sgn in (1.0, -1.0):
        qd = q + sgn * 0.05 * w
        got = o.spread_quirk(q, qd, F, 0.7)
        want = _quirk_by_definition(p, q, qd, F, 0.7)
        assert np.abs(got - want).max() < 1e-13
    # where the displaced and undisplaced positions choose the same nodes the variant is the plain
    # spread of q; for the straddling particle it is not
    dV = 1.0 * 1.0 * 0.75
    plain = o.spread(q, F) * dV
    one = np.array([0.0, 1.0, 1.0] * 3)  # particles 1 and 2 only (component-major layout)
    assert np.abs(o.spread_quirk(q, q + 1e-7 * w, F * one, 1.0) - o.spread(q, F * one) * dV).max() < 1e-14
    assert np.abs(o.spread_quirk(q, q - 0.05 * w, F, 1.0) - plain).max() > 1e-6


def test_stokes_reference_rfd_step_composes_pinned_blocks(tmp_path):
    """rfdMode 1: the Stokes step's drift term is (kT/delta)[S_ref(q, q + delta/2 w) w -
    S_ref(q, q - delta/2 w) w] with S_ref the reference's spread pinned above; the rest of the step
    is as in the default mode."""
    bonds = tmp_path / "traps.dat"
    bonds.write_text("0\n3\n0 0.7 0.0 3.0 -2.0 1.5\n1 0.4 0.5 -1.0 1.0 0.0\n2 1.1 0.0 0.0 0.0 -2.5\n")
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 12), box=(16.0, 12.0, 9.0), np=3, kernel_points=3,
                    nonbonded=0, temperature=1e-3, dt=0.2, seed=19, initial_positions=0, rfd_mode=RFD_REFERENCE,
                    rfd_delta=0.3, bonded_forces=str(bonds))
    o = O.Oracle(p)
    q0 = np.array([-4.3, 2.6, 1.0, 2.6, -1.2, 0.4, -1.2, 3.0, -2.0])
    o.set_particles(q0)
    want = _stokes_rebuild(o, p, q0, 0.0, quirk=True)
    displaced = _stokes_rebuild(o, p, q0, 0.0, quirk=False)
    o.step(1)
    got = o.particles()
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
    assert np.abs(want - displaced).max() > 1e-9
