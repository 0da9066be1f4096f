"""Pins of the oracle's inertial-coupling (QNB) step with excess mass, m_e != 0
(src/runSchemeQuasiNeutrallyBuoyant.cu:336-656, src/firstStepQuasiNeutrallyBuoyant.cu:35-1152,
src/quasiNeutrallyBuoyantFunctions2.cu:5-2290,2679-3149, src/quasiNeutrallyBuoyantFunctions.cu:
5259-5795), and of the projection it uses (src/projectionDivergenceFree.cu:21-72).

  * the projection: idempotent, divergence-free, zero on discrete gradients, the identity on
    divergence-free fields, mean kept;
  * m_e -> 0 reproduces the mass-0 step (positions and fluid);
  * total momentum rho dV sum v + m_e sum u is conserved (internal forces, zero-sum drift and
    noise): a dropped or mis-signed correction term breaks it;
  * at kT = 0 the bootstrap and Adams-Bashforth steps with the excess-mass corrections equal
    the scheme rebuilt from the separately pinned spread, interpolation, projection,
    Crank-Nicolson solve, Laplacian and advection blocks."""
import numpy as np
import pytest

import oracle_lib as O
from fluam_amd._abi import make_params, SCHEME_QUASI_NEUTRALLY_BUOYANT

WCA = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)


def _grad(o, phi, h):
    """Discrete gradient of a cell-centred scalar onto the +a faces (the staggered G)."""
    mz, my, mx = o.p.mz, o.p.my, o.p.mx
    f = phi.reshape(mz, my, mx)
    gx = (np.roll(f, -1, axis=2) - f) / h[0]
    gy = (np.roll(f, -1, axis=1) - f) / h[1]
    gz = (np.roll(f, -1, axis=0) - f) / h[2]
    return np.concatenate([gx.ravel(), gy.ravel(), gz.ravel()])


def test_project_is_a_projection():
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(12, 8, 6), box=(12.0, 6.0, 7.5))
    o = O.Oracle(p)
    rng = np.random.default_rng(3)
    W = rng.standard_normal(3 * o.N)
    PW = o.project(W)
    assert np.abs(o.project(PW) - PW).max() < 1e-12 * np.abs(PW).max()
    assert np.abs(o.divergence(PW)).max() < 1e-11
    for a in range(3):  # the mean (k = 0 mode) is kept
        assert abs(PW[a * o.N:(a + 1) * o.N].mean() - W[a * o.N:(a + 1) * o.N].mean()) < 1e-13
    G = _grad(o, rng.standard_normal(o.N), (1.0, 0.75, 1.25))
    assert np.abs(o.project(G)).max() < 1e-11 * np.abs(G).max()
    # a divergence-free field (the part of W that P keeps) is left alone, and W = P W + G phi
    assert np.abs(o.divergence(W - PW - 0.0)).max() > 1e-3
    assert np.abs(o.project(PW + G) - PW).max() < 1e-11 * np.abs(G).max()


def _qnb(mass, kernel=3, **kw):
    base = dict(cells=(16, 16, 16), np=24, kernel_points=kernel, temperature=1e-3, dt=0.05, seed=21,
                initial_positions=3, mass=mass, **WCA)
    base.update(kw)
    return make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, **base)


@pytest.mark.parametrize("kernel", [3, 4])
def test_excess_mass_zero_limit(kernel):
    a, b = O.Oracle(_qnb(0.0, kernel)), O.Oracle(_qnb(1e-12, kernel))
    a.step(4)
    b.step(4)
    qa, qb = a.particles(), b.particles()
    va, vb = a.velocity(), b.velocity()
    assert np.abs(qa - qb).max() < 1e-9 * np.abs(qa - O.Oracle(_qnb(0.0, kernel)).particles()).max() + 1e-12
    assert np.abs(va - vb).max() < 1e-9 * np.abs(va).max()


@pytest.mark.parametrize("kernel", [3, 4])
def test_excess_mass_conserves_total_momentum(kernel):
    p = _qnb(2.5, kernel, initfluid=4)  # uniform vx = 0.425 (src/initializeFluid.cpp): momentum along x
    o = O.Oracle(p)
    dV = 1.0
    me = p.mass

    def momentum():
        v = o.velocity().reshape(3, -1).sum(axis=1) * p.densfluid * dV
        u = o.particle_velocities().reshape(3, -1).sum(axis=1) * me
        return v + u, u

    P0, _ = momentum()
    assert P0[0] > 100.0
    for n in range(5):
        o.step(1)
        P, pu = momentum()
        assert np.abs(P - P0).max() < 1e-11 * np.abs(P0).max(), (n, P - P0)
    assert np.abs(pu).max() > 1e-6 * np.abs(P0).max()  # the particles carry a share of it


def test_excess_mass_steps_compose_pinned_blocks(tmp_path):
    """Three steps at kT = 0 from rest with trapped particles (mass 1.7, rho 1.3): the bootstrap
    predictor-corrector and two Adams-Bashforth steps, each with the excess-mass corrections,
    rebuilt from the pinned blocks:
      du = lag + J(qh) v - J(q) v;  dp = m_e (u - J(qh) vt - du);
      dvt = P[f dV S(dp)], f = 1.5 Vp / (1.5 Vp dV rho + m_e);  g = S(dp - m_e J(qh) dvt) / rho;
      u' = J(qh) vt + J(qh) dvt + du;  dv = CN(g);  v' = vt + dv;  q' = q + dt J(qh) (v + v') / 2;
      next lag = J(qh) lap(dv).  The bootstrap's predictor runs the same corrections with no lag,
    keeps u, and leaves the lag J(q2) lap(dv_p) at q2 = q + dt/2 J(qh) v_half."""
    bonds = tmp_path / "traps.dat"
    bonds.write_text("0\n2\n0 0.9 0.0 2.0 -1.0 1.5\n1 0.6 0.3 -2.0 2.0 -1.0\n")
    rho, dt, me = 1.3, 0.1, 1.7
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(16, 12, 12), box=(16.0, 12.0, 9.0), np=2,
                    kernel_points=3, nonbonded=0, temperature=0.0, densfluid=rho, dt=dt, seed=13,
                    initial_positions=0, bonded_forces=str(bonds), mass=me)
    o = O.Oracle(p)
    dV = 1.0 * 1.0 * 0.75
    Vp = p.volume_particle
    fct = 1.5 * Vp / (1.5 * Vp * dV * rho + me)
    q = np.array([-3.1, 3.4, 1.7, -0.6, 0.9, 2.2])
    o.set_particles(q)
    v = np.zeros(3 * o.N)
    u = np.zeros(6)
    lag = np.zeros(6)
    a_prev = None

    def corrections(qh, vt, lag):
        du = lag + o.interp(qh, v) - o.interp(q, v)
        dp = me * (u - o.interp(qh, vt) - du)
        dvt = o.project(fct * dV * o.spread(qh, dp))
        g = o.spread(qh, dp - me * o.interp(qh, dvt)) / rho
        un = o.interp(qh, vt) + o.interp(qh, dvt) + du
        return o.solve_incompressible(g), un

    for n in range(3):
        qh = q + 0.5 * dt * o.interp(q, v)
        b = v + O.laplacian_term(o, v) + dt * o.spread(qh, o.bonded(qh)) / rho
        A = O.advection_term(o, v)
        if n == 0:
            vp = o.solve_incompressible(b - A)
            dvp, _ = corrections(qh, vp, np.zeros(6))
            vh = 0.5 * (v + vp + dvp)
            q2 = q + 0.5 * dt * o.interp(qh, vh)
            lag = o.interp(q2, O.laplacian_term(o, dvp))
            vt = o.solve_incompressible(b - O.advection_term(o, vh))
        else:
            vt = o.solve_incompressible(b - 1.5 * A + 0.5 * a_prev)
        a_prev = A
        dv, u = corrections(qh, vt, lag)
        vn = vt + dv
        q = q + dt * o.interp(qh, 0.5 * (v + vn))
        lag = o.interp(qh, O.laplacian_term(o, dv))
        v = vn
        o.step(1)
        assert np.abs(o.particles() - q).max() < 1e-11 * np.abs(q).max(), n
        assert np.abs(o.velocity() - v).max() < 1e-11 * max(np.abs(v).max(), 1e-300), n
        assert np.abs(o.particle_velocities() - u).max() < 1e-11 * max(np.abs(u).max(), 1e-300), n
    assert np.abs(v).max() > 1e-4 and np.abs(u).max() > 1e-5
