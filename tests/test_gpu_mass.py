"""The inertial-coupling (QNB) step with excess mass m_e != 0 on the GPU
(src/runSchemeQuasiNeutrallyBuoyant.cu:336-656 and the bootstrap of
src/firstStepQuasiNeutrallyBuoyant.cu:35-1152), against the FP64 oracle:

  * trajectories, fluid velocity and particle velocities over 30 steps (3- and 4-point kernels;
    a heavy and a light particle species), positions to 1e-5 relative, displacements to 1e-3;
  * total momentum rho dV sum v + m_e sum u conserved over 60 steps in FP32;
  * graph replay and merged batches give the bits of eager solo runs."""
import numpy as np
import pytest

import fluam_amd as F
import oracle_lib as O
from fluam_amd._abi import make_params, SCHEME_QUASI_NEUTRALLY_BUOYANT

pytestmark = pytest.mark.gpu

WCA = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _announce(name, p):
    print(f"[gpu-case] {name} cells=({p.mx},{p.my},{p.mz}) np={p.np} mass={p.mass} k={p.kernel_points} "
          f"seed={p.seed}", flush=True)


def _params(mass, kernel, **kw):
    base = dict(cells=(32, 24, 16), np=200, kernel_points=kernel, temperature=1e-3, dt=0.05, seed=41 + kernel,
                initial_positions=3, mass=mass, **WCA)
    base.update(kw)
    return make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, **base)


@pytest.mark.parametrize("mass,kernel", [(2.0, 3), (2.0, 4), (-5.0, 3)])
def test_excess_mass_parity(mass, kernel):
    # volumeParticle 8, dV 1: m_f = 12, so -5 is a particle lighter than the fluid it carries
    p = _params(mass, kernel)
    _announce("qnb_mass", p)
    g, o = F.Simulation(p), O.Oracle(p)
    q0 = o.particles()
    g.step(30)
    g.sync()
    o.step(30)
    qg, qo = g.particles(), o.particles()
    assert np.isfinite(qg).all()
    e = (rel(qg, qo), rel(qg - q0, qo - q0), rel(g.velocity(), o.velocity()),
         rel(g.particle_velocities(), o.particle_velocities()))
    print(f"  pos {e[0]:.2e} disp {e[1]:.2e} vel {e[2]:.2e} u {e[3]:.2e}", flush=True)
    assert e[0] < 1e-5 and e[1] < 1e-3 and e[2] < 1e-4 and e[3] < 1e-4
    assert g.max_divergence() < 1e-6


def test_excess_mass_conserves_momentum_gpu():
    p = _params(3.0, 3, initfluid=4)  # uniform vx = 0.425
    _announce("qnb_mass_momentum", p)
    g = F.Simulation(p)

    def momentum():
        v = g.velocity().reshape(3, -1).sum(axis=1) * p.densfluid
        u = g.particle_velocities().reshape(3, -1).sum(axis=1) * p.mass
        return v + u, u

    P0, _ = momentum()
    for _ in range(6):
        g.step(10)
        g.sync()
        P, pu = momentum()
        assert np.abs(P - P0).max() < 2e-5 * np.abs(P0).max(), P - P0
    assert np.abs(pu).max() > 1e-6 * np.abs(P0).max()


def test_excess_mass_graph_and_batch_equal_eager():
    ps = [_params(2.0, 3, seed=51), _params(1.0, 4, cells=(16, 16, 32), np=90, seed=52)]
    for p in ps:
        _announce("qnb_mass_graph_batch", p)
    eager = [F.Simulation(p) for p in ps]
    for s in eager:
        s.step(7)
    graph = F.Simulation(ps[0])
    graph.set_graph(True)
    graph.step(7)
    assert np.array_equal(graph.velocity(), eager[0].velocity())
    assert np.array_equal(graph.particles(), eager[0].particles())
    assert np.array_equal(graph.particle_velocities(), eager[0].particle_velocities())
    b = F.Batch(ps)
    b.set_mode(2)
    b.step(7)
    b.sync()
    for r, s in zip(b.replicas, eager):
        assert np.array_equal(r.velocity(), s.velocity())
        assert np.array_equal(r.particles(), s.particles())
        assert np.array_equal(r.particle_velocities(), s.particle_velocities())
    b.close()
