"""The benchmark configurations on the GPU, at their own sizes, against the FP64 oracle where it
finishes in seconds and through size-independent properties beyond that.

  * configuration 2: inertial coupling (QNB, mass 0), 64^3, 4,096 particles, 3-point kernel, WCA;
  * configuration 3: Stokes limit (midpoint), 128^3, 65,536 particles, 4-point kernel, WCA;
  * 100-step trajectories (Stokes 4-point + WCA, QNB 3-point + WCA) at 32^3;
  * configuration 4: the 256-replica batched ensemble (merged stages);
  * configuration 5: served jobs on 64^3-128^3 grids with up to 65,536 particles.

Trajectory tolerance: relative L2 error of the unwrapped positions < 1e-5 (FP32 against FP64),
and of the displacements from the start < DISP_TOL (they are ~1e-3 cells per step at kT = 1e-3,
so this is the sharper test).  Divergence of every GPU velocity < 1e-6 (absolute, 1/time)."""
import numpy as np
import pytest

import fluam_amd as F
import fluam_amd.serve as S
import oracle_lib as O
from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT

pytestmark = pytest.mark.gpu

WCA = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)
DISP_TOL = 2e-4


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _announce(name, p):
    print(f"[gpu-case] {name} cells=({p.mx},{p.my},{p.mz}) np={p.np} scheme={p.scheme} seed={p.seed}", flush=True)


def _trajectory_parity(p, nsteps, pos_tol=1e-5, disp_tol=DISP_TOL, vel_tol=1e-5):
    g = F.Simulation(p)
    o = O.Oracle(p)
    q0 = o.particles()
    assert np.abs(g.particles() - q0).max() < 1e-5
    g.step(nsteps)
    g.sync()
    o.step(nsteps)
    qg, qo = g.particles(), o.particles()
    assert np.isfinite(qg).all()
    e_pos, e_disp = rel(qg, qo), rel(qg - q0, qo - q0)
    e_vel = rel(g.velocity(), o.velocity())
    print(f"  pos {e_pos:.2e} disp {e_disp:.2e} vel {e_vel:.2e}", flush=True)
    assert e_pos < pos_tol
    assert e_disp < disp_tol
    assert e_vel < vel_tol
    assert g.max_divergence() < 1e-6
    return g, o


def config2_params(seed=5):
    return make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(64, 64, 64), np=4096, kernel_points=3, temperature=1e-3,
                       dt=0.05, seed=seed, initial_positions=3, **WCA)


def config3_params(seed=3):
    return make_params(SCHEME_STOKES_LIMIT, cells=(128, 128, 128), np=65536, kernel_points=4, temperature=1e-3,
                       dt=0.05, seed=seed, initial_positions=3, **WCA)


def test_config2_qnb64_matches_oracle_20_steps():
    p = config2_params()
    _announce("config2_parity", p)
    _trajectory_parity(p, 20)


def test_config2_qnb64_1000_steps_properties():
    """Configuration 2 over its full 1,000 steps: finite, divergence-free, and the total fluid
    momentum stays zero (internal WCA forces obey Newton's third law and the thermal drift of each
    particle integrates to zero over the grid, so nothing feeds the k = 0 mode)."""
    p = config2_params(seed=6)
    _announce("config2_1000", p)
    g = F.Simulation(p)
    g.set_graph(True)
    for _ in range(10):
        g.step(100)
        g.sync()
        v = g.velocity()
        assert np.isfinite(v).all()
        assert g.max_divergence() < 1e-6
        mom = np.abs(v.reshape(3, -1).mean(axis=1)).max()
        assert mom < 1e-6 * max(np.abs(v).max(), 1e-30), mom
    assert np.isfinite(g.particles()).all()
    assert g.nstep == 1000


def test_config3_stokes128_matches_oracle_10_steps():
    p = config3_params()
    _announce("config3_parity", p)
    g, o = _trajectory_parity(p, 10)
    # the last step's forces (WCA at q^{n+1/2}) agree too
    fg, fo = g.forces(), o.forces()
    assert rel(fg, fo) < 1e-4


@pytest.mark.parametrize("scheme,kernel", [(SCHEME_STOKES_LIMIT, 4), (SCHEME_QUASI_NEUTRALLY_BUOYANT, 3)])
def test_trajectory_parity_100_steps(scheme, kernel):
    """The 1e-5 over 100 steps bar for particle trajectories, with WCA on, at 32^3 and 1,000
    particles (a WCA neighbour for a fifth of them at this density)."""
    p = make_params(scheme, cells=(32, 32, 32), np=1000, kernel_points=kernel, temperature=1e-3, dt=0.05,
                    seed=100 + kernel, initial_positions=3, **WCA)
    _announce(f"parity100_s{scheme}", p)
    _trajectory_parity(p, 100)


def _ensemble_params(n, seed=4):
    """Configuration 4's replicas: grid edges per axis from {16,24,32,48,64}, 0-2048 particles on
    the simple-cubic start, 3-point kernel, own seeds (as bench.py's batched workload)."""
    rng = np.random.default_rng(seed)
    edges = (16, 24, 32, 48, 64)
    out = []
    for i in range(n):
        shape = tuple(int(rng.choice(edges)) for _ in range(3))
        out.append(make_params(SCHEME_STOKES_LIMIT, cells=shape, np=int(rng.integers(0, 2049)), kernel_points=3,
                               temperature=1e-3, dt=0.05, seed=1000 + i, initial_positions=1, **WCA))
    return out


def test_config4_ensemble_256_replicas():
    """The 256-replica ensemble in merged stages: a sample of replicas equals its solo run bit for
    bit, and two match the oracle."""
    params = _ensemble_params(256)
    print(f"[gpu-case] config4 256 replicas, {sum(p.mx * p.my * p.mz for p in params)} cells", flush=True)
    b = F.Batch(params)
    b.set_mode(2)
    b.step(4)
    b.sync()
    sample = [0, 17, 63, 128, 201, 255]
    for i in sample:
        alone = F.Simulation(params[i])
        alone.step(4)
        assert np.array_equal(alone.velocity(), b.replicas[i].velocity()), i
        if params[i].np:
            assert np.array_equal(alone.particles(), b.replicas[i].particles()), i
    for i in (3, 100):
        o = O.Oracle(params[i])
        o.step(4)
        assert rel(b.replicas[i].velocity(), o.velocity()) < 1e-5
        if params[i].np:
            assert rel(b.replicas[i].particles(), o.particles()) < 1e-5
    b.close()


@pytest.mark.parametrize("merged", [False, True])
def test_config5_served_jobs_at_full_size(merged):
    """Configuration 5's sizes: jobs on 128^3 (65,536 particles, Stokes 4-point), 96^3 (QNB 3-point)
    and 64^3 grids served together (per-job streams, or regrouped merged-stage batches) give each
    the bits of its solo run."""
    def params(i, edge, npart, scheme):
        return make_params(scheme, cells=(edge, edge, edge), np=npart, kernel_points=4 if scheme == SCHEME_STOKES_LIMIT else 3,
                           temperature=1e-3, dt=0.05, seed=3000 + i, initial_positions=1, **WCA)
    specs = [(128, 65536, SCHEME_STOKES_LIMIT, 12), (96, 30000, SCHEME_QUASI_NEUTRALLY_BUOYANT, 16),
             (64, 4096, SCHEME_STOKES_LIMIT, 20), (96, 0, SCHEME_STOKES_LIMIT, 10)]
    mix = [S.Job(f"j{i}", params(i, e, n, sc), steps, arrival=0.002 * i) for i, (e, n, sc, steps) in enumerate(specs)]
    for j in mix:
        _announce("config5_job", j.params)
    res = S.Server(keep_results=True, merged=merged).run(mix)
    assert res["jobs"] == len(specs) and res["evicted"] == 0
    for i, (e, n, sc, steps) in enumerate(specs):
        alone = S.Job(f"a{i}", params(i, e, n, sc), steps)
        S.Server(keep_results=True).run([alone])
        assert np.array_equal(alone.velocity, mix[i].velocity), i
        if n:
            assert np.array_equal(alone.particles, mix[i].particles), i
