"""The runtime-size FFT kernels (fft.hip), which serve every grid shape without a size-specialised
instantiation, forced on for every shape with FHD_FFT_GENERIC=1 (read when a context is created):
the transforms against numpy and the three schemes' steps against the FP64 oracle, to the same
tolerances as the specialised path (relative L2 < 1e-5 for fields and trajectories)."""
import numpy as np
import pytest

import fluam_amd as F
import oracle_lib as O
from fluam_amd._abi import (make_params, SCHEME_INCOMPRESSIBLE, SCHEME_STOKES_LIMIT,
                            SCHEME_QUASI_NEUTRALLY_BUOYANT)

pytestmark = pytest.mark.gpu

WCA = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)


@pytest.fixture(autouse=True)
def generic(monkeypatch):
    monkeypatch.setenv("FHD_FFT_GENERIC", "1")


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("shape", [(6, 12, 16), (24, 48, 16), (32, 32, 32), (128, 128, 128), (18, 24, 36)])
def test_generic_fft_matches_numpy(shape):
    print(f"[gpu-case] generic fft {shape}", flush=True)
    rng = np.random.default_rng(7 + sum(shape))
    a = rng.standard_normal(shape)
    got = F.fft_r2c(a)
    ref = np.fft.rfftn(a)
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()
    back = F.fft_c2r(ref, shape[2])
    assert np.abs(back - a * a.size).max() < 1e-5 * np.abs(a * a.size).max()


def _run(p, nsteps):
    print(f"[gpu-case] generic scheme={p.scheme} cells=({p.mx},{p.my},{p.mz}) np={p.np} k={p.kernel_points}",
          flush=True)
    g, o = F.Simulation(p), O.Oracle(p)
    q0 = o.particles() if p.np else None
    g.step(nsteps)
    g.sync()
    o.step(nsteps)
    assert rel(g.velocity(), o.velocity()) < (1e-5 if p.np == 0 else 1e-4)
    assert g.max_divergence() < 1e-6
    if p.np:
        qg, qo = g.particles(), o.particles()
        assert rel(qg, qo) < 1e-5
        assert rel(qg - q0, qo - q0) < 1e-3
    return g, o


def test_generic_incompressible():
    _run(make_params(SCHEME_INCOMPRESSIBLE, cells=(24, 16, 48), box=(24.0, 20.0, 36.0), temperature=1e-2, dt=0.1,
                     seed=3, background=(0.02, -0.01, 0.03)), 20)


@pytest.mark.parametrize("cells,kernel", [((32, 32, 32), 4), ((12, 24, 36), 3)])
def test_generic_stokes(cells, kernel):
    _run(make_params(SCHEME_STOKES_LIMIT, cells=cells, np=200, kernel_points=kernel, temperature=1e-3, dt=0.05,
                     seed=23, initial_positions=3, **WCA), 30)


def test_generic_qnb():
    g, o = _run(make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(24, 24, 36), np=150, kernel_points=3,
                            temperature=1e-3, dt=0.05, seed=33, initial_positions=3, mass=2.0, **WCA), 30)
    assert rel(g.particle_velocities(), o.particle_velocities()) < 1e-4
