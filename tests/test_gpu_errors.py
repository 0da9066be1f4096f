"""Device error flags (fluam's errorKernel, src/allocateErrorArray.cu:22-32, checked at sample
steps by src/gpuToHostStokesLimit.cu:23-30) and stream ordering around batch steps, on the GPU.

The error is raised from finite inputs: a harmonic trap so stiff that its force overflows FP32
(k = 3e38 at a distance of several cells).  Every kernel still exits normally: the overflowed
source saturates in the fixed-point spread, the field turns non-finite, and the interpolation
pass flags the displacement and leaves the particle where it was."""
import numpy as np
import pytest

import fluam_amd as F
from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT

pytestmark = pytest.mark.gpu

WCA = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)


def _announce(name, p):
    print(f"[gpu-case] {name} cells=({p.mx},{p.my},{p.mz}) np={p.np} scheme={p.scheme} seed={p.seed}", flush=True)


def _overflow_params(tmp_path, scheme=SCHEME_STOKES_LIMIT):
    bonds = tmp_path / "stiff.dat"
    bonds.write_text("0\n1\n0 3e38 0.0 5.0 -4.0 3.0\n")
    return make_params(scheme, cells=(16, 16, 16), np=8, kernel_points=4 if scheme == SCHEME_STOKES_LIMIT else 3,
                       temperature=1e-3, dt=0.05, seed=81, initial_positions=1, bonded_forces=str(bonds), **WCA)


@pytest.mark.parametrize("scheme", [SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT])
def test_overflowing_force_raises_the_device_flag(tmp_path, scheme):
    p = _overflow_params(tmp_path, scheme)
    _announce("overflow_flag", p)
    g = F.Simulation(p)
    q0 = g.particles()
    g.step(3)
    with pytest.raises(F.FhdError, match="non-finite"):
        g.sync()
    # sticky, and reported by the completion query as well
    with pytest.raises(F.FhdError, match="non-finite"):
        g.sync()
    g.record()
    with pytest.raises(F.FhdError, match="non-finite"):
        while not g.done():
            pass
    q = g.particles()
    assert np.all(np.isfinite(q))
    assert np.array_equal(q, q0)  # the flagged displacements were not applied


@pytest.mark.parametrize("mode", [0, 2])
def test_flagged_replica_leaves_its_batch_neighbour_unchanged(tmp_path, mode):
    """A replica whose force overflows flags its own context only; the replica stepped beside it
    (in merged dispatches for mode 2) keeps the bits of its solo run."""
    bad = _overflow_params(tmp_path)
    good = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=64, kernel_points=4, temperature=1e-3,
                       dt=0.05, seed=82, initial_positions=1, **WCA)
    _announce("flagged_batch", bad)
    b = F.Batch([bad, good])
    b.set_mode(mode)
    b.step(6)
    with pytest.raises(F.FhdError, match="non-finite"):
        b.sync()
    b.replicas[1].sync()
    alone = F.Simulation(good)
    alone.step(6)
    assert np.array_equal(alone.particles(), b.replicas[1].particles())
    assert np.array_equal(alone.velocity(), b.replicas[1].velocity())
    b.close()


def test_mode1_batch_orders_replica_stream_work():
    """Work queued on a replica's own stream between steps of a one-graph batch (mode 1) -- a
    structure-factor sample and a velocity upload -- is ordered with the batch graph, so the
    replica matches a solo run given the same calls."""
    ps = [make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(16, 16, 32), np=40, kernel_points=3, temperature=1e-3,
                      dt=0.05, seed=90 + i, initial_positions=1, **WCA) for i in range(3)]
    for p in ps:
        _announce("mode1_ordering", p)
    b = F.Batch(ps)
    b.set_mode(1)
    solo = [F.Simulation(p) for p in ps]
    for s in solo:
        s.set_graph(True)
    b.step(4)
    for s in solo:
        s.step(4)
    for r, s in zip(b.replicas, solo):
        r.sk_sample()
        s.sk_sample()
        v = s.velocity()
        r.set_velocity(1.01 * v)
        s.set_velocity(1.01 * v)
    b.step(5)
    for s in solo:
        s.step(5)
    b.sync()
    for r, s in zip(b.replicas, solo):
        s.sync()
        assert np.array_equal(r.velocity(), s.velocity())
        assert np.array_equal(r.particles(), s.particles())
        assert np.array_equal(r.structure_factor()[0], s.structure_factor()[0])
    b.close()
