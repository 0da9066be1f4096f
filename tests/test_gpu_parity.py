"""GPU parity: the HIP library (FP32) against the FP64 oracle on the same seeds and noise stream.

Tolerances (stated per test):
  * fields: relative L2 error of the velocity after N steps < 1e-5;
  * particles: relative L2 error of the unwrapped trajectories < 1e-5, plus a looser check on
    the displacements themselves;
  * discrete divergence of every GPU velocity < 1e-6 (absolute, 1/time units).
"""
import os

import numpy as np
import pytest

import fluam_amd as F
import oracle_lib as O
from fluam_amd._abi import (make_params, SCHEME_INCOMPRESSIBLE, SCHEME_STOKES_LIMIT,
                            SCHEME_QUASI_NEUTRALLY_BUOYANT)

pytestmark = pytest.mark.gpu

WCA = dict(cutoff=2 ** (1 / 6), wca_sigma=1.0)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _announce(name, p):
    print(f"[gpu-case] {name} cells=({p.mx},{p.my},{p.mz}) np={p.np} scheme={p.scheme} seed={p.seed}", flush=True)


# ---------------- building blocks ----------------
def test_normals_match_oracle():
    for stream, sub in ((1, 0), (1, 2), (3, 0)):
        g = F.normals(77, 5, stream, sub, 4096)
        o = np.array([O.normals4(77, 5, stream, sub, i) for i in range(4096)])
        # FP32 Box-Muller with the hardware log2 / sin / cos against the FP64 oracle
        assert np.abs(g - o).max() < 1e-5 * max(1.0, np.abs(o).max())


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 12, 16), (32, 32, 32), (24, 48, 16), (16, 96, 128),
                                   (128, 128, 128), (16, 64, 24), (24, 16, 64), (48, 32, 48), (64, 24, 32),
                                   (32, 64, 64), (16, 96, 96)])
def test_fft_r2c_matches_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape)
    got = F.fft_r2c(a)
    ref = np.fft.rfftn(a)
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()
    back = F.fft_c2r(ref, shape[2])
    assert np.abs(back - a * a.size).max() < 1e-5 * np.abs(a * a.size).max()


@pytest.mark.parametrize("shape,plan_xy,plan_z", [((128, 128, 128), k, k % 3) for k in range(10)]
                         + [((96, 96, 96), k, (k + 1) % 3) for k in range(9)]
                         + [((48, 32, 64), 0, k) for k in range(3)])
def test_fft_every_radix_plan_matches_numpy(shape, plan_xy, plan_z):
    """The radix plans the schedule autotuner chooses among, each forced: for 128^3 (x: 16 4,
    4 16, 8 8; y and z: 16 8, 8 16, 4 4 8; xy plan 9: the register-transform kernels), for 96^3
    (x: 16 3, 3 16, 4 4 3; y and z: 16 2 3, 8 4 3, 3 8 4) and 48-point columns (16 3, 3 16,
    4 4 3).  The transforms match numpy to FP32 accuracy."""
    rng = np.random.default_rng(100 + plan_xy + shape[0])
    a = rng.standard_normal(shape)
    os.environ["FHD_PLAN_XY"] = str(plan_xy)
    os.environ["FHD_PLAN_Z"] = str(plan_z)
    try:
        got = F.fft_r2c(a)
        ref = np.fft.rfftn(a)
        back = F.fft_c2r(ref, shape[2])
    finally:
        del os.environ["FHD_PLAN_XY"], os.environ["FHD_PLAN_Z"]
    assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()
    assert np.abs(back - a * a.size).max() < 1e-5 * np.abs(a * a.size).max()


# ---------------- incompressible ----------------
def test_incompressible_config1_parity_100_steps():
    """32^3, dx=1, rho=eta=1, kT=1e-3, dt=0.05, v0=0, 100 steps: |v_gpu - v_oracle| / |v| < 1e-5."""
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=11)
    _announce("incompressible_config1", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    g.step(100)
    g.sync()
    o.step(100)
    vg, vo = g.velocity(), o.velocity()
    assert rel(vg, vo) < 1e-5
    assert g.max_divergence() < 1e-6


def test_incompressible_noncubic_with_background():
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(24, 16, 48), box=(24.0, 20.0, 36.0), temperature=1e-2,
                    dt=0.1, seed=3, background=(0.02, -0.01, 0.03), shearviscosity=0.7, densfluid=1.3)
    _announce("incompressible_noncubic", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    g.step(20)
    o.step(20)
    assert rel(g.velocity(), o.velocity()) < 1e-5
    assert g.max_divergence() < 1e-6


def test_incompressible_structure_factor_gpu():
    """Equal-time structure factor over an ensemble: S_a(k) = <|v_hat_a(k)|^2> / N with v_hat the
    unnormalised DFT; the discrete theory for Crank-Nicolson with fluam's noise gives
    sum_a S_a(k) = trace P(k) kT / (rho dV) = 2 kT / (rho dV) for every k != 0 (at any dt for
    the linear part; at kT = 1e-3 advection is negligible).  Four replicas (own seeds), 32^3,
    120 snapshots each: the mode average must match to 1%, and every |k| shell from 2 up to 5%
    (shell 1 holds 18 slowly decorrelating modes, too few for a 5% statistical bound)."""
    m, kT, rho = 32, 1e-3, 1.0
    N = m ** 3
    acc = np.zeros((m, m, m))
    ns = 0
    for seed in (41, 42, 43, 44):
        p = make_params(SCHEME_INCOMPRESSIBLE, cells=(m, m, m), temperature=kT, dt=0.5, seed=seed)
        _announce(f"structure_factor_seed{seed}", p)
        g = F.Simulation(p)
        g.step(100)
        for _ in range(120):
            g.step(5)
            v = g.velocity().reshape(3, m, m, m)
            acc += sum(np.abs(np.fft.fftn(v[a])) ** 2 for a in range(3)) / N
            ns += 1
    S = acc / ns / (2 * kT / rho)
    S[0, 0, 0] = np.nan
    assert abs(np.nanmean(S) - 1.0) < 0.01
    k = np.fft.fftfreq(m) * m
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    shells = np.round(kk).astype(int)
    for sh in range(2, m // 2):
        sel = (shells == sh) & np.isfinite(S)
        assert abs(S[sel].mean() - 1.0) < 0.05, (sh, S[sel].mean())


def test_incompressible_deterministic_replay():
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=5)
    a = F.Simulation(p)
    b = F.Simulation(p)
    a.step(10)
    b.step(10)
    assert np.array_equal(a.velocity(), b.velocity())


# ---------------- Stokes limit ----------------
def _particle_parity(p, nsteps, tol=1e-5, disp_tol=1e-3):
    g = F.Simulation(p)
    o = O.Oracle(p)
    q0 = o.particles()
    np.testing.assert_allclose(g.particles(), q0, atol=1e-6)
    g.step(nsteps)
    g.sync()
    o.step(nsteps)
    qg, qo = g.particles(), o.particles()
    assert np.isfinite(qg).all()
    assert rel(qg, qo) < tol
    assert rel(qg - q0, qo - q0) < disp_tol
    assert rel(g.velocity(), o.velocity()) < 1e-4
    assert g.max_divergence() < 1e-6
    return g, o


@pytest.mark.parametrize("kernel", [3, 4])
def test_stokes_parity_no_wca(kernel):
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=300, kernel_points=kernel, nonbonded=0,
                    temperature=1e-3, dt=0.05, seed=21, initial_positions=0, **WCA)
    _announce(f"stokes_no_wca_k{kernel}", p)
    _particle_parity(p, 50)


@pytest.mark.parametrize("cells,scheme,kernel", [((12, 24, 8), SCHEME_STOKES_LIMIT, 4),
                                                  ((8, 16, 12), SCHEME_QUASI_NEUTRALLY_BUOYANT, 3),
                                                  ((36, 8, 16), SCHEME_STOKES_LIMIT, 3)])
def test_small_box_tiles_parity(cells, scheme, kernel):
    """Boxes whose tiles are not 8^3 with at least two per axis (tile edge 4 along 12- and
    36-cell axes, a single tile along an 8-cell axis) take the general spreading path, the
    folded node indexing and the small-box WCA cull; trajectories match the oracle."""
    p = make_params(scheme, cells=cells, np=60, kernel_points=kernel, temperature=1e-3, dt=0.05,
                    seed=41 + kernel, initial_positions=3, **WCA)
    _announce(f"small_box_{cells}", p)
    _particle_parity(p, 20)


def test_stokes_parity_with_wca_lattice():
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=512, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=22, initial_positions=1, **WCA)
    _announce("stokes_wca_lattice", p)
    _particle_parity(p, 50)


def test_stokes_wca_forces_match():
    rng = np.random.default_rng(7)
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 24), np=200, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=23, initial_positions=0, **WCA)
    _announce("stokes_wca_forces", p)
    # 100 pairs at separations 0.9..1.12 sigma, pair centres on a coarse lattice
    g = F.Simulation(p)
    o = O.Oracle(p)
    centres = np.array([[x, y, z] for x in (-6, -2, 2, 6) for y in (-6, -2, 2, 6)
                        for z in (-10, -7, -4, -1, 2, 5, 8)], dtype=float)[:100]
    d = rng.standard_normal((100, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= rng.uniform(0.9, 1.12, (100, 1)) / 2
    pts = np.concatenate([centres + d, centres - d])  # (200, 3)
    q = pts.T.ravel()
    g.set_particles(q)
    o.set_particles(q)
    g.step(1)
    o.step(1)
    fg, fo = g.forces(), o.forces()
    assert np.abs(fo).max() > 0
    assert rel(fg, fo) < 1e-4


def test_wca_list_overflow_falls_back_to_tiles():
    """48 particles packed within 0.9 sigma of one point (each has all the others within the
    list radius, more than the 32 a list holds) among 150 spread-out ones: the packed particles
    take the tile scan, and every force matches the oracle."""
    rng = np.random.default_rng(11)
    # dt small enough that the steep forces move nothing far: the step only measures them
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=198, kernel_points=4, temperature=1e-3,
                    dt=1e-9, seed=29, initial_positions=0, **WCA)
    _announce("wca_list_overflow", p)
    u = []
    while len(u) < 48:  # no two packed particles closer than 0.3 sigma
        c = rng.uniform(-0.9, 0.9, 3)
        if np.linalg.norm(c) <= 0.9 and all(np.linalg.norm(c - v) >= 0.3 for v in u):
            u.append(c)
    u = np.array(u)
    lattice = np.array([[x, y, z] for x in range(-7, 8, 2) for y in range(-7, 8, 2) for z in range(-7, 8, 2)],
                       dtype=float)
    far = lattice[np.linalg.norm(lattice - 2.5, axis=1) > 3.0][:150] + 0.1 * rng.standard_normal((150, 3))
    pts = np.concatenate([u + 2.5, far])
    q = pts.T.ravel()
    g, o = F.Simulation(p), O.Oracle(p)
    g.set_particles(q)
    o.set_particles(q)
    g.step(1)
    o.step(1)
    fg, fo = g.forces(), o.forces()
    assert np.isfinite(fg).all() and np.abs(fo).max() > 0
    assert rel(fg, fo) < 1e-4


def test_wca_lists_rebuilt_as_particles_move():
    """kT = 0.05 moves particles ~0.02 per step, so the neighbour lists (cutoff + a quarter
    cutoff) go stale every few steps and are rebuilt; 40 steps of 400 particles at 0.1 per cell
    (pairs form and break) match the oracle's trajectories."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=400, kernel_points=4, temperature=0.05,
                    dt=0.05, seed=31, initial_positions=3, **WCA)
    _announce("wca_list_rebuild", p)
    _particle_parity(p, 40)


def test_wca_lists_graph_replay_equals_eager():
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=400, kernel_points=4, temperature=0.05,
                    dt=0.05, seed=32, initial_positions=3, **WCA)
    _announce("wca_list_graph", p)
    a, b = F.Simulation(p), F.Simulation(p)
    b.set_graph(True)
    for _ in range(4):
        a.step(5)
        b.step(5)
        assert np.array_equal(a.particles(), b.particles())
        assert np.array_equal(a.forces(), b.forces())


def test_stokes_reference_rfd_mode_runs():
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=64, kernel_points=3, nonbonded=0,
                    temperature=1e-3, dt=0.05, seed=24, initial_positions=0, rfd_mode=F.RFD_REFERENCE, **WCA)
    _announce("stokes_rfd_reference", p)
    _particle_parity(p, 20)


def test_stokes_extra_mobility():
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=64, kernel_points=3, nonbonded=0,
                    temperature=1e-3, dt=0.05, seed=25, initial_positions=0, extra_mobility=0.3, **WCA)
    _announce("stokes_extra_mobility", p)
    _particle_parity(p, 20)


def test_stokes_crowded_tile_overflow():
    """1,500 particles inside one 8^3-cell tile: the tile's candidate list overflows its capacity
    (1,024), so spreading takes the re-culling path for that tile and its neighbours."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=1500, kernel_points=4, nonbonded=0,
                    temperature=1e-3, dt=0.05, seed=27, initial_positions=0, **WCA)
    _announce("stokes_crowded_tile", p)
    rng = np.random.default_rng(27)
    pts = rng.uniform(-7.9, -0.1, (1500, 3))  # cells 8..15 of the box [-16, 16)
    q = pts.T.ravel()
    g = F.Simulation(p)
    o = O.Oracle(p)
    g.set_particles(q)
    o.set_particles(q)
    q0 = o.particles()
    g.step(5)
    g.sync()
    o.step(5)
    qg, qo = g.particles(), o.particles()
    assert np.isfinite(qg).all()
    assert rel(qg, qo) < 1e-5
    assert rel(qg - q0, qo - q0) < 1e-3
    assert rel(g.velocity(), o.velocity()) < 1e-4


def test_stokes_deterministic_replay():
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=2000, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=26, initial_positions=0, **WCA)
    a = F.Simulation(p)
    b = F.Simulation(p)
    a.step(5)
    b.step(5)
    assert np.array_equal(a.particles(), b.particles())
    assert np.array_equal(a.velocity(), b.velocity())


# ---------------- QNB ----------------
@pytest.mark.parametrize("kernel", [3, 4])
def test_qnb_parity(kernel):
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(32, 32, 32), np=256, kernel_points=kernel,
                    temperature=1e-3, dt=0.05, seed=31, initial_positions=1, **WCA)
    _announce(f"qnb_k{kernel}", p)
    g, o = _particle_parity(p, 50)
    assert rel(g.particle_velocities(), o.particle_velocities()) < 1e-4


def test_qnb_no_particles_equals_incompressible_path():
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=32)
    _announce("qnb_no_particles", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    g.step(30)
    o.step(30)
    assert rel(g.velocity(), o.velocity()) < 1e-5


# ---------------- headline configuration (smoke at full size) ----------------
def test_stokes_config3_full_size_steps():
    """128^3, 65,536 seeded uniform-random particles (no pair inside sigma), 4-point kernel, WCA:
    the first step agrees with the oracle on the fluid velocity and the positions, and a few
    more steps stay finite and divergence-free."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(128, 128, 128), np=65536, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=3, initial_positions=3, **WCA)
    _announce("stokes_config3", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    assert np.abs(g.particles() - o.particles()).max() < 1e-5
    g.step(1)
    g.sync()
    o.step(1)
    assert rel(g.velocity(), o.velocity()) < 1e-5
    assert rel(g.particles(), o.particles()) < 1e-5
    g.step(2)
    g.sync()
    assert np.isfinite(g.particles()).all()
    assert g.max_divergence() < 1e-6


# ---------------- graph replay ----------------
@pytest.mark.parametrize("scheme,np_", [(SCHEME_INCOMPRESSIBLE, 0), (SCHEME_STOKES_LIMIT, 300),
                                        (SCHEME_QUASI_NEUTRALLY_BUOYANT, 200)])
def test_graph_replay_matches_eager(scheme, np_):
    """Steps replayed from captured graphs (noise step counter read from device memory) give the
    same bits as eager launches; replays are checked one by one, with host work and syncs in
    between, and across a switch back to eager and into graphs again."""
    kw = dict(cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=61)
    if np_:
        kw.update(np=np_, kernel_points=3, initial_positions=1, **WCA)
    p = make_params(scheme, **kw)
    _announce(f"graph_replay_{scheme}", p)
    e = F.Simulation(p)
    g = F.Simulation(p)
    g.set_graph(True)
    for k in range(7):
        e.step(1)
        g.step(1)
        g.sync()
        assert np.array_equal(e.velocity(), g.velocity()), k
        if np_:
            assert np.array_equal(e.particles(), g.particles()), k
    g.set_graph(False)
    e.step(3)
    g.step(3)
    g.set_graph(True)
    e.step(4)
    g.step(4)
    assert np.array_equal(e.velocity(), g.velocity())
    if np_:
        assert np.array_equal(e.particles(), g.particles())


# ---------------- batches and schedules ----------------
def _replica_params(i, shape, np_, scheme=SCHEME_STOKES_LIMIT):
    kw = dict(cells=shape, temperature=1e-3, dt=0.05, seed=100 + i)
    if np_:
        # simple-cubic starts: no overlapping pairs, so every replica stays finite
        kw.update(np=np_, kernel_points=3, initial_positions=1, **WCA)
    return make_params(scheme, **kw)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_replicas_equal_alone(mode):
    """Replicas of uneven shapes, particle counts and schemes stepped together in one batch
    (mode 0: each replica's own graphs on its stream; mode 1: one graph for the batch step;
    mode 2: one batched dispatch per stage and kernel instance) give the same bits as each run
    alone with eager launches."""
    specs = [((16, 24, 32), 100, SCHEME_STOKES_LIMIT), ((32, 32, 32), 0, SCHEME_STOKES_LIMIT),
             ((24, 16, 48), 300, SCHEME_QUASI_NEUTRALLY_BUOYANT), ((32, 32, 32), 50, SCHEME_STOKES_LIMIT),
             ((16, 16, 16), 0, SCHEME_INCOMPRESSIBLE)]
    params = [_replica_params(i, s, n, sc) for i, (s, n, sc) in enumerate(specs)]
    for p in params:
        _announce("batch_replica", p)
    b = F.Batch(params)
    b.set_mode(mode)
    b.step(5)
    b.sync()
    b.step(7)
    b.sync()
    for p, r in zip(params, b.replicas):
        alone = F.Simulation(p)
        alone.step(12)
        assert np.array_equal(alone.velocity(), r.velocity())
        if p.np:
            assert np.array_equal(alone.particles(), r.particles())
    b.close()


def test_autotuned_schedule_changes_no_bits():
    """The per-shape schedule (z-pass column tile, W-kernel z chunk) changes timing only."""
    p = _replica_params(7, (48, 48, 24), 200)
    os.environ["FHD_AUTOTUNE"] = "0"
    fixed = F.Simulation(p)
    os.environ["FHD_AUTOTUNE"] = "1"
    tuned = F.Simulation(p)
    fixed.step(6)
    tuned.step(6)
    assert np.array_equal(fixed.velocity(), tuned.velocity())
    assert np.array_equal(fixed.particles(), tuned.particles())



@pytest.mark.parametrize("btile,zchunk", [(2, 1), (2, 2), (4, 3), (16, 16)])
def test_every_schedule_changes_no_bits(btile, zchunk):
    """Forced schedules (z-pass column tiles 2-16, W-kernel chunks of 1-16 planes, including
    ones that do not divide the grid) give the bits of the fixed default schedule, for an
    incompressible 32^3 run and a Stokes run with particles."""
    cases = [make_params(SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=5),
             _replica_params(8, (48, 48, 24), 200)]
    for p in cases:
        _announce(f"schedule_bt{btile}_zc{zchunk}", p)
        os.environ["FHD_AUTOTUNE"] = "0"
        fixed = F.Simulation(p)
        os.environ["FHD_BTILE"] = str(btile)
        os.environ["FHD_ZCHUNK"] = str(zchunk)
        try:
            forced = F.Simulation(p)
        finally:
            del os.environ["FHD_BTILE"], os.environ["FHD_ZCHUNK"]
            os.environ["FHD_AUTOTUNE"] = "1"
        fixed.step(6)
        forced.step(6)
        assert np.array_equal(fixed.velocity(), forced.velocity())
        if p.np:
            assert np.array_equal(fixed.particles(), forced.particles())


def test_blown_up_start_moves_no_coordinate_by_2p24_cells():
    """A dense uniform-random start (pairs deep inside the WCA core) asks for displacements of
    ~1e16 cells.  Such a coordinate stays where it was, so no coordinate moves by 2^24 cells or
    more in a step, everything stays finite, and particles are not stacked on one another."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=2048, kernel_points=3, temperature=1e-3, dt=0.05,
                    seed=5, initial_positions=0, **WCA)
    _announce("blown_up_start", p)
    g = F.Simulation(p)
    q = g.particles()
    flagged = False
    for _ in range(4):
        g.step(1)
        try:
            g.sync()
        except F.FhdError as e:  # the clamped displacement is reported, as fluam's errorKernel would
            assert "2^24" in str(e)
            flagged = True
        qn = g.particles()
        assert np.all(np.isfinite(qn))
        assert np.abs(qn - q).max() < 2.0 ** 24
        q = qn
    pos = np.stack([q[:2048], q[2048:4096], q[4096:]], axis=1)
    assert len(np.unique(pos, axis=0)) == 2048
    assert flagged


def test_overlap_free_random_start_matches_oracle():
    """The GPU host's random sequential addition (cell list) places every particle where the
    oracle's direct search does."""
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(32, 32, 32), np=3000, kernel_points=3,
                    temperature=1e-3, dt=0.05, seed=21, initial_positions=3, **WCA)
    _announce("overlap_free_start", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    assert np.abs(g.particles() - o.particles()).max() < 1e-6



def test_stokes_resume_is_bit_exact():
    """A Stokes run's whole state is the particle positions and the step counter: a run
    restarted at step 10 (positions read back, initialStep 10) continues with the same bits."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=300, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=71, initial_positions=3, **WCA)
    _announce("stokes_resume", p)
    full = F.Simulation(p)
    full.step(20)
    half = F.Simulation(p)
    half.step(10)
    q10 = half.particles()
    r = p.copy()
    r.initial_step = 10
    resumed = F.Simulation(r)
    resumed.set_particles(q10)
    resumed.step(10)
    assert np.array_equal(resumed.particles(), full.particles())
    assert np.array_equal(resumed.velocity(), full.velocity())


def test_incompressible_restart_from_fluid_file(tmp_path):
    """fluam's `fluid` restart: a run restarted from the saved configuration (and initialStep)
    matches the oracle restarted from the same file, and the uninterrupted run to round-off
    (the loaded field is projected once more, as every fresh run projects its start)."""
    p = make_params(SCHEME_INCOMPRESSIBLE, cells=(32, 32, 32), temperature=1e-3, dt=0.05, seed=72)
    _announce("incompressible_restart", p)
    full = F.Simulation(p)
    full.step(10)
    v10 = full.velocity()
    f = tmp_path / "r.fluidFinalConfiguration"
    F.write_fluid_file(f, (32, 32, 32), (p.lx, p.ly, p.lz), v10, densfluid=p.densfluid)
    r = p.copy()
    r.fluid = str(f).encode()
    r.initial_step = 10
    g = F.Simulation(r)
    o = O.Oracle(r)
    assert np.array_equal(g.velocity(), v10)
    g.step(10)
    o.step(10)
    full.step(10)
    assert rel(g.velocity(), o.velocity()) < 1e-5
    assert rel(g.velocity(), full.velocity()) < 1e-5
    assert g.max_divergence() < 1e-6



def _temperature_expected(vs, p):
    """fluam's temperatureFunction (src/temperature.cu:56-132) on the sampled fields."""
    N = p.mx * p.my * p.mz
    dV = p.lx * p.ly * p.lz / N
    ts = []
    for v in vs:
        v = v.reshape(3, N)
        ta = [(p.densfluid * dV * (v[a] - v[a].mean()) ** 2).sum() / N for a in range(3)]
        ts.append(ta)
    ts = np.array(ts)
    t = ts.mean(axis=1)
    return t.mean(), np.sqrt(((t - p.temperature) ** 2).mean()), ts.mean(axis=0)


def test_scheme_run_writes_fluam_outputs(tmp_path):
    """The drop-in entry point end to end: a fluam data.main through fhd_main writes the .seed,
    .data.main, .fluidFinalConfiguration, .temperature and .time files, with the same fields
    and temperature summary as the context API run of the same parameters."""
    out = tmp_path / "run"
    dm = tmp_path / "data.main"
    dm.write_text("incompressible\ndensfluid 1\nshearviscosity 1\ntemperature 0.001\ncells 16 16 16\n"
                  "celldimension 16 16 16\ninitfluid 0\nbackgroundvelocity 0 0 0\nseed 7\n"
                  f"numstepsRelaxation 0\nnumsteps 20\nsamplefreq 5\ndt 0.05\noutputname {out}\n")
    p = F.load_datamain(str(dm))
    _announce("scheme_run_incompressible", p)
    assert F.main(str(dm))
    for ext in (".seed", ".data.main", ".fluidFinalConfiguration", ".temperature", ".time"):
        assert (tmp_path / ("run" + ext)).exists(), ext
    sim = F.Simulation(p)
    vs = []
    for _ in range(4):
        sim.step(5)
        vs.append(sim.velocity())
    v_file, dens = F.read_fluid_file(str(out) + ".fluidFinalConfiguration", (16, 16, 16))
    assert dens == 1.0
    assert np.abs(v_file - vs[-1]).max() <= 1e-12 * np.abs(vs[-1]).max()
    lines = open(str(out) + ".temperature").read().split("\n")
    head = lines[0].split()
    tm, dev, n = float(head[2]), float(head[4]), int(head[5])
    t_exp, dev_exp, tabc = _temperature_expected(vs, p)
    assert n == 4
    assert abs(tm - t_exp) <= 1e-9 * t_exp
    assert abs(dev - dev_exp / p.temperature) <= 1e-9 * max(dev, 1e-30)
    for k in range(3):
        assert abs(float(lines[1 + k].split()[1]) - tabc[k]) <= 1e-9 * tabc[k]


def test_stokes_scheme_run_writes_particle_trajectory(tmp_path):
    """fhd_main on a Stokes-limit data.main: the .particles trajectory holds the positions the
    context API reaches at every sample (saveParticles, src/saveParticles.cu:55-73)."""
    out = tmp_path / "st"
    dm = tmp_path / "data.main"
    dm.write_text("stokesLimit\nnumberparticles 50\ncutoff 1.122462048309373\nsigma 1\nkernel 3\n"
                  "initialPositions 3\nshearviscosity 1\ntemperature 0.001\ncells 16 16 16\n"
                  f"celldimension 16 16 16\nseed 9\nnumsteps 6\nsamplefreq 3\ndt 0.05\noutputname {out}\n")
    p = F.load_datamain(str(dm))
    _announce("scheme_run_stokes", p)
    assert F.main(str(dm))
    rows = [ln for ln in open(str(out) + ".particles").read().split("\n")[1:] if ln.strip()]
    # per sample: the time, then one line per particle
    assert len(rows) == 2 * (1 + 50)
    sim = F.Simulation(p)
    for k in range(2):
        sim.step(3)
        q = sim.particles().reshape(3, -1).T
        got = np.array([[float(x) for x in ln.split()] for ln in rows[k * 51 + 1:(k + 1) * 51]])
        assert abs(float(rows[k * 51]) - 3 * (k + 1) * p.dt) < 1e-12
        assert np.abs(got - q).max() < 1e-12 * max(1.0, np.abs(q).max())
    assert not (tmp_path / "st.temperature").exists()  # fluam's Stokes scheme writes none


# ---------------- harmonic bonds ----------------
def _chains(tmp_path, name, nchain=8, per=8, spacing=1.1, r0=1.0, k=1.0, trap_k=0.5):
    """nchain chains of `per` particles along x, starting near the +x face of a 32-cell box so
    the chains run through the periodic boundary (unwrapped), bonded neighbour to neighbour,
    and the first particle of each chain trapped half a cell off its start."""
    pts, pairs, fixed = [], [], []
    for c in range(nchain):
        y, z = -12.0 + 6.5 * (c % 4), -9.0 + 11.0 * (c // 4)
        for m in range(per):
            i = len(pts)
            pts.append((11.3 + spacing * m, y, z))
            if m:
                pairs.append((i - 1, i, k, r0))
        fixed.append((c * per, trap_k, 0.0, (11.8, y + 0.25, z)))
    path = tmp_path / name
    lines = [str(len(pairs))] + [f"{i} {j} {kk} {rr}" for i, j, kk, rr in pairs]
    lines += [str(len(fixed))] + [f"{i} {kk} {rr} {x} {y} {z}" for i, kk, rr, (x, y, z) in fixed]
    path.write_text("\n".join(lines) + "\n")
    return np.asarray(pts).T.ravel(), str(path)


@pytest.mark.parametrize("scheme", [SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT])
def test_bonded_chains_parity(tmp_path, scheme):
    """Bonded chains (springs with rest length, fixed-point traps, bonds across the periodic
    boundary) with WCA: forces and 20-step trajectories match the oracle."""
    q, path = _chains(tmp_path, "bonds.dat")
    p = make_params(scheme, cells=(32, 32, 32), np=64, kernel_points=4, temperature=1e-3, dt=0.05, seed=51,
                    initial_positions=0, bonded_forces=path, **WCA)
    _announce(f"bonded_chains_s{scheme}", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    g.set_particles(q)
    o.set_particles(q)
    g.step(1)
    o.step(1)
    fg, fo = g.forces(), o.forces()
    assert np.abs(fo).max() > 0.05
    assert rel(fg, fo) < 1e-4
    g.step(19)
    g.sync()
    o.step(19)
    qg, qo = g.particles(), o.particles()
    assert np.isfinite(qg).all()
    assert rel(qg, qo) < 1e-5
    assert rel(qg - q, qo - q) < 1e-3
    assert g.max_divergence() < 1e-6


def test_bonds_without_wca_and_bad_file(tmp_path):
    """Bond forces alone (computeNonBondedForces 0) match the oracle; a bond naming a particle
    outside 0..np-1 is an error at creation."""
    q, path = _chains(tmp_path, "bonds.dat", nchain=4, per=6)
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=24, kernel_points=3, nonbonded=0,
                    temperature=1e-3, dt=0.05, seed=52, initial_positions=0, bonded_forces=path, **WCA)
    _announce("bonded_no_wca", p)
    g = F.Simulation(p)
    o = O.Oracle(p)
    g.set_particles(q)
    o.set_particles(q)
    g.step(10)
    g.sync()
    o.step(10)
    assert rel(g.forces(), o.forces()) < 1e-4
    assert rel(g.particles(), o.particles()) < 1e-5
    bad = tmp_path / "bad.dat"
    bad.write_text("1\n0 24 1.0 0.0\n0\n")
    pb = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=24, kernel_points=3, initial_positions=1,
                     bonded_forces=str(bad), **WCA)
    with pytest.raises(F.FhdError, match="bond index"):
        F.Simulation(pb)


def test_bonded_scheme_run_from_datamain(tmp_path):
    """bondedForces in a fluam data.main, run through fhd_main: a trapped pair, a spring chain
    and extra mobility; the .particles trajectory equals the context API's on the same
    parameters, and the trap pulls its particle toward the fixed point."""
    bonds = tmp_path / "springs.dat"
    bonds.write_text("3\n0 1 2.0 1.0\n1 2 2.0 1.0\n2 3 2.0 0.0\n1\n4 0.5 0.0 3.0 3.0 3.0\n")
    out = tmp_path / "bd"
    dm = tmp_path / "data.main"
    dm.write_text("stokesLimit\nnumberparticles 20\ncutoff 1.122462048309373\nsigma 1\nkernel 4\n"
                  "initialPositions 1\nshearviscosity 1\ntemperature 0.0001\ncells 16 16 16\n"
                  f"celldimension 16 16 16\nseed 12\nnumsteps 40\nsamplefreq 20\ndt 0.05\nextraMobility 0.2\n"
                  f"bondedForces {bonds}\noutputname {out}\n")
    p = F.load_datamain(str(dm))
    assert p.bonded_forces == str(bonds).encode()
    _announce("scheme_run_bonded", p)
    assert F.main(str(dm))
    rows = [ln for ln in open(str(out) + ".particles").read().split("\n")[1:] if ln.strip()]
    assert len(rows) == 2 * (1 + 20)
    sim = F.Simulation(p)
    o = O.Oracle(p)
    q0 = sim.particles().reshape(3, -1).T.copy()
    sim.step(40)
    o.step(40)
    q = sim.particles().reshape(3, -1).T
    got = np.array([[float(x) for x in ln.split()] for ln in rows[22:]])
    assert np.abs(got - q).max() < 1e-12 * max(1.0, np.abs(q).max())
    assert rel(sim.particles(), o.particles()) < 1e-5
    x0 = np.array([3.0, 3.0, 3.0])
    assert np.linalg.norm(q[4] - x0) < np.linalg.norm(q0[4] - x0)


@pytest.mark.parametrize("text", ["2\n0 1 1.0 0.0\n", "1\n0 1 1.0 0.0\n1\n0 1.0 0.0 1.0 2.0\n", "-1\n0\n", ""])
def test_malformed_bond_file_is_an_error(tmp_path, text):
    """A truncated bond file or a negative count is reported at creation, nothing is run."""
    path = tmp_path / "bonds.dat"
    path.write_text(text)
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=2, kernel_points=3, initial_positions=1,
                    bonded_forces=str(path), **WCA)
    with pytest.raises(F.FhdError, match="bonds file"):
        F.Simulation(p)


@pytest.mark.parametrize("mode", [0, 2])
def test_batch_with_bonds_equals_alone(tmp_path, mode):
    """Replicas with and without bonds in one batch give the same bits as each run alone."""
    qa, pa = _chains(tmp_path, "a.dat", nchain=4, per=8)
    qb, pb = _chains(tmp_path, "b.dat", nchain=8, per=4, r0=0.9, k=2.0)
    params = [make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=32, kernel_points=4, temperature=1e-3,
                          dt=0.05, seed=61, initial_positions=0, bonded_forces=pa, **WCA),
              make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(32, 32, 32), np=32, kernel_points=3,
                          temperature=1e-3, dt=0.05, seed=62, initial_positions=0, bonded_forces=pb, **WCA),
              _replica_params(3, (16, 24, 32), 100)]
    qs = [qa, qb, None]
    for p in params:
        _announce("batch_bonded_replica", p)
    b = F.Batch(params)
    for r, q in zip(b.replicas, qs):
        if q is not None:
            r.set_particles(q)
    b.set_mode(mode)
    b.step(6)
    b.sync()
    for p, r, q in zip(params, b.replicas, qs):
        alone = F.Simulation(p)
        if q is not None:
            alone.set_particles(q)
        alone.step(6)
        assert np.array_equal(alone.velocity(), r.velocity())
        assert np.array_equal(alone.particles(), r.particles())
    b.close()


def test_noise_overlap_switch_changes_no_bits():
    """The Stokes noise built on the context's own stream instead of the second stream, switched
    mid-run and under graph replay, gives the same bits as the default."""
    p = make_params(SCHEME_STOKES_LIMIT, cells=(32, 32, 32), np=400, kernel_points=4, temperature=1e-3,
                    dt=0.05, seed=71, initial_positions=1, **WCA)
    _announce("noise_overlap_switch", p)
    a = F.Simulation(p)
    b = F.Simulation(p)
    b.set_graph(True)
    a.step(3)
    b.step(3)
    b.set_noise_overlap(False)
    a.step(4)
    b.step(4)
    b.set_noise_overlap(True)
    a.step(2)
    b.step(2)
    assert np.array_equal(a.particles(), b.particles())
    assert np.array_equal(a.velocity(), b.velocity())


# ---------------- equal-time structure factor (in situ) ----------------
def _sk_numpy(v, p):
    """rho dV / kT |fftn(v_a)|^2 / N over the half spectrum, layout [a][z][ky][kx]."""
    m = (p.mx, p.my, p.mz)
    N = p.mx * p.my * p.mz
    dV = (p.lx / p.mx) * (p.ly / p.my) * (p.lz / p.mz)
    v = v.reshape(3, p.mz, p.my, p.mx)
    return np.stack([np.abs(np.fft.rfftn(v[a])) ** 2 for a in range(3)]) * (p.densfluid * dV / p.temperature) / N


@pytest.mark.parametrize("scheme,np_", [(SCHEME_INCOMPRESSIBLE, 0), (SCHEME_STOKES_LIMIT, 200)])
def test_structure_factor_in_situ_matches_numpy(scheme, np_):
    """The GPU accumulator over 5 snapshots equals numpy's transform of the same snapshots
    (FP32 transform, FP64 accumulation: 1e-5 relative), on a non-cubic box with dV != 1."""
    kw = dict(np=np_, kernel_points=4, initial_positions=1, **WCA) if np_ else {}
    p = make_params(scheme, cells=(16, 24, 32), box=(8.0, 18.0, 40.0), temperature=2e-3, dt=0.05, seed=81,
                    densfluid=1.5, **kw)
    _announce(f"sk_in_situ_s{scheme}", p)
    g = F.Simulation(p)
    g.step(10)
    ref = 0.0
    g.sk_reset()
    for _ in range(5):
        g.step(3)
        g.sk_sample()
        ref = ref + _sk_numpy(g.velocity(), p)
    s, n = g.structure_factor()
    assert n == 5
    assert rel(s, ref / 5) < 1e-5
    g.sk_reset()
    s0, n0 = g.structure_factor()
    assert n0 == 0 and not s0.any()


def test_structure_factor_in_situ_equilibrium():
    """At equilibrium S_a(k) is the diagonal of the discrete projector, P_aa = 1 - g_a^2/|g|^2
    with g_a = (2/dx_a) sin(pi n_a/m_a): the in-situ accumulator over four replicas and 120
    snapshots each gives sum_a S_a = 2 to 1% (mode average), and per component the mean of S_a
    over the modes with P_aa < 1/4 and over those with P_aa > 3/4 within 0.03 of the mean P_aa
    there (this separates the components: a swapped axis fails it)."""
    m = (32, 16, 24)
    acc, ns = 0.0, 0
    for seed in (91, 92, 93, 94):
        p = make_params(SCHEME_INCOMPRESSIBLE, cells=m, box=(32.0, 24.0, 24.0), temperature=1e-3, dt=0.5,
                        seed=seed)
        _announce(f"sk_equilibrium_seed{seed}", p)
        g = F.Simulation(p)
        g.step(100)
        for _ in range(120):
            g.step(5)
            g.sk_sample()
        s, n = g.structure_factor()
        acc = acc + s * n
        ns += n
    S = acc / ns
    nx = np.arange(m[0] // 2 + 1)[None, None, :]
    ny = np.fft.fftfreq(m[1], 1.0 / m[1])[None, :, None]
    nz = np.fft.fftfreq(m[2], 1.0 / m[2])[:, None, None]
    h = (32.0 / m[0], 24.0 / m[1], 24.0 / m[2])
    g2 = [(2 / h[0] * np.sin(np.pi * nx / m[0])) ** 2 + 0 * ny + 0 * nz,
          (2 / h[1] * np.sin(np.pi * ny / m[1])) ** 2 + 0 * nx + 0 * nz,
          (2 / h[2] * np.sin(np.pi * nz / m[2])) ** 2 + 0 * nx + 0 * ny]
    G = g2[0] + g2[1] + g2[2]
    nonzero = G > 0
    tr = S.sum(axis=0)[nonzero]
    assert abs(tr.mean() - 2.0) < 0.02
    for a in range(3):
        P = 1 - g2[a][nonzero] / G[nonzero]
        Sa = S[a][nonzero]
        for sel in (P < 0.25, P > 0.75):
            assert sel.sum() > 100
            assert abs(Sa[sel].mean() - P[sel].mean()) < 0.03, (a, Sa[sel].mean(), P[sel].mean())


def test_scheme_run_writes_structure_factor(tmp_path):
    """structureFactor 1 in data.main: the run samples every samplefreq steps and writes
    <outputname>.structureFactor, equal to the context API's accumulator on the same run."""
    out = tmp_path / "run"
    dm = tmp_path / "data.main"
    dm.write_text("incompressible\ndensfluid 1\nshearviscosity 1\ntemperature 0.001\ncells 16 16 8\n"
                  "celldimension 16 16 8\nseed 9\nnumstepsRelaxation 0\nnumsteps 20\nsamplefreq 5\ndt 0.05\n"
                  f"structureFactor 1\noutputname {out}\n")
    p = F.load_datamain(str(dm))
    assert p.structure_factor == 1
    _announce("scheme_run_structure_factor", p)
    assert F.main(str(dm))
    lines = open(str(out) + ".structureFactor").read().split("\n")
    assert lines[0].split()[-1] == "4"
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[2:] if ln.strip()])
    assert rows.shape == (8 * 16 * 9, 9)
    sim = F.Simulation(p)
    for _ in range(4):
        sim.step(5)
        sim.sk_sample()
    s, n = sim.structure_factor()
    assert n == 4
    np.testing.assert_allclose(rows[:, 6:9], s.reshape(3, -1).T, rtol=1e-8, atol=1e-12)
    i = 1 + 9 * 1 + 9 * 16 * 7  # nx 1, ny 1, nz -1 (row z = 7)
    assert rows[i, :3].tolist() == [1, 1, -1]
    np.testing.assert_allclose(rows[i, 3:6], [2 * np.pi / 16, 2 * np.pi / 16, -2 * np.pi / 8])
