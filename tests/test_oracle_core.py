"""Pins the FP64 oracle's building blocks against known answers, closed forms and brute force.

Covers: Philox4x32-10 (Random123 known-answer vectors), Box-Muller normality, the 3D FFT
(numpy.fft golden fixtures + brute-force DFT), the Peskin 3-/4-point kernels (closed-form
values, moment conditions, derivative), WCA (zero force at 2^(1/6) sigma, Newton's third law,
brute-force pair sum) and spreading/interpolation (brute force over every grid node,
adjointness)."""
import os

import numpy as np
import pytest

import oracle_lib as O
from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")


# ---------------- RNG ----------------
# Random123 kat_vectors for philox4x32 with 10 rounds: (ctr, key) -> out.
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(O.philox(ctr, key)) == expect


def test_normals_are_standard_normal():
    z = np.concatenate([O.normals4(7, 3, 1, 0, i) for i in range(20000)])
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.02
    # fourth moment of a standard normal is 3
    assert abs((z ** 4).mean() - 3.0) < 0.15
    # distinct (step, stream, sub, seed) give distinct draws
    a = O.normals4(7, 3, 1, 0, 5)
    for other in [O.normals4(7, 4, 1, 0, 5), O.normals4(7, 3, 2, 0, 5), O.normals4(7, 3, 1, 1, 5),
                  O.normals4(8, 3, 1, 0, 5)]:
        assert not np.allclose(a, other)


def test_fluid_noise_layout_matches_draws():
    N = 10
    Z1 = O.fluid_noise(11, 2, 0, N).reshape(6, N)
    Z2 = O.fluid_noise(11, 2, 1, N).reshape(6, N)
    for i in range(N):
        n = np.concatenate([O.normals4(11, 2, 1, sub, i) for sub in range(3)])
        np.testing.assert_array_equal(Z1[:, i], n[:6])
        np.testing.assert_array_equal(Z2[:, i], n[6:])


# ---------------- FFT ----------------
def test_fft_against_numpy_golden():
    g = np.load(GOLDEN)
    shapes = g["shapes"]
    for k in range(len(shapes)):
        a = g[f"in_{k}"]
        f = O.fft3d(a, -1)
        np.testing.assert_allclose(f, g[f"fwd_{k}"], rtol=0, atol=1e-10 * np.abs(g[f"fwd_{k}"]).max())
        b = O.fft3d(a, +1)
        np.testing.assert_allclose(b, g[f"inv_{k}"], rtol=0, atol=1e-10 * np.abs(g[f"inv_{k}"]).max())


@pytest.mark.parametrize("shape", [(1, 1, 7), (2, 3, 4), (3, 1, 6)])
def test_fft_brute_force_dft(shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    mz, my, mx = shape
    ref = np.zeros(shape, dtype=complex)
    for kz in range(mz):
        for ky in range(my):
            for kx in range(mx):
                s = 0
                for z in range(mz):
                    for y in range(my):
                        for x in range(mx):
                            s += a[z, y, x] * np.exp(-2j * np.pi * (kx * x / mx + ky * y / my + kz * z / mz))
                ref[kz, ky, kx] = s
    np.testing.assert_allclose(O.fft3d(a, -1), ref, atol=1e-11)


# ---------------- Peskin kernels ----------------
def test_delta3_closed_form_values():
    assert O.delta3(0.0) == pytest.approx(2.0 / 3.0, abs=1e-15)
    assert O.delta3(0.5) == pytest.approx(0.5, abs=1e-15)
    assert O.delta3(0.5 - 1e-12) == pytest.approx(0.5, abs=1e-9)  # both branches agree
    assert O.delta3(1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert O.delta3(1.5) == pytest.approx(0.0, abs=1e-15)
    assert O.delta3(-0.3) == O.delta3(0.3)
    assert O.delta3(1.6) == 0.0


@pytest.mark.parametrize("x", np.linspace(0.0, 0.999, 23))
def test_delta3_moments(x):
    js = np.arange(-3, 4)
    phi = np.array([O.delta3(x - j) for j in js])
    assert phi.sum() == pytest.approx(1.0, abs=1e-14)
    assert ((x - js) * phi).sum() == pytest.approx(0.0, abs=1e-14)
    assert (phi ** 2).sum() == pytest.approx(0.5, abs=1e-14)


@pytest.mark.parametrize("x", np.linspace(0.0, 0.999, 23))
def test_delta4_moments(x):
    js = np.arange(-3, 5)
    phi = np.array([O.delta4(x - j) for j in js])
    assert phi.sum() == pytest.approx(1.0, abs=1e-14)
    assert phi[js % 2 == 0].sum() == pytest.approx(0.5, abs=1e-14)
    assert ((x - js) * phi).sum() == pytest.approx(0.0, abs=1e-14)
    assert (phi ** 2).sum() == pytest.approx(3.0 / 8.0, abs=1e-14)


def test_delta4_matches_fluam_1d_weights():
    # src/deltaGPU.cu:44-53 weights for offset x in [0,1): nodes at distance 1+x, x, 1-x, 2-x
    for x in np.linspace(0, 0.99, 12):
        s = 0.125 * (3 - 2 * x + np.sqrt(1 + 4 * x * (1 - x)))
        w = [0.25 * (3 - 2 * x) - s, s, 0.25 * (2 * x - 1) + s, 0.5 - s]
        got = [O.delta4(x + 1), O.delta4(x), O.delta4(x - 1), O.delta4(x - 2)]
        np.testing.assert_allclose(got, w, atol=1e-14)


@pytest.mark.parametrize("r", [-1.3, -0.7, -0.2, 0.1, 0.45, 0.8, 1.2, 1.45])
def test_ddelta3_is_derivative(r):
    h = 1e-6
    fd = (O.delta3(r + h) - O.delta3(r - h)) / (2 * h)
    assert O.ddelta3(r) == pytest.approx(fd, rel=1e-6, abs=1e-8)


# ---------------- particles ----------------
def _stokes_params(n=(8, 10, 12), np_=40, kernel=3, box=None, **kw):
    return make_params(SCHEME_STOKES_LIMIT, cells=n, box=box, np=np_, kernel_points=kernel,
                       cutoff=2 ** (1 / 6), wca_sigma=1.0, **kw)


def test_wca_zero_at_minimum_and_third_law():
    p = _stokes_params(np_=2)
    o = O.Oracle(p)
    rmin = 2 ** (1 / 6)
    q = np.array([0.0, rmin * 0.999999999, 0, 0, 0, 0])
    f = o.wca(q).reshape(3, 2)
    assert abs(f[0, 0]) < 1e-6 * 48e-3
    # close pair: repulsive, equal and opposite
    q = np.array([0.1, 0.95, 0.2, 0.2, -0.3, -0.3])
    f = o.wca(q).reshape(3, 2)
    np.testing.assert_allclose(f[:, 0], -f[:, 1], atol=1e-15)
    assert f[0, 0] < 0  # particle 0 is at smaller x -> pushed to -x


def test_wca_against_brute_force_pairs():
    p = _stokes_params(n=(10, 10, 10), np_=120)
    o = O.Oracle(p)
    rng = np.random.default_rng(3)
    q = (rng.random((3, 120)) - 0.5) * 10.0 + 10.0 * rng.integers(-2, 3, (3, 120))  # unwrapped
    F = o.wca(q.ravel()).reshape(3, 120)
    L = 10.0
    sig, eps, rc = 1.0, p.wca_epsilon, 2 ** (1 / 6)
    ref = np.zeros((3, 120))
    for i in range(120):
        for j in range(120):
            if i == j:
                continue
            r = q[:, i] - q[:, j]
            r -= L * np.round(r / L)
            r2 = r @ r
            if r2 < rc * rc:
                ref[:, i] += 48 * eps * ((sig ** 2 / r2) ** 6 - 0.5 * (sig ** 2 / r2) ** 3) / r2 * r
    np.testing.assert_allclose(F, ref, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(F.sum(axis=1), 0, atol=1e-9 * np.abs(F).max())


def _node_positions(p, comp):
    """Coordinates of every node of component comp: face-centred along comp, cell-centred
    along the other axes; box centred at the origin."""
    m = (p.mx, p.my, p.mz)
    L = (p.lx, p.ly, p.lz)
    h = [L[a] / m[a] for a in range(3)]
    ax = []
    for a in range(3):
        off = 1.0 if a == comp else 0.5
        ax.append((np.arange(m[a]) + off) * h[a] - 0.5 * L[a])
    return ax, h, L


@pytest.mark.parametrize("kernel", [3, 4])
def test_spread_and_interp_against_all_nodes(kernel):
    p = _stokes_params(n=(6, 7, 8), box=(6.0, 8.4, 8.0), np_=5, kernel=kernel)
    o = O.Oracle(p)
    rng = np.random.default_rng(9)
    q = (rng.random((3, 5)) - 0.5) * np.array([[6.0], [8.4], [8.0]]) + np.array([[6.0], [0], [-16.0]])
    F = rng.standard_normal((3, 5))
    S = o.spread(q.ravel(), F.ravel()).reshape(3, p.mz, p.my, p.mx)
    v = rng.standard_normal(3 * p.mx * p.my * p.mz)
    U = o.interp(q.ravel(), v).reshape(3, 5)
    V = v.reshape(3, p.mz, p.my, p.mx)
    phi = O.delta3 if kernel == 3 else O.delta4
    dV = (p.lx / p.mx) * (p.ly / p.my) * (p.lz / p.mz)
    for comp in range(3):
        ax, h, L = _node_positions(p, comp)
        ref = np.zeros((p.mz, p.my, p.mx))
        uref = np.zeros(5)
        for n in range(5):
            w = []
            for a in range(3):
                d = q[a, n] - ax[a]
                d -= L[a] * np.round(d / L[a])
                w.append(np.array([phi(x / h[a]) for x in d]))
            W3 = w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
            ref += W3 * F[comp, n] / dV
            uref[n] = (W3 * V[comp]).sum()
        np.testing.assert_allclose(S[comp], ref, atol=1e-13)
        np.testing.assert_allclose(U[comp], uref, atol=1e-13)
    # adjointness <J v, F> = dV <v, S F>
    assert (U * F).sum() == pytest.approx(dV * (v * S.ravel()).sum(), rel=1e-12)


def test_initial_positions_uniform_and_cubic():
    p = _stokes_params(n=(16, 16, 16), np_=4000)
    q = O.initial_positions(p).reshape(3, -1)
    assert q.min() >= -8.0 and q.max() < 8.0
    assert abs(q.mean()) < 0.3
    p.initial_positions = 1
    p.np = 27
    q = O.initial_positions(p).reshape(3, -1)
    assert sorted(set(np.round(q[0], 9))) == pytest.approx([-16 / 3, 0.0, 16 / 3])



@pytest.mark.parametrize("scheme", [SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT])
def test_blown_up_step_leaves_the_coordinate(scheme):
    """A dense uniform-random start puts pairs deep inside the WCA core, and the capped pair
    force then asks for displacements of ~1e16 cells.  No coordinate may move by 2^24 cells or
    more in one step (step_disp keeps it where it was), and everything stays finite."""
    p = make_params(scheme, cells=(8, 8, 8), np=512, kernel_points=3, temperature=1e-3, dt=0.05, seed=7,
                    initial_positions=0, cutoff=2 ** (1 / 6), wca_sigma=1.0)
    o = O.Oracle(p)
    q0 = o.particles()
    o.step(1)
    q1 = o.particles()
    assert np.all(np.isfinite(q1))
    assert np.abs(q1 - q0).max() < 2.0 ** 24
    # the start is blown up: the Stokes half step dt/2 J L^-1 S F asks for >= 2^24 cells somewhere
    u = o.interp(q0, o.solve_stokes(o.spread(q0, o.wca(q0))))
    assert np.abs(0.5 * p.dt * u).max() > 2.0 ** 24


def test_overlap_free_random_start():
    """initialPositions 3: the first draw of every particle is the uniform start's draw, and a
    particle is redrawn only while it lies closer than sigma (minimum image) to an earlier one."""
    p0 = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=600, kernel_points=3, temperature=1e-3,
                     dt=0.05, seed=9, initial_positions=0, cutoff=2 ** (1 / 6), wca_sigma=1.0)
    p3 = p0.copy()
    p3.initial_positions = 3
    q0 = O.Oracle(p0).particles().reshape(3, -1).T
    q3 = O.Oracle(p3).particles().reshape(3, -1).T
    L = np.array([p0.lx, p0.ly, p0.lz])

    def dmin_to_earlier(q, i):
        d = q[:i] - q[i]
        d -= L * np.floor(d / L + 0.5)
        return np.sqrt((d * d).sum(axis=1)).min() if i else np.inf

    moved = 0
    for i in range(len(q3)):
        assert dmin_to_earlier(q3, i) >= 1.0
        if not np.array_equal(q3[i], q0[i]):
            moved += 1
            # redrawn only because the first draw overlapped an earlier accepted particle
            d = q3[:i] - q0[i]
            d -= L * np.floor(d / L + 0.5)
            assert np.sqrt((d * d).sum(axis=1)).min() < 1.0
    assert moved > 0  # at this density some first draws overlap
