"""Pins the oracle's harmonic bonds (src/forceBondedGPU.cu:21-121, file layout
bin/inputfiles/README:55-108) with checks that do not restate its formula:

  * closed forms: a particle in a trap (r0 = 0) feels -k (q - x0); a stretched spring pulls
    with magnitude k |r - r0| along the bond;
  * F = -grad V, by central differences of V = sum_b k_b/2 (|r_b| - r_b0)^2 over a random
    network of particle-particle and fixed-point bonds;
  * Newton's third law for particle-particle bonds, linearity in the file's bonds
    (a bond listed twice doubles its force);
  * the separation is the plain difference of the unwrapped positions, with no periodic
    image (fluam reads the raw coordinates);
  * the special cases r = 0: no force for r0 != 0, and -k r = 0 for r0 = 0.
"""
import numpy as np
import pytest

import oracle_lib as O
from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT


def _oracle(tmp_path, np_, pairs=(), fixed=(), box=16.0):
    path = tmp_path / "bonds.dat"
    lines = [str(len(pairs))] + [f"{i} {j} {k!r} {r0!r}" for i, j, k, r0 in pairs]
    lines += [str(len(fixed))] + [f"{i} {k!r} {r0!r} {x!r} {y!r} {z!r}" for i, k, r0, (x, y, z) in fixed]
    path.write_text("\n".join(lines) + "\n")
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), box=(box, box, box), np=np_, nonbonded=0,
                    initial_positions=0, seed=3, bonded_forces=str(path))
    return O.Oracle(p)


def _q(points):
    return np.asarray(points, dtype=float).T.ravel()


def _energy(q, np_, pairs, fixed):
    Q = q.reshape(3, np_)
    e = 0.0
    for i, j, k, r0 in pairs:
        e += 0.5 * k * (np.linalg.norm(Q[:, i] - Q[:, j]) - r0) ** 2
    for i, k, r0, x in fixed:
        e += 0.5 * k * (np.linalg.norm(Q[:, i] - np.asarray(x)) - r0) ** 2
    return e


def test_trap_at_fixed_point_is_hookean(tmp_path):
    x0 = (1.5, -2.0, 0.25)
    o = _oracle(tmp_path, 2, fixed=[(0, 2.0, 0.0, x0), (1, 0.5, 0.0, (0.0, 0.0, 0.0))])
    pts = [(3.0, 1.0, -4.0), (-7.0, 6.5, 2.0)]
    f = o.bonded(_q(pts)).reshape(3, 2)
    np.testing.assert_allclose(f[:, 0], -2.0 * (np.array(pts[0]) - x0), rtol=1e-15)
    np.testing.assert_allclose(f[:, 1], -0.5 * np.array(pts[1]), rtol=1e-15)


def test_stretched_spring_magnitude_and_direction(tmp_path):
    o = _oracle(tmp_path, 2, pairs=[(0, 1, 3.0, 2.0)])
    a, b = np.array([0.5, 0.0, 0.0]), np.array([0.5, 0.0, 3.5])  # r = 3.5, stretched by 1.5
    f = o.bonded(_q([a, b])).reshape(3, 2)
    np.testing.assert_allclose(f[:, 0], [0.0, 0.0, 3.0 * 1.5], atol=1e-14)
    np.testing.assert_allclose(f[:, 1], [0.0, 0.0, -3.0 * 1.5], atol=1e-14)
    b2 = np.array([0.5, 1.0, 0.0])  # r = 1, compressed by 1: pushed apart
    f = o.bonded(_q([a, b2])).reshape(3, 2)
    np.testing.assert_allclose(f[:, 0], [0.0, -3.0, 0.0], atol=1e-14)


def test_force_is_minus_energy_gradient_random_network(tmp_path):
    rng = np.random.default_rng(11)
    n = 12
    pairs = [(int(i), int(j), float(rng.uniform(0.5, 3)), float(rng.choice([0.0, rng.uniform(0.5, 2)])))
             for i, j in rng.integers(0, n, (20, 2)) if i != j]
    fixed = [(int(rng.integers(0, n)), float(rng.uniform(0.5, 3)), float(rng.choice([0.0, 1.0])),
              tuple(float(x) for x in rng.uniform(-8, 8, 3))) for _ in range(6)]
    o = _oracle(tmp_path, n, pairs, fixed)
    q = rng.uniform(-7, 7, 3 * n)
    f = o.bonded(q)
    h = 1e-6
    g = np.zeros_like(q)
    for c in range(q.size):
        e = np.zeros_like(q)
        e[c] = h
        g[c] = (_energy(q + e, n, pairs, fixed) - _energy(q - e, n, pairs, fixed)) / (2 * h)
    np.testing.assert_allclose(f, -g, rtol=1e-6, atol=1e-6)


def test_pair_bonds_obey_third_law_and_add_linearly(tmp_path):
    rng = np.random.default_rng(12)
    q = rng.uniform(-5, 5, 3 * 6)
    pairs = [(0, 1, 1.0, 0.7), (1, 2, 2.0, 0.0), (3, 4, 0.5, 1.3), (5, 0, 1.5, 2.2)]
    once = _oracle(tmp_path, 6, pairs).bonded(q).reshape(3, 6)
    np.testing.assert_allclose(once.sum(axis=1), 0.0, atol=1e-12)
    twice = _oracle(tmp_path, 6, pairs + pairs).bonded(q).reshape(3, 6)
    np.testing.assert_allclose(twice, 2 * once, rtol=1e-14, atol=1e-14)


def test_bond_uses_unwrapped_separation(tmp_path):
    # box 16: the particles sit 14 apart along x, 2 apart through the periodic boundary; the
    # spring sees 14 (fluam uses the raw coordinates, no minimum image)
    o = _oracle(tmp_path, 2, pairs=[(0, 1, 1.0, 0.0)])
    f = o.bonded(_q([(-7.0, 0.0, 0.0), (7.0, 0.0, 0.0)])).reshape(3, 2)
    np.testing.assert_allclose(f[0], [14.0, -14.0], rtol=1e-15)


def test_coincident_particles(tmp_path):
    o = _oracle(tmp_path, 4, pairs=[(0, 1, 1.0, 1.0), (2, 3, 1.0, 0.0)])
    f = o.bonded(_q([(1.0, 2.0, 3.0)] * 4))
    assert np.array_equal(f, np.zeros(12))


def test_bad_bond_file_is_rejected(tmp_path):
    with pytest.raises(RuntimeError):
        _oracle(tmp_path, 2, pairs=[(0, 2, 1.0, 0.0)])  # index outside 0..np-1


@pytest.mark.parametrize("text", ["2\n0 1 1.0 0.0\n", "1\n0 1 1.0 0.0\n1\n0 1.0 0.0 1.0 2.0\n", "-1\n0\n", ""])
def test_truncated_or_malformed_bond_file_is_rejected(tmp_path, text):
    path = tmp_path / "bonds.dat"
    path.write_text(text)
    p = make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16), np=2, nonbonded=0, initial_positions=0, seed=3,
                    bonded_forces=str(path))
    with pytest.raises(RuntimeError):
        O.Oracle(p)
