"""Pins of the Stokes step's stochastic stress, drawn in Fourier space (the oracle's
kspace_sigma / kspace_stress_div; the GPU's z pass draws the same modes from the same counters).

The reference draws the stress Sigma cell by cell in real space (white N(0,1) per component,
src/stokesLimitFunctions.cu:2281-2378,3922-4069) and applies the staggered divergence
D.Sigma (src/kernelConstructW.cu:148-198).  Drawing the transform of a white field instead gives
a field with the same law, and the divergence becomes a Fourier symbol.  Pinned here:

  * Hermitian symmetry: the inverse transform of Sigma_hat is real;
  * the real fields are white N(0,1): mean, variance, lag-1 and cross-component correlations;
    Sigma_1 and Sigma_2 are independent; per-mode second moments E|S(k)|^2 = N on generic and
    self-conjugate modes, and E[S(k)^2] = 0 on generic ones (real and imaginary parts balanced);
  * the divergence symbol equals the pinned real-space noise divergence applied to the real
    fields (non-trace-free stress with prefactors f1, f4), on a box with unequal cell sizes."""
import numpy as np
import pytest

import oracle_lib as O
from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT


def _oracle(cells, box=None, seed=5):
    return O.Oracle(make_params(SCHEME_STOKES_LIMIT, cells=cells, box=box, seed=seed))


@pytest.mark.parametrize("cells", [(12, 8, 6), (8, 9, 6), (6, 5, 3)])
def test_stress_is_real_after_inverse_transform(cells):
    o = _oracle(cells)
    for nset in (1, 2):
        sig, _, mi = O.kspace_stress(o, 3, nset, 1.0, 1.0)
        assert mi < 1e-13 * max(1.0, np.abs(sig).max())


def test_stress_fields_are_white_unit_normals():
    o = _oracle((24, 16, 12))
    N = o.N
    s1, _, _ = O.kspace_stress(o, 7, 1, 1.0, 1.0)
    s12, _, _ = O.kspace_stress(o, 7, 2, 1.0, 1.0)
    s2 = s12 - s1
    n = 6 * N
    for s in (s1, s2):
        assert abs(s.mean()) < 4.0 / np.sqrt(n)
        assert abs(s.var() - 1.0) < 4.0 * np.sqrt(2.0 / n)
        f = s.reshape(6, 12, 16, 24)
        for ax in (1, 2, 3):  # lag-1 correlation along z, y, x of every component
            r = (f * np.roll(f, 1, axis=ax)).mean()
            assert abs(r) < 4.0 / np.sqrt(n), ax
        c = np.corrcoef(f.reshape(6, -1))
        assert np.abs(c - np.eye(6)).max() < 4.5 / np.sqrt(N)
    assert abs((s1 * s2).mean()) < 4.0 / np.sqrt(n)  # the two stress sets are independent


def test_stress_modes_have_the_moments_of_a_white_field():
    """Over 300 steps on a 6 x 4 x 4 grid: E|S(k)|^2 = N at every mode, self-conjugate ones
    included (their values are real N(0, N)), and E[S(k)^2] = 0 at generic modes."""
    cells = (6, 4, 4)
    o = _oracle(cells, seed=17)
    N = o.N
    acc2 = np.zeros((6, N))
    accsq = np.zeros((6, N), dtype=complex)
    nstep = 300
    for st in range(nstep):
        sig, _, _ = O.kspace_stress(o, st, 1, 1.0, 1.0)
        S = np.fft.fftn(sig.reshape(6, 4, 4, 6), axes=(1, 2, 3)).reshape(6, N)
        acc2 += np.abs(S) ** 2
        accsq += S * S
    m2 = acc2 / (nstep * N)
    kz, ky, kx = np.meshgrid(np.arange(4), np.arange(4), np.arange(6), indexing="ij")
    selfc = (((2 * kx) % 6 == 0) & ((2 * ky) % 4 == 0) & ((2 * kz) % 4 == 0)).ravel()
    assert selfc.sum() == 8
    # self-conjugate modes: 6 components x 8 modes x 300 draws of a real N(0, N) squared
    assert abs(m2[:, selfc].mean() - 1.0) < 4.0 * np.sqrt(2.0 / (6 * 8 * nstep))
    assert abs(m2[:, ~selfc].mean() - 1.0) < 4.0 * np.sqrt(1.0 / (6 * (N - 8) * nstep))
    r = np.abs(accsq[:, ~selfc].mean()) / (nstep * N)
    assert r < 4.0 * np.sqrt(1.0 / (6 * (N - 8) * nstep))


@pytest.mark.parametrize("nset,f1,f4", [(1, 1.0, 1.0), (2, 0.7, 0.3)])
def test_stress_divergence_symbol_equals_the_real_space_divergence(nset, f1, f4):
    o = _oracle((12, 10, 8), box=(12.0, 7.5, 10.0), seed=9)
    sig, w, _ = O.kspace_stress(o, 2, nset, f1, f4)
    ref = O.noise_divergence(o, sig, f1, f4, False)
    assert np.abs(w - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(ref).max() > 0.1


# ---- the projected noise source, drawn directly (orc_kspace_noise) ----
# The Stokes solve keeps only P_e D.Sigma of the noise source (P_e: the solve's projection in
# the frame of the staggering shifts e_a).  With the step's prefactors (f1^2 = 2 f4^2) that part
# has per-mode covariance N f4^2 |g|^2 E P E*, which two complex normals per mode reproduce.

@pytest.mark.parametrize("cells,box", [((6, 4, 4), None), ((8, 6, 4), (8.0, 9.0, 3.0)), ((5, 6, 3), None),
                                       ((7, 5, 4), (7.0, 4.0, 6.0))])
def test_projected_noise_is_real_divergence_free_and_projected(cells, box):
    o = _oracle(cells, box=box)
    for nset in (1, 2):
        u, mi = O.kspace_noise(o, 4, nset, 1.3)
        top = np.abs(u).max()
        assert top > 0.1
        assert mi < 1e-13 * top                             # Hermitian spectrum
        assert np.abs(o.divergence(u)).max() < 1e-12 * top  # discretely divergence-free
        assert np.abs(o.project(u) - u).max() < 1e-12 * top  # the solve's projection keeps it


def _mode_cov(fields, cells):
    mx, my, mz = cells
    n = mx * my * mz
    X = np.fft.fftn(fields.reshape(3, mz, my, mx), axes=(1, 2, 3)).reshape(3, n)
    return X[:, None] * X[None].conj()


def test_projected_noise_has_the_law_of_the_projected_stress_divergence():
    """Over 800 draws on a box with unequal cell sizes: the per-mode 3x3 covariance of the new
    draw and of the pinned projection of the pinned stress divergence (f1 = sqrt(2) f4) both
    equal N f4^2 |g|^2 E P E*, and each other, to sampling accuracy."""
    cells, box = (6, 4, 4), (6.0, 5.0, 3.0)
    o = _oracle(cells, box=box, seed=23)
    mx, my, mz = cells
    N = o.N
    h = (box[0] / mx, box[1] / my, box[2] / mz)
    kz, ky, kx = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    k = (kx, ky, kz)
    m = (mx, my, mz)
    g = np.stack([2 / h[a] * np.sin(np.pi * k[a] / m[a]) for a in range(3)]).reshape(3, N)
    e = np.exp(1j * np.pi * np.stack([k[a] / m[a] for a in range(3)])).reshape(3, N)
    g2 = (g ** 2).sum(0)
    safe = np.where(g2 > 0, g2, 1.0)
    f4 = 0.8
    C = np.zeros((3, 3, N), complex)
    for a in range(3):
        for b in range(3):
            C[a, b] = N * f4 ** 2 * (g2 * (a == b) - g[a] * g[b] * (g2 > 0)) * e[a] * np.conj(e[b])
    nd = 800
    A = np.zeros((3, 3, N), complex)
    B = np.zeros((3, 3, N), complex)
    for st in range(nd):
        u, _ = O.kspace_noise(o, st, 1, f4)
        _, w, _ = O.kspace_stress(o, st, 1, np.sqrt(2) * f4, f4)
        A += _mode_cov(u, cells)
        B += _mode_cov(o.project(w), cells)
    A /= nd
    B /= nd
    tr = np.real(C[0, 0] + C[1, 1] + C[2, 2])
    live = tr > 0
    assert not np.any(np.abs(A[:, :, ~live]) > 1e-20) and live.sum() == N - 1  # only the mean is empty
    tol = 4.0 / np.sqrt(nd)
    for X in (A, B):
        dev = np.abs(X[:, :, live] - C[:, :, live]).max(axis=(0, 1)) / tr[live]
        assert dev.max() < tol, dev.max()
    dev = np.abs(A[:, :, live] - B[:, :, live]).max(axis=(0, 1)) / tr[live]
    assert dev.max() < 1.5 * tol
