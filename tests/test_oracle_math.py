"""Pins the oracle's dense building blocks (restated un-vendored linear algebra) against
mathematics: scipy's expm, exact Krylov exactness, unitarity, Householder QR identities."""
import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sla

from helpers import random_hermitian


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def orc_expm(orc, M):
    n = M.shape[0]
    m = np.ascontiguousarray(M, dtype=complex)
    e = np.zeros_like(m)
    assert orc.fn["expm"](n, _dp(m.view(np.float64)), _dp(e.view(np.float64))) == 0
    return e


def orc_expv(orc, H, x, c, tau, kdim, tol):
    n = H.shape[0]
    h = np.ascontiguousarray(H, dtype=complex)
    v = np.ascontiguousarray(x, dtype=complex).copy()
    na = C.c_int64()
    rc = orc.fn["expv"](n, _dp(h.view(np.float64)), _dp(v.view(np.float64)), c.real, c.imag, tau, kdim, tol,
                        C.byref(na))
    assert rc == 0
    return v, na.value


def orc_qr(orc, A):
    m, n = A.shape
    a = np.ascontiguousarray(A, dtype=complex)
    Q = np.zeros((m, n), dtype=complex)
    R = np.zeros((n, n), dtype=complex)
    assert orc.fn["qr"](m, n, _dp(a.view(np.float64)), _dp(Q.view(np.float64)), _dp(R.view(np.float64))) == 0
    return Q, R


@pytest.mark.parametrize("n", [1, 2, 5, 12, 16])
@pytest.mark.parametrize("scale", [1e-6, 0.3, 4.0, 60.0])
def test_expm_matches_scipy(orc, n, scale):
    rng = np.random.default_rng(n * 1000 + int(scale * 10))
    M = scale * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    E = orc_expm(orc, M)
    ref = sla.expm(M)
    assert np.abs(E - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_expm_skew_hermitian_is_unitary(orc):
    rng = np.random.default_rng(3)
    H = random_hermitian(rng, 12, 5.0)
    U = orc_expm(orc, -1j * H)
    assert np.abs(U.conj().T @ U - np.eye(12)).max() < 1e-13


@pytest.mark.parametrize("n,kdim,tau", [(40, 12, 0.05), (40, 12, 0.5), (64, 16, 0.025), (30, 6, 1.0)])
def test_expv_matches_dense_propagator(orc, n, kdim, tau):
    rng = np.random.default_rng(n + kdim)
    H = random_hermitian(rng, n, 2.0)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y, napply = orc_expv(orc, H, x, -1j, tau, kdim, 1e-12)
    ref = sla.expm(-1j * tau * H) @ x
    assert np.abs(y - ref).max() <= 1e-10 * np.linalg.norm(x)
    # unitary propagation keeps the norm
    assert abs(np.linalg.norm(y) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)
    assert napply >= 1


def test_expv_exact_when_krylov_space_is_invariant(orc):
    # rank-3 Hamiltonian acting on a vector inside its range: the Krylov space closes
    # after 3 vectors (happy breakdown) and the result is exact for any step length
    rng = np.random.default_rng(7)
    U, _ = np.linalg.qr(rng.standard_normal((20, 3)) + 1j * rng.standard_normal((20, 3)))
    H = U @ np.diag([1.0, -2.0, 5.0]) @ U.conj().T
    x = U @ np.array([1.0, 0.5j, -0.25])
    y, napply = orc_expv(orc, H, x, -1j, 3.7, 12, 1e-12)
    assert napply <= 4
    assert np.abs(y - sla.expm(-3.7j * H) @ x).max() < 1e-13


def test_expv_rabi_closed_form(orc):
    # H = Delta sigma_x: exp(-i t H)|0> = cos(Delta t)|0> - i sin(Delta t)|1>
    D, t = 1.3, 0.9
    H = np.array([[0, D], [D, 0]], dtype=complex)
    y, _ = orc_expv(orc, H, np.array([1, 0], dtype=complex), -1j, t, 4, 1e-14)
    assert np.abs(y - np.array([np.cos(D * t), -1j * np.sin(D * t)])).max() < 1e-15


def test_expv_substeps_when_krylov_dim_is_small(orc):
    # large ||tau H|| with a small Krylov space forces step halving; still accurate
    rng = np.random.default_rng(11)
    H = random_hermitian(rng, 50, 10.0)
    x = rng.standard_normal(50) + 0j
    y, napply = orc_expv(orc, H, x, -1j, 1.0, 8, 1e-12)
    assert napply > 8
    assert np.abs(y - sla.expm(-1j * H) @ x).max() < 1e-9


def test_expv_backward_time_inverts_forward(orc):
    rng = np.random.default_rng(5)
    H = random_hermitian(rng, 24)
    x = rng.standard_normal(24) + 1j * rng.standard_normal(24)
    y, _ = orc_expv(orc, H, x, -1j, 0.3, 12, 1e-13)
    z, _ = orc_expv(orc, H, y, -1j, -0.3, 12, 1e-13)
    assert np.abs(z - x).max() < 1e-12


@pytest.mark.parametrize("m,n", [(8, 8), (64, 8), (100, 17), (5, 1), (256, 16)])
def test_householder_qr_identities(orc, m, n):
    rng = np.random.default_rng(m * 31 + n)
    A = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    Q, R = orc_qr(orc, A)
    assert np.abs(Q @ R - A).max() < 1e-13 * np.abs(A).max() * m
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-13 * m
    assert np.abs(np.tril(R, -1)).max() == 0.0
    assert np.abs(np.diag(R).imag).max() == 0.0
    # LAPACK sign convention: beta = -sign(Re alpha) * norm, so R_ii carries the
    # opposite sign of the column's leading entry after the previous reflections
    ref_R = np.linalg.qr(A, mode="r")
    assert np.allclose(np.abs(np.diag(R)), np.abs(np.diag(ref_R)), rtol=1e-12)


def test_qr_rank_deficient_completion_is_deterministic(orc):
    # an exactly zero column leaves its reflector as the identity (tau = 0), so the
    # completing column of Q is fixed by the previous reflectors only
    A = np.zeros((6, 3), dtype=complex)
    A[:, 0] = [1, 0, 0, 0, 0, 0]
    A[:, 2] = [0, 0, 3, 0, 0, 0]
    Q, R = orc_qr(orc, A)
    assert np.abs(Q @ R - A).max() < 1e-15
    assert np.abs(Q.conj().T @ Q - np.eye(3)).max() < 1e-15
    # column 1 is e_1 (reflector 0 is the identity on e_1 since A[1:,0] = 0 and A00 > 0)
    assert np.allclose(Q[:, 1], np.eye(6)[1])


def test_qr_graded_columns_stay_finite(orc):
    # columns scaled from 1 down to 1e-300: squares of the trailing columns underflow, so
    # the reflectors need the scaled norm and zlarfg's 1/safmin rescaling.  Householder QR
    # is columnwise backward stable, and LAPACK (numpy) gives the |R_jj| to compare with.
    rng = np.random.default_rng(11)
    m, n = 40, 16
    G = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    A = G * np.logspace(0, -300, n)[None, :]
    Q, R = orc_qr(orc, A)
    assert np.isfinite(Q).all() and np.isfinite(R).all()
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() < 1e-13
    s = np.abs(A).max(axis=0)  # column scales (numpy's norm would underflow too)
    colerr = np.linalg.norm((Q @ R - A) / s, axis=0) / np.linalg.norm(A / s, axis=0)
    assert colerr.max() < 1e-13
    ref = np.abs(np.diag(np.linalg.qr(A, mode="r")))
    assert np.allclose(np.abs(np.diag(R)), ref, rtol=1e-10, atol=0)


def test_qr_column_with_underflowing_squares(orc):
    # every square in column 1 underflows to zero: the plain formula gives beta = -0 and
    # tau = inf; the scaled norm and safmin rescaling give an exact reflector
    A = np.zeros((3, 2), dtype=complex)
    A[:, 0] = [1, 0, 0]
    A[:, 1] = [0, 1e-170 * (1 + 1j), 2e-170]
    Q, R = orc_qr(orc, A)
    assert np.isfinite(Q).all() and np.isfinite(R).all()
    assert np.abs(Q.conj().T @ Q - np.eye(2)).max() < 1e-15
    assert abs(abs(R[1, 1]) - np.sqrt(6) * 1e-170) < 1e-15 * 1e-170 * 4
    assert np.abs(Q @ R - A)[:, 1].max() < 1e-185


def test_qr_wide_matrix_is_zero_padded(orc):
    rng = np.random.default_rng(2)
    A = rng.standard_normal((2, 5)) + 1j * rng.standard_normal((2, 5))
    Q, R = orc_qr(orc, A)
    assert Q.shape == (2, 5) and R.shape == (5, 5)
    assert np.abs(Q[:, 2:]).max() == 0 and np.abs(R[2:, :]).max() == 0
    assert np.abs(Q @ R - A).max() < 1e-14
