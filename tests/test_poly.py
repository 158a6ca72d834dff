"""Host side of the polynomial inner preconditioner (svfsi_poly_factors):
the factors must carry exactly the harmonic Ritz values of the Hessenberg
matrix (numpy's eigenvalues of H_d + h^2 H_d^-T e_d e_d^T), each conjugate
pair in one real quadratic factor, in Leja order (largest modulus first)."""
import ctypes as C

import numpy as np
import pytest

from svfsi_amd import _lib


def harmonic_ritz(Hu):
    d = Hu.shape[1]
    Hd = Hu[:d, :d]
    e = np.zeros(d)
    e[-1] = 1.0
    f = np.linalg.solve(Hd.T, e)
    return np.linalg.eigvals(Hd + Hu[d, d - 1] ** 2 * np.outer(f, e))


def factors(Hu):
    d = Hu.shape[1]
    a = np.asfortranarray(Hu)  # column-major, ld = d + 1
    coef = np.zeros(2 * d)
    kind = np.zeros(d, np.int32)
    nf = _lib.lib().svfsi_poly_factors(d, d + 1, a.ctypes.data_as(C.POINTER(C.c_double)),
                                       coef.ctypes.data_as(C.POINTER(C.c_double)),
                                       kind.ctypes.data_as(C.POINTER(C.c_int)), d)
    assert nf > 0
    return coef[:2 * nf].reshape(nf, 2), kind[:nf]


def roots_of(coef, kind):
    out = []
    for (c1, c2), k in zip(coef, kind):
        if k == 1:
            out.append(1.0 / c1)
        else:  # 1 - c1 a + c2 a^2 = 0
            out.extend(np.roots([c2, -c1, 1.0]))
    return np.array(out, dtype=complex)


def arnoldi_hessenberg(A, d, seed):
    n = A.shape[0]
    Q = np.zeros((n, d + 1))
    H = np.zeros((d + 1, d))
    v = np.random.default_rng(seed).normal(size=n)
    Q[:, 0] = v / np.linalg.norm(v)
    for j in range(d):
        w = A @ Q[:, j]
        for _ in range(2):
            h = Q[:, :j + 1].T @ w
            w -= Q[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        Q[:, j + 1] = w / H[j + 1, j]
    return H


@pytest.mark.parametrize("seed,d", [(0, 8), (1, 20), (2, 30), (3, 40)])
def test_poly_factors_carry_the_harmonic_ritz_values(seed, d):
    rng = np.random.default_rng(seed)
    # nonsymmetric, eigenvalues in the right half plane, some complex
    n = 200
    A = np.diag(rng.uniform(0.1, 4.0, n)) + 0.3 * rng.normal(size=(n, n)) / np.sqrt(n)
    Hu = arnoldi_hessenberg(A, d, seed)
    ref = harmonic_ritz(Hu)
    coef, kind = factors(Hu)
    got = roots_of(coef, kind)
    assert len(got) == d
    # every reference value matched by a distinct computed root
    used = np.zeros(d, bool)
    scale = np.abs(ref).max()
    for z in ref:
        dist = np.where(used, np.inf, np.abs(got - z))
        k = int(np.argmin(dist))
        assert dist[k] <= 1e-8 * scale, (z, got[k])
        used[k] = True
    # Leja start: the first factor holds the root of largest modulus
    first = roots_of(coef[:1], kind[:1])
    assert np.isclose(np.abs(first).max(), np.abs(ref).max(), rtol=1e-8)


def test_poly_factors_rejects_singular_input():
    Hu = np.zeros((5, 4))
    coef = np.zeros(8)
    kind = np.zeros(4, np.int32)
    a = np.asfortranarray(Hu)
    rc = _lib.lib().svfsi_poly_factors(4, 5, a.ctypes.data_as(C.POINTER(C.c_double)),
                                       coef.ctypes.data_as(C.POINTER(C.c_double)),
                                       kind.ctypes.data_as(C.POINTER(C.c_int)), 4)
    assert rc < 0
