"""The strided FP64 MFMA GEMM kernel against the plain kernel and numpy."""
import numpy as np
import pytest

from libuw12_amd import _lib

pytestmark = pytest.mark.gpu


def M2(s0, s1=0, n0=0):
    return _lib.Map2(s0, s1, n0)


def host_gemm(M, N, K, batch, alpha, beta, A, Am, Ak, Ab, B, Bk, Bn, Bb, C, Cm, Cn, Cb, ks=None, Km=None):
    def off(m, i):
        return i * m.s0 if m.n0 == 0 else (i % m.n0) * m.s0 + (i // m.n0) * m.s1
    C = C.copy()
    mi, ki, ni = np.arange(M), np.arange(K), np.arange(N)
    for b in range(batch):
        Amat = A[b * Ab + off(Am, mi)[:, None] + off(Ak, ki)[None, :]]
        if ks is not None:
            Amat = Amat * ks[off(Km, ki)][None, :]
        Bmat = B[b * Bb + off(Bk, ki)[:, None] + off(Bn, ni)[None, :]]
        idx = b * Cb + off(Cm, mi)[:, None] + off(Cn, ni)[None, :]
        C[idx] = alpha * (Amat @ Bmat) + (beta * C[idx] if beta != 0 else 0)
    return C


def run(M, N, K, batch, alpha, beta, A, Am, Ak, Ab, B, Bk, Bn, Bb, C, Cm, Cn, Cb, ks=None, Km=None, naive=0):
    L = _lib.load()
    out = C.copy()
    p = lambda a: a.ctypes.data_as(_lib.dp)
    rc = L.uw12g_test_gemm(0, M, N, K, batch, alpha, beta, p(A), A.size, Am, Ak, Ab, p(B), B.size, Bk, Bn, Bb,
                           p(out), out.size, Cm, Cn, Cb, p(ks) if ks is not None else None,
                           ks.size if ks is not None else 0, Km if Km is not None else M2(0), naive)
    _lib.check(rc)
    return out


CASES = [
    # name, M, N, K, batch, maps
    ('nn_square', 130, 70, 45, 1),
    ('big', 257, 300, 129, 1),
    ('thin_k_split', 96, 80, 4000, 1),
    ('batched', 100, 33, 64, 5),
    ('tiny', 3, 5, 2, 2),
]


@pytest.mark.parametrize('name,M,N,K,batch', CASES)
@pytest.mark.parametrize('layout', ['NN', 'TN', 'NT', 'TT'])
def test_plain_layouts(name, M, N, K, batch, layout):
    rng = np.random.default_rng(hash((name, layout)) % 2**32)
    A = rng.standard_normal(M * K * batch)
    B = rng.standard_normal(K * N * batch)
    C = rng.standard_normal(M * N * batch)
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    args = (M, N, K, batch, 1.25, -0.5, A, Am, Ak, M * K, B, Bk, Bn, K * N, C, M2(1), M2(M), M * N)
    ref = host_gemm(*args)
    got = run(*args)
    naive = run(*args, naive=1)
    scale = np.abs(ref).max()
    assert np.abs(naive - ref).max() <= 1e-12 * scale
    assert np.abs(got - ref).max() <= 1e-12 * scale


def test_two_level_maps_and_kscale():
    # C[(b,k), n] = sum_{(j,A)} X[b + 7 j + 7*5*A] s[A] Y[...]: compound K of (j, A) and compound M
    rng = np.random.default_rng(3)
    nb, nk, nj, nA, N = 7, 4, 5, 9, 11
    M, K = nb * nk, nj * nA
    A = rng.standard_normal(nb * 13 * nA * 2 + 500)
    B = rng.standard_normal(K * N * 3 + 100)
    ks = rng.standard_normal(nA) + 2
    Am = M2(1, 13 * 0 + 9, nb)            # m = (b, k): b stride 1, k stride 9
    Ak = M2(60, 500 // nA, nj)            # k = (j, A)
    Bk = M2(N, 0, 0)
    Bn = M2(1)
    Km = M2(0, 1, nj)                     # scale by s[A], A = k div nj
    C = np.zeros(M * N * 2 + 50)
    Cm = M2(1, 2 * nb, nb)
    Cn = M2(2 * nb * nk)
    args = (M, N, K, 1, 1.0, 0.0, A, Am, Ak, 0, B, Bk, Bn, 0, C, Cm, Cn, 0, ks, Km)
    ref = host_gemm(*args)
    assert np.abs(run(*args) - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(run(*args, naive=1) - ref).max() <= 1e-12 * np.abs(ref).max()


def test_zero_k_scales_c():
    C = np.arange(12.0)
    A = np.zeros(1)
    B = np.zeros(1)
    out = run(3, 4, 0, 1, 1.0, 2.0, A, M2(1), M2(3), 0, B, M2(1), M2(1), 0, C, M2(1), M2(3), 0)
    assert np.array_equal(out, 2 * C)
