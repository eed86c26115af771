"""The strided MFMA GEMM kernel (FP64, FP32, BF16 and FP16 split-product forms) against the plain kernel and numpy."""
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
    # large outputs on 128-wide tiles with a partly filled last row / column of tiles
    ('edge_rows_cols', 900, 800, 70, 1),
    ('edge_cols_batched', 1000, 960, 33, 2),
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


def fp32_bound(M, N, K, batch, alpha, A, Am, Ak, Ab, B, Bk, Bn, Bb, Cm, Cn, Cb, ks=None, Km=None, chunk=4096):
    """Elementwise bound on |C_fp32 - C_exact| for the FP32 form: operands rounded
    to FP32 (relative 2^-24 each) and FP32 recursive summation over k chunks of at
    most `chunk` (gamma_n = n u / (1 - n u), Higham, Accuracy and Stability of
    Numerical Algorithms, 2nd ed., eq. 3.5), chunks added in FP64."""
    u = 2.0 ** -24
    n = min(K, chunk) + 2
    g = n * u / (1 - n * u)
    absC = host_gemm(M, N, K, batch, abs(alpha), 0.0, np.abs(A), Am, Ak, Ab, np.abs(B), Bk, Bn, Bb,
                     np.zeros(_c_size(M, N, batch, Cm, Cn, Cb)), Cm, Cn, Cb,
                     None if ks is None else np.abs(ks), Km)
    return g * absC


# Per-product relative error of the BF16 split forms (test hook modes 3 and 4) and
# the FP32 terms summed per k.  Mode 3 (two planes, hh + hl + lh): each operand
# is h + l to 2^-16 (h = bf16(v), l = bf16(v - h), unit roundoff 2^-8 per plane)
# and the dropped ll is below 2^-16: 3 * 2^-16 in all.  Mode 4 (three planes, six
# products hh hm mh hl mm lh): operands to about 2^-24, dropped ml + lm + ll below
# 2^-23: 2^-22 in all.  Small factors cover the FP32 roundings inside the splits.
# Mode 5 (FP16x3: rows scaled by powers of two, two half planes, hh + hl + lh):
# each operand is h + l to 2^-22 (11 significant bits per plane) and the dropped
# ll is below 2^-22: 3 * 2^-22; the in-kernel fallback (no workspace) runs the
# three-plane BF16 form, so the FP32 terms are counted as six per k.
BF16_FORMS = {3: (3.1 * 2.0 ** -16, 3), 4: (1.1 * 2.0 ** -22, 6), 5: (3.1 * 2.0 ** -22, 6)}


def reduced_bound(mode, M, N, K, batch, alpha, A, Am, Ak, Ab, B, Bk, Bn, Bb, Cm, Cn, Cb, ks=None, Km=None,
                  chunk=4096):
    """Elementwise bound on |C - C_exact| for a reduced form: mode 2 the FP32 form
    (fp32_bound), modes 3 / 4 / 5 the split forms: (per-product error + gamma_n)
    times |A| |B|, with n the FP32 additions of one k chunk (products per k times
    the chunk length).  Mode 5 adds the half planes' subnormal floor: elements of
    a row scaled to a maximum in [2^14, 2^15) keep an absolute error below 2^-25
    in l, i.e. 2^-38 of the row maximum, bounded here by 2^-37 K max|A| max|B|."""
    if mode == 2:
        return fp32_bound(M, N, K, batch, alpha, A, Am, Ak, Ab, B, Bk, Bn, Bb, Cm, Cn, Cb, ks, Km, chunk)
    rel, per_k = BF16_FORMS[mode]
    u = 2.0 ** -24
    n = per_k * min(K, chunk) + 2
    g = n * u / (1 - n * u)
    absC = host_gemm(M, N, K, batch, abs(alpha), 0.0, np.abs(A), Am, Ak, Ab, np.abs(B), Bk, Bn, Bb,
                     np.zeros(_c_size(M, N, batch, Cm, Cn, Cb)), Cm, Cn, Cb,
                     None if ks is None else np.abs(ks), Km)
    floor = 0.0
    if mode == 5:
        ka = 1.0 if ks is None else np.abs(ks).max()
        floor = 2.0 ** -37 * K * abs(alpha) * np.abs(A).max() * ka * np.abs(B).max()
    return (rel + g) * absC + floor


def _c_size(M, N, batch, Cm, Cn, Cb):
    def mx(m, e):
        if e <= 0:
            return 0
        last = e - 1
        if m.n0 == 0:
            return last * m.s0
        q, r = divmod(last, m.n0)
        best = r * m.s0 + q * m.s1
        if q >= 1:
            best = max(best, (m.n0 - 1) * m.s0 + (q - 1) * m.s1)
        return best
    return Cb * (batch - 1) + mx(Cm, M) + mx(Cn, N) + 1


@pytest.mark.parametrize('name,M,N,K,batch', CASES)
@pytest.mark.parametrize('layout', ['NN', 'TN', 'NT', 'TT'])
def test_fp32_form_within_rounding_bound(name, M, N, K, batch, layout):
    rng = np.random.default_rng(hash(('fp32', name, layout)) % 2**32)
    A = rng.standard_normal(M * K * batch)
    B = rng.standard_normal(K * N * batch)
    C = np.zeros(M * N * batch)
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    args = (M, N, K, batch, 1.25, 0.0, A, Am, Ak, M * K, B, Bk, Bn, K * N, C, M2(1), M2(M), M * N)
    ref = host_gemm(*args)
    got = run(*args, naive=2)
    bound = fp32_bound(M, N, K, batch, 1.25, A, Am, Ak, M * K, B, Bk, Bn, K * N, M2(1), M2(M), M * N)
    err = np.abs(got - ref)
    assert np.all(err <= bound + 1e-300), (err / np.maximum(bound, 1e-300)).max()
    # and it is really FP32: not bit-equal to the FP64 result on a generic input
    assert np.abs(got - ref).max() > 0


@pytest.fixture
def bf16_path(request):
    """Operand splitting of the BF16 forms: 'planes' (plane tensors in the
    workspace, the default), 'inkernel' (split inside the GEMM kernel) or
    'pieces' (plane tensors with a 256 KiB cap, so products run in pieces)."""
    mode = getattr(request, 'param', 'planes')
    L = _lib.load()
    L.uw12g_test_set_bf16_planes(0 if mode == 'inkernel' else 1, 262144 if mode == 'pieces' else 0)
    yield mode
    L.uw12g_test_set_bf16_planes(1, 0)


@pytest.mark.parametrize('name,M,N,K,batch', CASES)
@pytest.mark.parametrize('layout', ['NN', 'TN', 'NT', 'TT'])
@pytest.mark.parametrize('mode', [3, 4, 5])
@pytest.mark.parametrize('bf16_path', ['planes', 'inkernel', 'pieces'], indirect=True)
def test_bf16_forms_within_rounding_bound(name, M, N, K, batch, layout, mode, bf16_path):
    """Split forms (two and three bfloat16 planes, two half planes) within their
    error bound, with the operands split into plane tensors, inside the kernel,
    or in pieces; the three-plane BF16 and the FP16 forms are within a small
    factor of the FP32 form's error."""
    rng = np.random.default_rng(hash(('bf16', name, layout, mode)) % 2**32)
    A = rng.standard_normal(M * K * batch)
    B = rng.standard_normal(K * N * batch)
    C = rng.standard_normal(M * N * batch)
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    args = (M, N, K, batch, 1.25, 0.5, A, Am, Ak, M * K, B, Bk, Bn, K * N, C, M2(1), M2(M), M * N)
    ref = host_gemm(*args)
    got = run(*args, naive=mode)
    bound = reduced_bound(mode, M, N, K, batch, 1.25, A, Am, Ak, M * K, B, Bk, Bn, K * N, M2(1), M2(M), M * N)
    err = np.abs(got - ref)
    bound = bound + 1e-15 * np.abs(ref)  # the beta C term, in FP64
    assert np.all(err <= bound), (err / bound).max()
    assert err.max() > 0
    if mode in (4, 5) and K >= 16:
        e32 = np.abs(run(*args, naive=2) - ref).max()
        assert err.max() <= 8 * e32, (err.max(), e32)


@pytest.mark.parametrize('layout', ['NN', 'TN', 'NT', 'TT'])
@pytest.mark.parametrize('mode', [3, 5])
@pytest.mark.parametrize('store', [1, 2, 4, 7])
def test_split_forms_fp32_storage(layout, mode, store):
    """Split forms reading A / B from FP32 storage and writing C to FP32 storage
    (test-hook mode bits 8 / 16 / 32; the host values rounded to float): the
    same error bound as FP64 storage on the rounded operands, plus the final
    rounding of C to float; beta C read from FP32 storage too."""
    M, N, K, batch = 300, 200, 150, 2
    rng = np.random.default_rng(hash(('f32', layout, mode, store)) % 2**32)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    A = rng.standard_normal(M * K * batch)
    B = rng.standard_normal(K * N * batch)
    C = rng.standard_normal(M * N * batch)
    A, B = (f32(A) if store & 1 else A), (f32(B) if store & 2 else B)
    C = f32(C) if store & 4 else C
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    args = (M, N, K, batch, 1.25, 0.5, A, Am, Ak, M * K, B, Bk, Bn, K * N, C, M2(1), M2(M), M * N)
    ref = host_gemm(*args)
    got = run(*args, naive=mode + 8 * store)
    bound = reduced_bound(mode, M, N, K, batch, 1.25, A, Am, Ak, M * K, B, Bk, Bn, K * N, M2(1), M2(M), M * N)
    bound = bound + (2.0 ** -24 if store & 4 else 1e-15) * np.abs(ref)
    err = np.abs(got - ref)
    assert np.all(err <= bound), (err / bound).max()
    if store & 4:
        assert np.array_equal(got, f32(got))  # C came back from FP32 storage


@pytest.mark.parametrize('mode', [3, 5])
@pytest.mark.parametrize('store', [0, 4])
@pytest.mark.parametrize('N', [200, 204])
def test_split_forms_row_major_output(mode, store, N):
    """Split forms writing an output whose columns are contiguous (row-major C, so each lane's
    four columns are one vector store; beta C read the same way), FP64 or FP32 output storage,
    batched, with a partial last tile of columns."""
    M, K, batch = 300, 150, 2
    rng = np.random.default_rng(hash(('rowmajor', mode, store, N)) % 2**32)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    A = rng.standard_normal(M * K * batch)
    B = rng.standard_normal(K * N * batch)
    C = rng.standard_normal(M * N * batch)
    C = f32(C) if store & 4 else C
    args = (M, N, K, batch, 1.25, 0.5, A, M2(1), M2(M), M * K, B, M2(1), M2(K), K * N, C, M2(N), M2(1), M * N)
    ref = host_gemm(*args)
    got = run(*args, naive=mode + 8 * store)
    bound = reduced_bound(mode, M, N, K, batch, 1.25, A, M2(1), M2(M), M * K, B, M2(1), M2(K), K * N, M2(N), M2(1),
                          M * N)
    bound = bound + (2.0 ** -24 if store & 4 else 1e-15) * np.abs(ref)
    err = np.abs(got - ref)
    assert np.all(err <= bound), (err / bound).max()


def test_fp32_storage_needs_the_split_forms():
    """FP32 operand storage is read only by the plane-tensor path: the FP64 and
    FP32 forms refuse such descriptors before launching anything."""
    A, B, C = np.ones(64), np.ones(64), np.zeros(64)
    for mode in (0, 2):
        with pytest.raises(_lib.UW12Error):
            run(8, 8, 8, 1, 1.0, 0.0, A, M2(1), M2(8), 0, B, M2(1), M2(8), 0, C, M2(1), M2(8), 0, naive=mode + 8)


@pytest.mark.parametrize('layout', ['NN', 'TN', 'NT', 'TT'])
def test_fp16_form_rows_of_wide_range(layout):
    """FP16x3 with operand rows and columns spread over 2^-40 .. 2^40 (beyond the
    half range of 2^-24 .. 65504 for any single scale): every row of A and column
    of B gets its own power-of-two scale, so each output element stays within
    the split forms' bound of its own |A| |B| and the error is far below the
    two-plane BF16 form's on the same operands."""
    M, N, K = 200, 136, 300
    rng = np.random.default_rng(hash(('wide', layout)) % 2**32)
    A = rng.standard_normal((M, K)) * np.exp2(rng.uniform(-40, 40, M))[:, None]
    B = rng.standard_normal((K, N)) * np.exp2(rng.uniform(-40, 40, N))[None, :]
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    Af = A.flatten(order='F') if layout[0] == 'N' else A.flatten(order='C')
    Bf = B.flatten(order='F') if layout[1] == 'N' else B.flatten(order='C')
    args = (M, N, K, 1, 1.0, 0.0, Af, Am, Ak, 0, Bf, Bk, Bn, 0, np.zeros(M * N), M2(1), M2(M), 0)
    ref = host_gemm(*args)
    got = run(*args, naive=5)
    absC = (np.abs(A) @ np.abs(B)).flatten(order='F')
    u32 = 2.0 ** -24
    g = (6 * K + 2) * u32 / (1 - (6 * K + 2) * u32)
    rel = np.abs(got - ref) / absC
    assert np.all(rel <= 3.1 * 2.0 ** -22 + g), rel.max()
    rel3 = np.abs(run(*args, naive=3) - ref) / absC
    # the FP32 sums bound both forms' largest relative errors from below: measured 1.4e-7 (FP16x3)
    # against 1.5e-6 (BF16x3)
    assert rel.max() <= rel3.max() / 4, (rel.max(), rel3.max())


@pytest.mark.parametrize('mode', [0, 2, 3, 4, 5])
@pytest.mark.parametrize('K', [16 * 40, 16 * 40 + 7])
@pytest.mark.parametrize('kc', [False, True])
@pytest.mark.parametrize('n0', [32, 16, 12, 20])
@pytest.mark.parametrize('bf16_path', ['planes', 'inkernel'], indirect=True)
def test_affine_two_level_k_maps(mode, K, kc, n0, bf16_path):
    """Compound k = (r, q) with an inner period n0: with n0 a multiple of the slab
    depth (16 for FP64 / FP32, 32 for the BF16 forms) the unpredicated buffer-load
    path, in both the k-along-waves form (kc False) and the k-along-lanes form (kc
    True, where a k remainder runs as a second launch), with a k scale; with n0 = 16
    the BF16 forms take the general path; with n0 = 12 some 8-k units of the plane
    split cross a period (its per-k offsets); with n0 = 20 and whole periods (K = 640)
    the FP64 / FP32 forms run the first 16 of every period on the unpredicated path
    and the other 4 as a second product."""
    if n0 == 16 and mode < 3:
        pytest.skip('n0 = 16 is already slab-affine for the 16-deep slabs')
    if mode < 3 and bf16_path != 'planes':
        pytest.skip('operand splitting applies to the BF16 forms only')
    rng = np.random.default_rng(17 + K + 2 * kc + mode + n0)
    M, N = 200, 150
    q = (K + n0 - 1) // n0
    # A(m, k) = A[m*sm + (k % n0)*s0 + (k // n0)*s1]
    if kc:
        Am, Ak = M2(n0 * q + 3), M2(1, n0, n0)
    else:
        Am, Ak = M2(1), M2(M + 5, (M + 5) * n0 + 11, n0)
    A = rng.standard_normal(M * (n0 * q + 3) + (M + 5) * n0 * q + 11 * q + 1000)
    Bk, Bn = (M2(1, n0 + 0, n0), M2(n0 * q + 1)) if kc else (M2(N, N * n0 + 3, n0), M2(1))
    B = rng.standard_normal(N * (n0 * q + 1) + N * n0 * q + 3 * q + 1000)
    ks = rng.standard_normal(n0 * q + 10)
    Km = M2(1, n0, n0)
    C = rng.standard_normal(M * N)
    args = (M, N, K, 1, 0.75, 0.5 if mode == 0 else 0.0, A, Am, Ak, 0, B, Bk, Bn, 0, C, M2(1), M2(M), 0, ks, Km)
    ref = host_gemm(*args)
    got = run(*args, naive=mode)
    if mode == 0:
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        bound = reduced_bound(mode, M, N, K, 1, 0.75, A, Am, Ak, 0, B, Bk, Bn, 0, M2(1), M2(M), 0, ks, Km)
        # the beta C term is exact
        assert np.all(np.abs(got - ref) <= bound + 1e-15 * np.abs(ref))


@pytest.mark.parametrize('side', ['A', 'B'])
@pytest.mark.parametrize('mode', [0, 2, 3, 4, 5])
@pytest.mark.parametrize('bf16_path', ['planes', 'inkernel'], indirect=True)
def test_operand_span_past_lane_offsets(side, mode, bf16_path):
    """An operand whose row (A) or column (B) offsets pass 2^28 doubles, the limit of the
    buffer path's 32-bit lane offsets: the launcher splits the output into row / column
    blocks launched from 64-bit base pointers.  The output map has an inner period of 30
    rows / columns, so the blocks must also be whole periods of it."""
    if mode < 3 and bf16_path != 'planes':
        pytest.skip('operand splitting applies to the BF16 forms only')
    rng = np.random.default_rng(17 + mode)
    big, K, n_small, stride = 300, 40, 50, 1 << 20
    Abig = np.zeros(big * stride)
    Abig.reshape(big, stride)[:, :K] = rng.standard_normal((big, K))
    small = rng.standard_normal(K * n_small)
    if side == 'A':
        M, N = big, n_small
        A, Am, Ak = Abig, M2(stride), M2(1)
        B, Bk, Bn = small, M2(1), M2(K)
        Cm, Cn = M2(1, 1000, 30), M2(10 * 1000)
        C = rng.standard_normal(10 * 1000 * N)
    else:
        M, N = n_small, big
        A, Am, Ak = small, M2(K), M2(1)
        B, Bk, Bn = Abig, M2(1), M2(stride)
        Cm, Cn = M2(1), M2(64, 10000, 30)
        C = rng.standard_normal(10000 * (big // 30) + 64 * 30)
    args = (M, N, K, 1, 0.75, 0.5, A, Am, Ak, 0, B, Bk, Bn, 0, C, Cm, Cn, 0)
    ref = host_gemm(*args)
    got = run(*args, naive=mode)
    if mode == 0:
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        bound = reduced_bound(mode, M, N, K, 1, 0.75, A, Am, Ak, 0, B, Bk, Bn, 0, Cm, Cn, 0)
        bound = np.pad(bound, (0, C.size - bound.size))  # elements past the output map are untouched
        assert np.all(np.abs(got - ref) <= bound + 1e-15 * np.abs(ref))


@pytest.mark.parametrize('mode', [3, 5])
@pytest.mark.parametrize('kc', [False, True])
@pytest.mark.parametrize('n0', [12, 20])
def test_split_forms_fp32_two_level_k(mode, kc, n0):
    """FP32 storage of both operands read through two-level k maps whose period is not a multiple
    of 8 (some 8-k units of the plane split cross it, and the last one passes K), for k strided with
    the rows unit stride (kc False) and for k unit stride within the period with the rows strided
    (kc True, the LDS-staged split): the same bound as FP64 storage on the rounded operands."""
    rng = np.random.default_rng(hash(('f32k', mode, kc, n0)) % 2**32)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    M, N, K, batch = 170, 150, 157, 2
    q = (K + n0 - 1) // n0
    if kc:
        Am, Ak = M2((n0 + 3) * q + 5), M2(1, n0 + 3, n0)
        Bk, Bn = M2(1, n0 + 1, n0), M2((n0 + 1) * q + 2)
    else:
        Am, Ak = M2(1), M2(M + 3, (M + 3) * n0 + 7, n0)
        Bk, Bn = M2(N + 1, (N + 1) * n0 + 5, n0), M2(1)
    def span(m_r, m_k, R):
        off = lambda m, i: i * m.s0 if m.n0 == 0 else (i % m.n0) * m.s0 + (i // m.n0) * m.s1
        return int(off(m_r, np.arange(R)).max() + off(m_k, np.arange(K)).max()) + 1
    sa, sb = span(Am, Ak, M), span(Bn, Bk, N)
    A = f32(rng.standard_normal(sa * batch))
    B = f32(rng.standard_normal(sb * batch))
    C = rng.standard_normal(M * N * batch)
    args = (M, N, K, batch, 0.75, 0.5, A, Am, Ak, sa, B, Bk, Bn, sb, C, M2(1), M2(M), M * N)
    ref = host_gemm(*args)
    got = run(*args, naive=mode + 8 + 16)
    bound = reduced_bound(mode, M, N, K, batch, 0.75, A, Am, Ak, sa, B, Bk, Bn, sb, M2(1), M2(M), M * N)
    err = np.abs(got - ref)
    assert np.all(err <= bound + 1e-15 * np.abs(ref)), (err / (bound + 1e-300)).max()
    assert err.max() > 0


@pytest.mark.parametrize('mode', [3, 4, 5])
@pytest.mark.parametrize('case', ['plain', 'kscale_batched', 'fp32_a'])
def test_kept_a_planes_match(mode, case):
    """A split once into a kept plane buffer (the path of a product whose A recurs, such as the
    rotated integrals of the t~ dX term) gives the same bits as splitting A within the product:
    column-major A, a k-scaled batched A with two-level maps, and A in FP32 storage."""
    rng = np.random.default_rng(hash(('kept', mode, case)) % 2**32)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    M, N, K, batch = 300, 170, 150, 1
    Am, Ak, Ab, ks, Km, store = M2(1), M2(M), 0, None, None, 0
    if case == 'kscale_batched':
        batch = 2
        Am, Ak, Ab = M2(1, 2 * M, 150), M2(150), 0
        ks, Km = rng.standard_normal(K + 3), M2(1)
    if case == 'fp32_a':
        store = 1
    A = rng.standard_normal(M * K * 2 + 2 * M * 2)
    if store:
        A = f32(A)
    B = rng.standard_normal(K * N * batch)
    C = rng.standard_normal(M * N * batch)
    args = (M, N, K, batch, 1.25, 0.5, A, Am, Ak, Ab, B, M2(1), M2(K), K * N, C, M2(1), M2(M), M * N, ks, Km)
    ref = run(*args, naive=mode + 8 * store)
    got = run(*args, naive=mode + 8 * store + 64)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    exact = host_gemm(*args)
    bound = reduced_bound(mode, M, N, K, batch, 1.25, A, Am, Ak, Ab, B, M2(1), M2(K), K * N, M2(1), M2(M), M * N,
                          ks, Km)
    assert np.all(np.abs(got - exact) <= bound + 1e-15 * np.abs(exact))


def test_kept_a_planes_need_a_split_form():
    A, B, C = np.ones(64), np.ones(64), np.zeros(64)
    with pytest.raises(_lib.UW12Error):
        run(8, 8, 8, 1, 1.0, 0.0, A, M2(1), M2(8), 0, B, M2(1), M2(8), 0, C, M2(1), M2(8), 0, naive=2 + 64)


@pytest.mark.parametrize('mode', [0, 2])
@pytest.mark.parametrize('layout', ['NN', 'TT'])
def test_k_scale_period_with_one_level_maps(mode, layout):
    """A k scale repeating with a period (20) that is not a multiple of the slab depth while A and B
    have one-level k maps (the direct Fock sum over (i, A) with weights per occupied i): the product
    runs the first 16 of every period on the unpredicated path and the rest as a second product."""
    rng = np.random.default_rng(hash(('kper', mode, layout)) % 2**32)
    M, N, K = 150, 130, 20 * 33
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    A = rng.standard_normal(M * K)
    B = rng.standard_normal(K * N)
    C = rng.standard_normal(M * N)
    ks = rng.standard_normal(20)
    args = (M, N, K, 1, 0.75, 0.5, A, Am, Ak, 0, B, Bk, Bn, 0, C, M2(1), M2(M), 0, ks, M2(1, 0, 20))
    ref = host_gemm(*args)
    got = run(*args, naive=mode)
    if mode == 0:
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    else:
        bound = fp32_bound(M, N, K, 1, 0.75, A, Am, Ak, 0, B, Bk, Bn, 0, M2(1), M2(M), 0, ks, M2(1, 0, 20))
        assert np.all(np.abs(got - ref) <= bound + 1e-15 * np.abs(ref))


@pytest.fixture
def tiles():
    """Sets the forced tile shape with tiles(bm, bn); restores the automatic choice afterwards."""
    L = _lib.load()
    yield lambda bm, bn: _lib.check(L.uw12g_test_force_tile(bm, bn))
    L.uw12g_test_force_tile(0, 0)


WIDE_CASES = [
    # name, M, N, K, batch, layout
    ('nn', 600, 520, 200, 1, 'NN'),
    ('ragged_tn', 257, 300, 129, 1, 'TN'),
    ('ragged_nt', 513, 250, 45, 1, 'NT'),
    ('batched_tt', 300, 290, 150, 2, 'TT'),
    ('small', 40, 33, 70, 3, 'NN'),
]


@pytest.mark.parametrize('name,M,N,K,batch,layout', WIDE_CASES)
@pytest.mark.parametrize('mode', [3, 5])
@pytest.mark.parametrize('extra', [0, 8 * 7, 64])
def test_wide_tiles_match(tiles, name, M, N, K, batch, layout, mode, extra):
    """The 256 x 256 tiles of the two-plane split forms (operands moved straight
    into LDS, eight waves of 128 x 64) give the bits of the 128 x 128 tiles: ragged
    edges, all layouts, batches, FP32 storage of A, B and C, and a kept A plane
    buffer; K below 512 keeps both on one k chunk."""
    rng = np.random.default_rng(hash(('wide', name, mode, extra)) % 2**32)
    f32 = lambda x: x.astype(np.float32).astype(np.float64)
    A = rng.standard_normal(M * K * batch)
    B = rng.standard_normal(K * N * batch)
    C = rng.standard_normal(M * N * batch)
    if extra == 56:
        A, B, C = f32(A), f32(B), f32(C)
    Am, Ak = (M2(1), M2(M)) if layout[0] == 'N' else (M2(K), M2(1))
    Bk, Bn = (M2(1), M2(K)) if layout[1] == 'N' else (M2(N), M2(1))
    args = (M, N, K, batch, 1.25, 0.5, A, Am, Ak, M * K, B, Bk, Bn, K * N, C, M2(1), M2(M), M * N)
    tiles(128, 128)
    ref = run(*args, naive=mode + extra)
    tiles(256, 256)
    got = run(*args, naive=mode + extra)
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    exact = host_gemm(*args)
    bound = reduced_bound(mode, M, N, K, batch, 1.25, A, Am, Ak, M * K, B, Bk, Bn, K * N, M2(1), M2(M), M * N)
    bound = bound + (2.0 ** -24 if extra == 56 else 1e-15) * np.abs(exact)
    assert np.all(np.abs(got - exact) <= bound)


@pytest.mark.parametrize('mode', [3, 5])
@pytest.mark.parametrize('bf16_path', ['planes', 'pieces'], indirect=True)
def test_wide_tiles_split_k_within_bound(tiles, mode, bf16_path):
    """256 x 256 tiles with split-K partials and several FP32 k chunks (K = 5000 on few tiles),
    whole or in pieces, within the form's rounding bound."""
    rng = np.random.default_rng(hash(('wide_k', mode, bf16_path)) % 2**32)
    M, N, K, batch = 520, 300, 5000, 1
    A = rng.standard_normal(M * K)
    B = rng.standard_normal(K * N)
    C = rng.standard_normal(M * N)
    args = (M, N, K, batch, 1.25, 0.5, A, M2(1), M2(M), 0, B, M2(N), M2(1), 0, C, M2(1), M2(M), 0)
    tiles(256, 256)
    got = run(*args, naive=mode)
    exact = host_gemm(*args)
    bound = reduced_bound(mode, M, N, K, batch, 1.25, A, M2(1), M2(M), 0, B, M2(N), M2(1), 0, M2(1), M2(M), 0)
    assert np.all(np.abs(got - exact) <= bound + 1e-15 * np.abs(exact))
    tiles(128, 128)
    ref = run(*args, naive=mode)
    assert np.abs(got - ref).max() <= 2 * bound.max()
