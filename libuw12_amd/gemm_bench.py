"""Time the FP64 strided GEMM on the shapes that dominate a c2 evaluation.

Usage (on a GPU): python -m libuw12_amd.gemm_bench

Each line gives TFLOP/s (2·M·N·K·batch per launch over the average of 3
launches after one warm-up, timed with HIP events on the library's own
stream through uw12g_time_gemm), next to torch.mm on a square FP64 matrix
as the library reference.  Operands are uniform random device tensors.
"""
import ctypes

from ._lib import Map2, load


def lin(s):
    return Map2(s, 0, 0)


def two(n0, s0, s1):
    return Map2(s0, s1, n0)


def main():
    import torch

    L = load()
    dev = torch.device('cuda:0')
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)

    def buf(n):
        return torch.rand(n, dtype=torch.float64, device=dev, generator=gen)

    def run(name, M, N, K, batch, A, Am, Ak, Ab, B, Bk, Bn, Bb, C, Cm, Cn, Cb, ks=None, Km=lin(0)):
        ms = ctypes.c_double()
        rc = L.uw12g_time_gemm(0, M, N, K, batch, 1.0, 0.0, A.data_ptr(), Am, Ak, Ab, B.data_ptr(), Bk, Bn, Bb,
                               C.data_ptr(), Cm, Cn, Cb, ks.data_ptr() if ks is not None else None, Km, 3,
                               ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(L.uw12g_last_error().decode())
        print(f'{name:44s} {2.0 * M * N * K * batch / ms.value / 1e9:6.1f} TF/s  ({ms.value:.2f} ms)', flush=True)

    # c2 shapes: n_ao 320, RI 1280 (M = 1600 rows of Phi), 80 occupied, DF 960 / 360 kept, j-block 6
    Nao, Nri, o, a, jn = 320, 1280, 80, 80, 6
    M = Nao + Nri
    MN = M * Nao
    Kt = buf(MN * o * jn)
    for DX in (360, 960):
        X2sT = buf(a * o * DX)
        Lb = buf(MN * DX)
        run(f'3e Fock L   (MN x {DX} x o*jn)', MN, DX, o * jn, 1, Kt, lin(1), lin(MN), 0, X2sT,
            two(o, DX * a, DX), lin(1), 0, Lb, lin(1), lin(MN), 0)
        del X2sT, Lb
    for DY in (960, 360):
        phit, phi, sgn = buf(M * a * DY), buf(M * o * DY), buf(DY)
        run(f'3e Fock K~  (M x N*o x {DY}, batch {jn})', M, Nao * o, DY, jn, phit, lin(1), lin(M * a), M, phi,
            lin(M * o), two(Nao, 1, M), 0, Kt, lin(1), lin(M), MN * o, ks=sgn, Km=lin(1))
        del phit, phi
    DW = 360
    phiW, Qt, R = buf(M * o * DW), buf(M * a * Nao * jn), buf(DW * Nao * jn)
    run(f'3e Fock R   ({DW} x N x M*a, batch {jn})', DW, Nao, M * a, jn, phiW, lin(M * o), lin(1), 0, Qt, lin(1),
        lin(M * a), M * a * Nao, R, lin(1), lin(DW), DW * Nao)
    del phiW, Qt, R, Kt
    n = 8192
    A, B, C = buf(n * n), buf(n * n), buf(n * n)
    run('square 8192 NN', n, n, n, 1, A, lin(1), lin(n), 0, B, lin(n), lin(1), 0, C, lin(1), lin(n), 0)
    run('square 8192 TN (k contiguous in A and B)', n, n, n, 1, A, lin(n), lin(1), 0, B, lin(1), lin(n), 0, C,
        lin(1), lin(n), 0)
    A2, B2, C2 = A.view(n, n), B.view(n, n), C.view(n, n)
    torch.mm(A2, B2, out=C2)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        torch.mm(A2, B2, out=C2)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    print(f'{"torch.mm square 8192 (library DGEMM)":44s} {2 * n ** 3 / ms / 1e9:6.1f} TF/s  ({ms:.2f} ms)', flush=True)


if __name__ == '__main__':
    main()
