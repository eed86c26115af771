"""Time the strided GEMM on the shapes that dominate a c2 evaluation.

Usage (on a GPU): python -m libuw12_amd.gemm_bench

Each line gives TFLOP/s for the FP64 matrix cores and for the FP32 form
(2·M·N·K·batch per launch over the average of 3 launches after one warm-up,
timed with HIP events on the library's own stream through uw12g_time_gemm),
next to torch.mm on a square FP64 matrix as the library reference.  Operands
are uniform random device tensors.  Shapes: n_ao 320, 1360 rotated ABS rows,
80 occupied, 960 / 360 kept DF directions, a 40-wide j-block.
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
        rates = []
        for prec in (0, 1):
            ms = ctypes.c_double()
            rc = L.uw12g_time_gemm(0, M, N, K, batch, 1.0, 0.0, A.data_ptr(), A.numel(), Am, Ak, Ab, B.data_ptr(),
                                   B.numel(), Bk, Bn, Bb, C.data_ptr(), C.numel(), Cm, Cn, Cb,
                                   ks.data_ptr() if ks is not None else None, ks.numel() if ks is not None else 0,
                                   Km, prec, 3, ctypes.byref(ms))
            if rc != 0:
                raise RuntimeError(L.uw12g_last_error().decode())
            rates.append(f'{2.0 * M * N * K * batch / ms.value / 1e9:6.1f} TF/s ({ms.value:7.2f} ms)')
        print(f'{name:40s} FP64 {rates[0]}   FP32 {rates[1]}', flush=True)

    Nao, r, o, a, jn = 320, 1360, 80, 80, 40
    rN = r * Nao
    Kt = buf(rN * o * jn)
    for DX in (360, 960):
        X2sT = buf(a * o * DX)
        Lb = buf(rN * DX)
        run(f'L    (rN x {DX} x o*jn)', rN, DX, o * jn, 1, Kt, lin(1), lin(rN), 0, X2sT,
            two(o, DX * a, DX), lin(1), 0, Lb, lin(1), lin(rN), 0)
        del X2sT, Lb
    for DY in (960, 360):
        phit, phi = buf(r * a * DY), buf(Nao * o * DY)
        run(f'K~   (r x N*o x {DY}, batch {jn})', r, Nao * o, DY, jn, phit, lin(1), lin(r * a), r, phi,
            lin(Nao * o), lin(1), 0, Kt, lin(1), lin(r), rN * o)
        del phit, phi
    DW = 360
    psiW, Rb = buf(r * o * DW), buf(DW * Nao * jn)
    run(f'R    ({DW} x N x r*a, batch {jn}; TN)', DW, Nao, r * a, jn, psiW, lin(r * o), lin(1), 0, Kt,
        two(r, 1, rN), lin(r), rN * o, Rb, lin(1), lin(DW), DW * Nao)
    del psiW, Rb, Kt
    DX = 960
    Gt, Lb, F = buf(rN * DX), buf(rN * DX), buf(Nao * Nao)
    run(f'F += Gt^T L (N x N x r*{DX}; TN)', Nao, Nao, r * DX, 1, Gt, lin(r), two(r, 1, rN), 0, Lb,
        two(r, 1, rN), lin(r), 0, F, lin(1), lin(Nao), 0)
    Co, Psi = buf(Nao * o), buf(r * o * DX)
    run(f'Psi  (r x o x N, batch {DX}; B k-contig)', r, o, Nao, DX, Gt, lin(1), lin(r), rN, Co, lin(1), lin(Nao), 0,
        Psi, lin(1), lin(r), r * o)
    del Gt, Lb, F, Co, Psi
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
    print(f'{"torch.mm square 8192 (library DGEMM)":40s} FP64 {2 * n ** 3 / ms / 1e9:6.1f} TF/s ({ms:7.2f} ms)',
          flush=True)


if __name__ == '__main__':
    main()
