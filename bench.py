#!/usr/bin/env python
"""UW12 energy + Fock evaluations/s on one MI355X.

One step = one full uw12_fock evaluation (2-, 3- and 4-electron energy terms
and the Fock matrix) of a synthetic s-Gaussian cluster whose density-fitted
integrals are already resident in HBM (set up once, as an SCF caller would).

The timed form is the mixed one by default (FP32 matrix cores for the large
intermediate tensors of the Fock terms, FP64 for the orbital transforms, the
energies and every accumulation into the Fock matrix).  The plain FP64 path is
timed beside it on the same inputs, and the mixed result is gated against it:
`accuracy_gate` reports the largest energy-term and Fock deviations and
whether both are within 1e-6.  `--form fp64` times the plain path instead.

Prints one JSON line (see the keys below).  --gpus > 1 is rejected: this
build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'UW12 E+Fock evals/s, mixed vs FP64 on 1 MI355X at ≤1e-6 Eh; MFMA % peak'
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix peak (AMD published)
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix peak (AMD published)
PEAK_HBM_GBS = 8000.0
GATE = 1e-6  # energy (Eh, per term and total) and Fock max-abs bar of the reduced-precision forms
TERM_KEYS = ('e2_direct', 'e2_indirect', 'e3_direct', 'e3_indirect', 'e4_direct', 'e4_indirect', 'total')


def flop_model(N, Nd, Nr, occ, ss):
    """Reference contraction order restated as GEMMs (per evaluation)."""
    M = N + Nr
    t = 0.0
    for o in occ:
        a = o
        t += 2 * (2 * N * N * o * Nd + 2 * N * o * Nd * Nd)
        t += 2 * (2 * Nr * N * a * Nd + 2 * Nr * a * Nd * Nd)
        t += 2 * (2 * a * N * o * Nd)
        t += 2 * N * N * a * Nd + 2 * N * a * Nd * Nd + (2 * a * N * a * Nd if ss else 0)
        t += 2 * N * N * a + N * N * Nd + 2 * Nd * Nd + N * N * Nd + 2 * Nd * Nd
        if ss:
            t += 2 * a * a * Nd + 2 * N * N * a * Nd
        t += 2 * a * o * Nd * Nd + 2 * Nd * Nd
        t += 2 * N * o * Nd * Nd + 2 * N * N * o * Nd
        if ss:
            t += 2 * 2 * (a * o) ** 2 * Nd + 2 * (a * o) ** 2
            t += 2 * 2 * (N * o) * (a * o) * Nd + 2 * N * N * a * a * o
        t += 2 * a * M * M * Nd + 2 * a * M * Nd * Nd
        t += 2 * N * o * Nd * Nd + 2 * N * N * o * Nd
        t += (N * N + 2 * N * Nr) * Nd * Nd + 2 * N * M * M * Nd + 2 * N * N * M * Nd
        if ss:
            t += a * a * (2 * M * Nd * Nd + 2 * o * Nd * Nd + 2 * Nd * Nd)
            t += 2 * a * o * (2 * N * Nd * (2 * Nr + 1.5 * N) + 2 * N * M * M + 2 * N * N * M)
            t += a * a * (2 * N * Nd * (2 * Nr + 2 * N) + 2 * N * M * M + 2 * N * N * M)
    return t


def roofline_from_profile(records):
    """records: list of (stage|kernel, ms, flops, bytes) over the timed steps."""
    agg = {}
    for name, ms, fl, by in records:
        kernel = name.split('|', 1)[1]
        a = agg.setdefault(kernel, [0.0, 0.0, 0.0, 0])
        a[0] += ms
        a[1] += fl
        a[2] += by
        a[3] += 1
    dom = max(agg, key=lambda k: agg[k][0])
    ms, fl, by, n = agg[dom]
    total_ms = sum(v[0] for v in agg.values())
    table = {k: dict(ms_total=round(v[0], 3), launches=v[3], avg_us=round(1e3 * v[0] / v[3], 2),
                     tflops=round(v[1] / (v[0] * 1e-3) / 1e12, 2) if v[1] and v[0] > 0 else None)
             for k, v in sorted(agg.items(), key=lambda kv: -kv[1][0])}
    stages = {}
    for name, ms_, fl_, by_ in records:
        st = stages.setdefault(name, [0.0, 0.0, 0])
        st[0] += ms_
        st[1] += fl_
        st[2] += 1
    table['_by_stage'] = {k: dict(ms_total=round(v[0], 3), launches=v[2],
                                  tflops=round(v[1] / (v[0] * 1e-3) / 1e12, 2) if v[1] and v[0] > 0 else None)
                          for k, v in sorted(stages.items(), key=lambda kv: -kv[1][0])}
    if fl > 0:
        achieved = fl / (ms * 1e-3) / 1e12
        # gemm_f64_kernel<BM, BN, AKC, BKC, KS, PREC>: PREC 1 runs on the FP32 matrix cores
        peak = PEAK_FP32_TFLOPS if dom.startswith('gemm_f64_kernel<') and dom.rstrip('>').endswith(',1') \
            else PEAK_FP64_TFLOPS
        roof = dict(bound='mfma', kernel=dom, achieved=round(achieved, 3), peak=peak, unit='TFLOP/s',
                    frac=round(achieved / peak, 4), traffic=None,
                    launches=n, avg_launch_us=round(1e3 * ms / n, 2),
                    algorithmic_flops_per_launch=fl / n, share_of_gpu_time=round(ms / total_ms, 3))
    else:
        achieved = by / (ms * 1e-3) / 1e9
        roof = dict(bound='hbm', kernel=dom, achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit='GB/s',
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=None, launches=n,
                    avg_launch_us=round(1e3 * ms / n, 2), share_of_gpu_time=round(ms / total_ms, 3))
    return roof, table


def traffic_profile(config, form):
    return f'profiles/gemm_f64_{config}_{form}_traffic.md'


def measured_traffic(kernel, config, form):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (profiles/gemm_f64_<config>_<form>_traffic.md), or None when there is none."""
    path = os.path.join(ROOT, traffic_profile(config, form))
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == 'traffic_bytes_per_launch' and parts[1] == kernel:
            return float(parts[2])
    return None


def cpu_baseline(system, k0, k1, pair_samples=(1, 4), full=False, gpu_result=None):
    """Time the CPU oracle (reference contraction order, FP64, OpenMP + MKL) on the
    same inputs, with the same-spin pair loops on two uniform samples of the outer
    pair index (pair_samples indices each), extrapolated linearly.  With
    full=True the whole evaluation runs once (no sampling) and its energy pieces
    and Fock are compared with each GPU form in gpu_result = {form: (terms, Fock list)}."""
    from oracle import cpu
    threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))
    t = time.perf_counter()
    sysc = cpu.System(system['ints'], system['S'])
    print(f'[bench] cpu oracle setup {time.perf_counter() - t:.1f}s', file=sys.stderr, flush=True)
    C, occ, act = system['orbitals'], system['occ'], system['n_active']
    if full:
        t = time.perf_counter()
        cpu.set_progress(True)
        try:
            eo, Fo = sysc.evaluate(C, occ, act, k0, k1, fock=True)
        finally:
            cpu.set_progress(False)
        t_full = time.perf_counter() - t
        print(f'[bench] cpu oracle full evaluation {t_full:.1f}s', file=sys.stderr, flush=True)
        out = dict(value=round(1.0 / t_full, 6), unit='evals/s', cores=threads, kind='port',
                   sample='one full oracle evaluate on the same inputs (no sampling); setup (metric eig, J3) excluded',
                   seconds_per_eval=round(t_full, 2), blas=cpu.blas_kind())
        if gpu_result:
            out['parity'] = {'oracle_energy': eo['total'], 'fock_max_abs': max(float(np.abs(b).max()) for b in Fo)}
            for form, (eg, Fg) in gpu_result.items():
                out['parity'][form] = {'term_diff_eh': {k: eg[k] - eo[k] for k in TERM_KEYS},
                                       'fock_max_abs_diff': max(float(np.abs(a - b).max()) for a, b in zip(Fg, Fo))}
        return out
    # The same-spin pair loops run on every s-th outer pair index; timing the whole
    # evaluation at two strides s1 > s2 separates the fixed part T0 from the cost p
    # of one outer index (t(s) = T0 + p ceil(a / s)); the full evaluation takes
    # T0 + p a.
    a = max(act)
    s1, s2 = (max(1, -(-a // n)) for n in pair_samples)
    if k1 == 0.0 or s2 >= s1:
        t = time.perf_counter()
        sysc.evaluate(C, occ, act, k0, k1, fock=True)
        t_full = time.perf_counter() - t
        sample = 'one full oracle evaluate on the same inputs; setup (metric eig, J3) excluded'
    else:
        times = {}
        try:
            # a short first run also carries one-off start-up (threads, BLAS
            # dispatch): time it again
            for stride in (s1, s1, s2):
                if stride in times and times[stride] > 10.0:
                    continue
                cpu.set_pair_stride(stride)
                t = time.perf_counter()
                sysc.evaluate(C, occ, act, k0, k1, fock=True)
                times[stride] = time.perf_counter() - t
                print(f'[bench] cpu oracle evaluation with pair stride {stride}: {times[stride]:.1f}s',
                      file=sys.stderr, flush=True)
        finally:
            cpu.set_pair_stride(1)
        # outer pair indices actually visited at each stride
        n1, n2 = -(-a // s1), -(-a // s2)
        per_index = max(times[s2] - times[s1], 0.0) / max(n2 - n1, 1)
        fixed = max(times[s1] - n1 * per_index, 0.0)
        t_full = fixed + a * per_index
        sample = (f'oracle evaluate on the same inputs with the same-spin pair loops on every {s1}-th and every '
                  f'{s2}-th outer pair index ({times[s1]:.1f}s, {times[s2]:.1f}s), extrapolated linearly to all '
                  f'{a}; setup (metric eig, J3) excluded')
    return dict(value=round(1.0 / t_full, 6), unit='evals/s', cores=threads, kind='port', sample=sample,
                seconds_per_eval=round(t_full, 2), blas=cpu.blas_kind())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='c2', choices=['c1', 'c2', 'c3', 'c4'])
    ap.add_argument('--kappa', default='1.0,0.5')
    ap.add_argument('--form', default='mixed', choices=['mixed', 'fp64', 'fp32'],
                    help='precision form of the timed steps (the plain FP64 path is run beside it)')
    ap.add_argument('--side-reps', type=int, default=3,
                    help='timed evaluations of each side form (after one untimed one when more than 1)')
    ap.add_argument('--side-forms', default='auto', choices=['auto', 'gate', 'skip'],
                    help='gate: only the FP64 evaluation the accuracy gate needs; skip: no FP64 / FP32 / mixed '
                         'evaluations beside the timed ones (no gate), for profiler runs that should see only the '
                         'timed form')
    ap.add_argument('--cpu-baseline', default='auto', choices=['auto', 'skip', 'full'],
                    help='auto: sampled oracle timing; full: one complete oracle evaluation, with parity report')
    ap.add_argument('--profile-table', action='store_true', help='also print the per-kernel table to stderr')
    ap.add_argument('--verbose', action='store_true',
                    help='library progress lines on stderr during each evaluation (long configurations)')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='only copy 1 GiB device to device 3 times (reference for HBM byte counters) and exit')
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'calibration': 'device copy', 'bytes_read': 1 << 30, 'bytes_written': 1 << 30,
                          'launches': 3}), flush=True)
        return
    if args.gpus != 1:
        print(f'bench.py measures one GPU; --gpus {args.gpus} is not supported by this build '
              '(the multi-GPU design shards the DF index and is documented, not run).', file=sys.stderr)
        sys.exit(2)
    k0, k1 = (float(x) for x in args.kappa.split(','))

    import libuw12_amd as u
    from libuw12_amd import synth

    # the larger configurations stream their integrals per DF shell (ThreeIndexFn form)
    # instead of holding X3 / X3_ri on the host
    stream = args.config in ('c3', 'c4')
    t = time.perf_counter()
    system = synth.make_system(args.config, with_integrals=not stream)
    t_gen = time.perf_counter() - t
    N, Nd, Nr = system['N'], system['Nd'], system['Nr']
    occ_counts = [len(x) for x in system['occ']]
    print(f'[bench] {args.config}: n_ao={N} n_df={Nd} n_ri={Nr} n_occ={occ_counts} generated in {t_gen:.1f}s',
          file=sys.stderr, flush=True)
    ctx = u.Context(0)
    if args.verbose:
        ctx.set_verbosity(1)
    t = time.perf_counter()
    if stream:
        handles = []
        for kind in ('W', 'V', 'WV'):
            X2, sizes, fn, fn_ri = synth.shell_integrals(system, kind)
            handles.append(u.setup_base_integrals_fn(X2, fn, fn_ri, sizes, N, Nr if fn_ri else 0, context=ctx))
            print(f'[bench] {kind} integrals streamed ({len(sizes)} DF shells)', file=sys.stderr, flush=True)
        W, V, WV = handles
    else:
        W = u.setup_base_integrals(*system['ints']['W'], n_ao=N, context=ctx)
        V = u.setup_base_integrals(*system['ints']['V'], n_ao=N, context=ctx)
        WV = u.setup_base_integrals(system['ints']['WV'][0], system['ints']['WV'][1], None, n_ao=N, context=ctx)
    P = u.setup_abs_projectors(system['S'], N, Nr, context=ctx)
    t_setup = time.perf_counter() - t
    print(f'[bench] device setup {t_setup:.1f}s (kept DF directions W={W.n_kept} V={V.n_kept} WV={WV.n_kept})',
          file=sys.stderr, flush=True)
    indirect = abs(k1) > 4 * np.finfo(float).eps
    prec = {'fp64': u.FP64, 'fp32': u.FP32, 'mixed': u.MIXED}[args.form]

    def evaluation(precision):
        return u.form_fock(W, V, WV, P, system['orbitals'], system['occ'], system['n_active'],
                           indirect_term=indirect, calculate_fock=True, scale_opp_spin=k0, scale_same_spin=k1,
                           precision=precision)

    for i in range(args.warmup):
        evaluation(prec)
        print(f'[bench] warmup {i + 1}/{args.warmup} done', file=sys.stderr, flush=True)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    ctx.set_profiling(True)
    records = []
    sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        terms, F = evaluation(prec)
        records.extend(ctx.profile())
        print(f'[bench] step {i + 1}/{args.steps} done', file=sys.stderr, flush=True)
    sync()
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(False)
    value = args.steps / elapsed
    roof, table = roofline_from_profile(records)
    roof['traffic'] = measured_traffic(roof['kernel'], args.config, args.form)
    if roof['traffic'] is not None:
        roof['traffic_source'] = f'{traffic_profile(args.config, args.form)} (rocprofv3 FETCH_SIZE + WRITE_SIZE)'
    gpu_flops = sum(r[2] for r in records) / args.steps
    model = flop_model(N, Nd, Nr, occ_counts, indirect)
    if args.profile_table:
        print(json.dumps(table, indent=1), file=sys.stderr)

    def deviation(tr, Fr, t_ref, F_ref):
        return {'term_error_eh': {k: tr[k] - t_ref[k] for k in TERM_KEYS},
                'fock_max_abs_error': max(float(np.abs(a - b).max()) for a, b in zip(Fr, F_ref))}

    def timed_form(precision, name, reps=max(1, args.side_reps)):
        if reps > 1:
            evaluation(precision)
        sync()
        t = time.perf_counter()
        for _ in range(reps):
            tr, Fr = evaluation(precision)
        sync()
        tp = (time.perf_counter() - t) / reps
        print(f'[bench] {name} form: {1e3 * tp:.1f} ms per evaluation', file=sys.stderr, flush=True)
        return tr, Fr, tp

    # the plain FP64 path on the same inputs: reference for the reduced-precision forms
    side = args.side_forms in ('auto', 'gate')
    if args.form == 'fp64' or not side:
        terms64, F64, t64 = terms, F, elapsed / args.steps
    else:
        terms64, F64, t64 = timed_form(u.FP64, 'FP64')
    fp64 = {'evals_per_s': round(1.0 / t64, 6), 'ms_per_eval': round(1e3 * t64, 3),
            'energy': terms64['total']} if side or args.form == 'fp64' else None
    gate = None
    if args.form != 'fp64' and side:
        dev = deviation(terms, F, terms64, F64)
        worst_e = max(abs(v) for v in dev['term_error_eh'].values())
        gate = dict(dev, bar=GATE, max_energy_error_eh=worst_e,
                    met=bool(worst_e <= GATE and dev['fock_max_abs_error'] <= GATE),
                    speedup_vs_fp64=round(t64 * value, 3))
        print(f'[bench] {args.form} vs FP64: max energy error {worst_e:.3e} Eh, Fock max-abs error '
              f'{dev["fock_max_abs_error"]:.3e}, gate {"met" if gate["met"] else "NOT met"}',
              file=sys.stderr, flush=True)
    others = {}
    for name, precision in (('fp32', u.FP32), ('mixed', u.MIXED)):
        if name == args.form or args.side_forms != 'auto':
            continue
        tr, Fr, tp = timed_form(precision, name)
        others[f'{name}_form'] = dict(deviation(tr, Fr, terms64, F64), evals_per_s=round(1.0 / tp, 6),
                                      speedup_vs_fp64=round(t64 / tp, 3))
    if args.side_forms == 'auto':
        # the energy-only call (uw12_energy, libuw12.hpp:129-140) in the timed form: no Fock, the
        # same-spin 3e energy from its own FP64 products instead of the Fock's L
        def energy_only():
            return u.form_fock(W, V, WV, P, system['orbitals'], system['occ'], system['n_active'],
                               indirect_term=indirect, calculate_fock=False, scale_opp_spin=k0,
                               scale_same_spin=k1, precision=prec)[0]
        reps = max(1, args.side_reps)
        energy_only()
        sync()
        t = time.perf_counter()
        for _ in range(reps):
            te = energy_only()
        sync()
        tp = (time.perf_counter() - t) / reps
        print(f'[bench] energy-only ({args.form}): {1e3 * tp:.1f} ms per evaluation', file=sys.stderr, flush=True)
        others['energy_only'] = {'form': args.form, 'evals_per_s': round(1.0 / tp, 6), 'ms_per_eval': round(1e3 * tp, 3),
                                 'term_error_eh': {k: te[k] - terms64[k] for k in TERM_KEYS}}
    cpu = None
    if args.cpu_baseline in ('auto', 'full') and not stream:
        print('[bench] timing CPU baseline (oracle)', file=sys.stderr, flush=True)
        cpu = cpu_baseline(system, k0, k1, full=args.cpu_baseline == 'full',
                           gpu_result={'fp64': (terms64, F64), args.form: (terms, F)} if side
                           else {args.form: (terms, F)})
    form_text = {'fp64': 'plain FP64',
                 'fp32': 'FP32 form (every contraction on FP32 matrix cores, FP64 storage and reductions)',
                 'mixed': 'mixed form (FP32 matrix cores for the Fock intermediates K~, L, R, Xi, G, X_i/Y_i; FP64 '
                          'for the orbital transforms, energies and Fock accumulation), gated at <=1e-6 Eh per '
                          'term and <=1e-6 Fock max-abs against the plain FP64 path of the same run'}[args.form]
    out = {
        'metric': METRIC,
        'value': round(value, 6),
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * elapsed / args.steps, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'fp64': 'f64', 'fp32': 'f32', 'mixed': 'f32+f64'}[args.form],
        'data': 'synthetic (seeded s-Gaussian cluster, analytic Boys-F0 integrals)',
        'config': {'workload': f'{args.config}: UW12 energy+Fock, n_ao={N} n_df={Nd} n_ri={Nr} n_occ={occ_counts}, '
                               f'kappa=({k0},{k1}), {form_text}',
                   'form': args.form, 'kappa': [k0, k1], 'parallelism': 'single-gpu'},
        'roofline': roof,
        'cpu_baseline': cpu,
        'accuracy_gate': gate,
        'fp64_form': fp64,
        'energy': terms['total'],
        'gpu_tflops_executed': round(gpu_flops * value / 1e12, 3),
        'model_flops_per_eval': model,
        'executed_flops_per_eval': gpu_flops,
        'setup_s': round(t_setup, 2),
    }
    out.update(others)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
