#!/usr/bin/env python
"""UW12 energy + Fock evaluations/s on one MI355X.

One step = one full uw12_fock evaluation (2-, 3- and 4-electron energy terms
and the Fock matrix) of a synthetic s-Gaussian cluster whose density-fitted
integrals are already resident in HBM (set up once, as an SCF caller would).

The timed form is by default the mixed form with FP16x3 split products
(`--form fp16x3`): the large intermediate tensors of the Fock terms run on the
FP16 matrix cores as three products of two IEEE half planes per operand (rows
scaled by powers of two, FP32 sums), while the orbital transforms, the
energies and every accumulation into the Fock matrix stay FP64.  The plain
FP64 path is timed beside it on the same inputs, and the timed result is gated
against it: `accuracy_gate` reports the largest energy-term and Fock
deviations and whether both are within 1e-6.  `--form fp64` times the plain
path instead; `--form mixed` the mixed form with those intermediates on the
FP32 matrix cores; `--form bf16x3` / `bf16x6` with BF16 split products.  The
side forms (FP32, mixed, BF16x3, BF16x6, FP16x3) are timed and compared with
FP64 after the timed steps, and so is the differentiation form (forward then
adjoint, `adjoint_form`).

Workloads (--config): c1, c2 (host integral arrays, the default timed
workload c2), c3 / c4 (integrals generated on the device and stored), c5
(n_ao = 1500: integrals generated on the device and not stored -- every
evaluation regenerates its three-index blocks).  One c5 evaluation takes about
half a minute, so the default run times c2 for the requested steps and then
measures c5 at kappa = (1, 0) once per form (`c5_form`: the timed form and
FP64, the accuracy gate, its own roofline); `--c5-side skip` leaves that out.  Before
that, `c3_forms` (`--c3-side skip` leaves it out) gives one evaluation per
precision form at n_ao = 800 (unrestricted, kappa = (1, 0)): time and per-term
and Fock deviation from FP64.

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
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak (AMD published)
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
        # gemm_f64_kernel<BM, BN, AKC, BKC, KS, PREC>: PREC 1 runs on the FP32 matrix cores.
        # gemm_bf16x_kernel<..., planes>, gemm_planes_kernel<BM, BN, planes, half> and
        # gemm_planes_wide_kernel<half> (two planes): 3 (two planes) or 6 (three planes) BF16 / FP16
        # products per multiply-add, so the peak of the
        # algorithmic FLOP rate is the BF16 (= FP16) dense peak over that count.
        args_ = dom.rstrip('>').split('<', 1)[-1].split(',')
        last = args_[-1]
        if dom.startswith('gemm_bf16x_kernel<'):
            peak = round(PEAK_BF16_TFLOPS / (3 if last == '2' else 6), 1)
        elif dom.startswith('gemm_planes_kernel<'):
            peak = round(PEAK_BF16_TFLOPS / (3 if args_[2] == '2' else 6), 1)
        elif dom.startswith('gemm_planes_wide_kernel<'):  # two planes only
            peak = round(PEAK_BF16_TFLOPS / 3, 1)
        else:
            peak = PEAK_FP32_TFLOPS if dom.startswith('gemm_f64_kernel<') and last == '1' else PEAK_FP64_TFLOPS
        roof = dict(bound='mfma', kernel=dom, achieved=round(achieved, 3), peak=peak, unit='TFLOP/s',
                    frac=round(achieved / peak, 4), traffic=None,
                    launches=n, avg_launch_us=round(1e3 * ms / n, 2),
                    algorithmic_flops_per_launch=fl / n, share_of_gpu_time=round(ms / total_ms, 3))
        if by > 0:
            # the library's count of each launch's compulsory HBM bytes (operands once, output, partials)
            roof['compulsory_bytes_per_launch'] = by / n
    else:
        achieved = by / (ms * 1e-3) / 1e9
        roof = dict(bound='hbm', kernel=dom, achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit='GB/s',
                    frac=round(achieved / PEAK_HBM_GBS, 4), traffic=None, launches=n,
                    avg_launch_us=round(1e3 * ms / n, 2), share_of_gpu_time=round(ms / total_ms, 3))
    return roof, table


def traffic_profile(config, form):
    """The committed traffic summary of this workload and form (profiles/<kernel>_<config>_<form>_traffic.md)."""
    import glob
    hits = sorted(glob.glob(os.path.join(ROOT, 'profiles', f'*_{config}_{form}_traffic.md')))
    return os.path.relpath(hits[0], ROOT) if hits else f'profiles/gemm_f64_{config}_{form}_traffic.md'


def measured_traffic(kernel, config, form):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (profiles/<kernel>_<config>_<form>_traffic.md), or None when there is none."""
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


CONFIG_DEFAULTS = {
    # name: (kappa, integrals)  host: X3 / X3_ri arrays from the host generator;
    # device: generated on the device and stored; streamed: generated, not stored
    'c1': ((1.0, 0.5), 'host'),
    'c2': ((1.0, 0.5), 'host'),
    'c3': ((1.0, 0.0), 'device'),
    'c4': ((1.0, 0.0), 'device'),
    'c5': ((1.0, 0.0), 'streamed'),
}


def setup_config(name, integrals, ctx):
    import libuw12_amd as u
    from libuw12_amd import synth
    t = time.perf_counter()
    system = synth.make_system(name, with_integrals=integrals == 'host')
    t_gen = time.perf_counter() - t
    N, Nd, Nr = system['N'], system['Nd'], system['Nr']
    print(f'[bench] {name}: n_ao={N} n_df={Nd} n_ri={Nr} n_occ={[len(x) for x in system["occ"]]} inputs in '
          f'{t_gen:.1f}s', file=sys.stderr, flush=True)
    t = time.perf_counter()
    if integrals == 'host':
        W = u.setup_base_integrals(*system['ints']['W'], n_ao=N, context=ctx)
        V = u.setup_base_integrals(*system['ints']['V'], n_ao=N, context=ctx)
        WV = u.setup_base_integrals(system['ints']['WV'][0], system['ints']['WV'][1], None, n_ao=N, context=ctx)
    else:
        W, V, WV = synth.device_integrals(system, ctx, stored=integrals == 'device')
    P = u.setup_abs_projectors(system['S'], N, Nr, context=ctx)
    t_setup = time.perf_counter() - t
    print(f'[bench] {name}: device setup {t_setup:.1f}s ({integrals} integrals; kept DF directions W={W.n_kept} '
          f'V={V.n_kept} WV={WV.n_kept})', file=sys.stderr, flush=True)
    return system, (W, V, WV, P), t_setup


def deviation(tr, Fr, t_ref, F_ref):
    return {'term_error_eh': {k: tr[k] - t_ref[k] for k in TERM_KEYS},
            'fock_max_abs_error': max(float(np.abs(a - b).max()) for a, b in zip(Fr, F_ref))}


def measure(name, system, handles, ctx, kappa, form, steps, warmup, side_forms, side_reps, energy_only, label):
    """Time `steps` evaluations in `form` after `warmup` untimed ones; then the side forms.
    Returns the measurement dict (value = evaluations/s over the timed steps)."""
    import libuw12_amd as u
    k0, k1 = kappa
    W, V, WV, P = handles
    N, Nd, Nr = system['N'], system['Nd'], system['Nr']
    occ_counts = [len(x) for x in system['occ']]
    indirect = abs(k1) > 4 * np.finfo(float).eps
    streamed = not getattr(W, 'stored', True)
    prec = PRECISION(u)[form]

    def evaluation(precision, fock=True):
        return u.form_fock(W, V, WV, P, system['orbitals'], system['occ'], system['n_active'],
                           indirect_term=indirect, calculate_fock=fock, scale_opp_spin=k0, scale_same_spin=k1,
                           precision=precision)

    for i in range(warmup):
        evaluation(prec)
        print(f'[bench] {label}: warmup {i + 1}/{warmup} done', file=sys.stderr, flush=True)
    # each evaluation returns its energies and Fock to the host, so the library's stream has drained
    # when form_fock returns: the wall clock brackets exactly the timed evaluations
    ctx.set_profiling(True)
    records = []
    t0 = time.perf_counter()
    for i in range(steps):
        terms, F = evaluation(prec)
        records.extend(ctx.profile())
        print(f'[bench] {label}: step {i + 1}/{steps} done', file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(False)
    value = steps / elapsed
    roof, table = roofline_from_profile(records)
    gpu_flops = sum(r[2] for r in records) / steps
    model = flop_model(N, Nd, Nr, occ_counts, indirect)

    def timed_form(precision, fname, reps=max(1, side_reps)):
        if reps > 1:
            evaluation(precision)
        t = time.perf_counter()
        for _ in range(reps):
            tr, Fr = evaluation(precision)
        tp = (time.perf_counter() - t) / reps
        print(f'[bench] {label}: {fname} form {1e3 * tp:.1f} ms per evaluation', file=sys.stderr, flush=True)
        return tr, Fr, tp

    side = side_forms in ('auto', 'gate')
    if form == 'fp64' or not side:
        terms64, F64, t64 = terms, F, elapsed / steps
    else:
        terms64, F64, t64 = timed_form(u.FP64, 'FP64')
    fp64 = {'evals_per_s': round(1.0 / t64, 6), 'ms_per_eval': round(1e3 * t64, 3),
            'energy': terms64['total']} if side or form == 'fp64' else None
    gate = None
    if form != 'fp64' and side:
        dev = deviation(terms, F, terms64, F64)
        worst_e = max(abs(v) for v in dev['term_error_eh'].values())
        gate = dict(dev, bar=GATE, max_energy_error_eh=worst_e,
                    met=bool(worst_e <= GATE and dev['fock_max_abs_error'] <= GATE),
                    speedup_vs_fp64=round(t64 * value, 3))
        print(f'[bench] {label}: {form} vs FP64: max energy error {worst_e:.3e} Eh, Fock max-abs error '
              f'{dev["fock_max_abs_error"]:.3e}, gate {"met" if gate["met"] else "NOT met"}',
              file=sys.stderr, flush=True)
    others = {}
    for fname, precision in PRECISION(u).items():
        if fname in (form, 'fp64') or side_forms != 'auto' or (streamed and fname == 'fp32'):
            continue
        tr, Fr, tp = timed_form(precision, fname)
        dev = deviation(tr, Fr, terms64, F64)
        worst_e = max(abs(v) for v in dev['term_error_eh'].values())
        others[f'{fname}_form'] = dict(dev, evals_per_s=round(1.0 / tp, 6), speedup_vs_fp64=round(t64 / tp, 3),
                                       meets_gate=bool(worst_e <= GATE and dev['fock_max_abs_error'] <= GATE))
    if side_forms == 'auto' and not streamed:
        # differentiation form: forward (energies, intermediates kept or marked for
        # recomputation) then the adjoint Fock, in the timed form and in FP64
        def fwd_adj(precision, reps):
            kw = dict(indirect_term=indirect, scale_opp_spin=k0, scale_same_spin=k1, precision=precision)
            args = (W, V, WV, P, system['orbitals'], system['occ'], system['n_active'])
            u.adjoint(u.forward(*args, **kw)[1])
            t = time.perf_counter()
            for _ in range(reps):
                tr, tape = u.forward(*args, **kw)
                Fa = u.adjoint(tape)
            return tr, Fa, (time.perf_counter() - t) / reps
        reps = max(1, side_reps)
        ta64, Fa64, tp64 = fwd_adj(u.FP64, reps)
        ta, Fa, tp = fwd_adj(prec, reps)
        print(f'[bench] {label}: forward+adjoint {1e3 * tp:.1f} ms ({form}), {1e3 * tp64:.1f} ms (FP64)',
              file=sys.stderr, flush=True)
        dev = deviation(ta, Fa, terms64, F64)
        others['adjoint_form'] = dict(
            dev, form=form, evals_per_s=round(1.0 / tp, 6), ms_per_eval=round(1e3 * tp, 3),
            fp64_ms_per_eval=round(1e3 * tp64, 3), speedup_vs_fp64=round(tp64 / tp, 3),
            fp64_adjoint_vs_fused_fock_max_abs=max(float(np.abs(a - b).max()) for a, b in zip(Fa64, F64)))
    if energy_only and side_forms == 'auto':
        # the energy-only call (uw12_energy, libuw12.hpp:129-140) in the timed form
        reps = max(1, side_reps)
        evaluation(prec, fock=False)
        t = time.perf_counter()
        for _ in range(reps):
            te = evaluation(prec, fock=False)[0]
        tp = (time.perf_counter() - t) / reps
        print(f'[bench] {label}: energy-only ({form}) {1e3 * tp:.1f} ms per evaluation', file=sys.stderr, flush=True)
        others['energy_only'] = {'form': form, 'evals_per_s': round(1.0 / tp, 6), 'ms_per_eval': round(1e3 * tp, 3),
                                 'term_error_eh': {k: te[k] - terms64[k] for k in TERM_KEYS}}
    return dict(value=value, elapsed=elapsed, terms=terms, F=F, terms64=terms64, F64=F64, roofline=roof,
                table=table, gate=gate, fp64=fp64, others=others, gpu_flops=gpu_flops, model=model,
                occ_counts=occ_counts)


def PRECISION(u):
    return {'fp64': u.FP64, 'fp32': u.FP32, 'mixed': u.MIXED, 'bf16x3': u.MIXED_BF16X3, 'bf16x6': u.MIXED_BF16X6,
            'fp16x3': u.MIXED_FP16X3}


_GATE_TEXT = 'gated at <=1e-6 Eh per term and <=1e-6 Fock max-abs against the plain FP64 path of the same run'
FORM_TEXT = {'fp64': 'plain FP64',
             'fp32': 'FP32 form (every contraction on FP32 matrix cores, FP64 storage and reductions)',
             'mixed': 'mixed form (FP32 matrix cores for the Fock intermediates K~, L, R, Xi, G, X_i/Y_i; FP64 '
                      'for the orbital transforms, the integral generation products, energies and Fock '
                      'accumulation), ' + _GATE_TEXT,
             'bf16x3': 'mixed form with the Fock intermediates as BF16x3 split products (two bfloat16 planes per '
                       'operand, three products, FP32 sums) on BF16 matrix cores, ' + _GATE_TEXT,
             'bf16x6': 'mixed form with the Fock intermediates as BF16x6 split products (three bfloat16 planes per '
                       'operand, six products, FP32 sums) on BF16 matrix cores, ' + _GATE_TEXT,
             'fp16x3': 'mixed form with the Fock intermediates as FP16x3 split products (operand rows scaled by '
                       'powers of two, two IEEE half planes per operand, three products, FP32 sums) on FP16 matrix '
                       'cores, ' + _GATE_TEXT}
INTEGRALS_TEXT = {'host': 'host integral arrays copied to HBM',
                  'device': 'integrals generated on the device, stored in HBM',
                  'streamed': 'integrals generated on the device, not stored: every evaluation regenerates its '
                              'three-index blocks (the rotated W tensor is kept between evaluations)'}


def c3_forms(ctx):
    """One evaluation of c3 (n_ao = 800, unrestricted, kappa = (1, 0), device-generated stored
    integrals) per precision form after one untimed FP64 one: ms per evaluation and the
    deviation of each reduced form from FP64."""
    import libuw12_amd as u
    kappa, integrals = CONFIG_DEFAULTS['c3']
    system, (W, V, WV, P), t_setup = setup_config('c3', integrals, ctx)

    def evaluation(precision):
        t = time.perf_counter()
        e, F = u.form_fock(W, V, WV, P, system['orbitals'], system['occ'], system['n_active'],
                           indirect_term=False, scale_opp_spin=kappa[0], scale_same_spin=kappa[1],
                           precision=precision)
        return e, F, time.perf_counter() - t

    evaluation(u.FP64)
    e64, F64, t64 = evaluation(u.FP64)
    out = {'workload': workload('c3', system, kappa, 'fp64', integrals).split(', plain')[0],
           'setup_s': round(t_setup, 2), 'fp64_ms_per_eval': round(1e3 * t64, 3)}
    for fname, precision in PRECISION(u).items():
        if fname == 'fp64':
            continue
        e, F, tp = evaluation(precision)
        dev = deviation(e, F, e64, F64)
        worst_e = max(abs(v) for v in dev['term_error_eh'].values())
        out[fname] = dict(dev, ms_per_eval=round(1e3 * tp, 3), speedup_vs_fp64=round(t64 / tp, 3),
                          meets_gate=bool(worst_e <= GATE and dev['fock_max_abs_error'] <= GATE))
        print(f'[bench] c3: {fname} {1e3 * tp:.1f} ms, Fock max-abs error {dev["fock_max_abs_error"]:.2e}',
              file=sys.stderr, flush=True)
    del W, V, WV, P
    return out


def workload(name, system, kappa, form, integrals):
    N, Nd, Nr = system['N'], system['Nd'], system['Nr']
    occ_counts = [len(x) for x in system['occ']]
    return (f'{name}: UW12 energy+Fock, n_ao={N} n_df={Nd} n_ri={Nr} n_occ={occ_counts}, kappa=({kappa[0]},'
            f'{kappa[1]}), {INTEGRALS_TEXT[integrals]}, {FORM_TEXT[form]}')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--config', default='c2', choices=sorted(CONFIG_DEFAULTS))
    ap.add_argument('--kappa', default=None, help='k0,k1 (default: (1,0.5) for c1/c2, (1,0) for c3-c5)')
    ap.add_argument('--integrals', default=None, choices=['host', 'device', 'streamed'],
                    help='default: host for c1/c2, device for c3/c4, streamed for c5')
    ap.add_argument('--form', default='fp16x3', choices=['mixed', 'fp64', 'fp32', 'bf16x3', 'bf16x6', 'fp16x3'],
                    help='precision form of the timed steps (the plain FP64 path is run beside it)')
    ap.add_argument('--side-reps', type=int, default=3,
                    help='timed evaluations of each side form (after one untimed one when more than 1)')
    ap.add_argument('--side-forms', default='auto', choices=['auto', 'gate', 'skip'],
                    help='gate: only the FP64 evaluation the accuracy gate needs; skip: no FP64 / FP32 / mixed '
                         'evaluations beside the timed ones (no gate), for profiler runs that should see only the '
                         'timed form')
    ap.add_argument('--c3-side', default='auto', choices=['auto', 'skip'],
                    help='auto: after the timed c2 steps, one c3 evaluation per precision form')
    ap.add_argument('--c5-side', default='auto', choices=['auto', 'skip'],
                    help='auto: after the timed c2 steps, one c5 evaluation per form (the timed form, FP64 gate)')
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
    kappa_default, integrals_default = CONFIG_DEFAULTS[args.config]
    kappa = tuple(float(x) for x in args.kappa.split(',')) if args.kappa else kappa_default
    integrals = args.integrals or integrals_default
    if integrals == 'streamed' and args.form == 'fp32':
        print('the FP32 form needs stored integrals (use --integrals device)', file=sys.stderr)
        sys.exit(2)

    import libuw12_amd as u

    ctx = u.Context(0)
    if args.verbose:
        ctx.set_verbosity(1)
    system, handles, t_setup = setup_config(args.config, integrals, ctx)
    m = measure(args.config, system, handles, ctx, kappa, args.form, args.steps, args.warmup, args.side_forms,
                args.side_reps, energy_only=integrals != 'streamed', label=args.config)
    roof = m['roofline']
    roof['traffic'] = measured_traffic(roof['kernel'], args.config, args.form)
    if roof['traffic'] is not None:
        roof['traffic_source'] = f'{traffic_profile(args.config, args.form)} (rocprofv3 FETCH_SIZE + WRITE_SIZE)'
        if roof.get('compulsory_bytes_per_launch'):
            roof['traffic_over_compulsory'] = round(roof['traffic'] / roof['compulsory_bytes_per_launch'], 3)
    if args.profile_table:
        print(json.dumps(m['table'], indent=1), file=sys.stderr)
    cpu = None
    if args.cpu_baseline in ('auto', 'full') and integrals == 'host':
        print('[bench] timing CPU baseline (oracle)', file=sys.stderr, flush=True)
        side = args.side_forms in ('auto', 'gate')
        cpu = cpu_baseline(system, kappa[0], kappa[1], full=args.cpu_baseline == 'full',
                           gpu_result={'fp64': (m['terms64'], m['F64']), args.form: (m['terms'], m['F'])} if side
                           else {args.form: (m['terms'], m['F'])})
    value = m['value']
    out = {
        'metric': METRIC,
        'value': round(value, 6),
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * m['elapsed'] / args.steps, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'fp64': 'f64', 'fp32': 'f32', 'mixed': 'f32+f64', 'bf16x3': 'bf16x3+f64',
                  'bf16x6': 'bf16x6+f64', 'fp16x3': 'fp16x3+f64'}[args.form],
        'data': 'synthetic (seeded s-Gaussian cluster, analytic Boys-F0 integrals)',
        'config': {'workload': workload(args.config, system, kappa, args.form, integrals),
                   'form': args.form, 'kappa': list(kappa), 'integrals': integrals, 'parallelism': 'single-gpu'},
        'roofline': roof,
        'cpu_baseline': cpu,
        'accuracy_gate': m['gate'],
        'fp64_form': m['fp64'],
        'energy': m['terms']['total'],
        'gpu_tflops_executed': round(m['gpu_flops'] * value / 1e12, 3),
        'model_flops_per_eval': m['model'],
        'executed_flops_per_eval': m['gpu_flops'],
        'setup_s': round(t_setup, 2),
    }
    out.update(m['others'])
    if args.config == 'c2' and args.c3_side == 'auto':
        del handles, m
        handles = m = None
        ctx.close()
        ctx = u.Context(0)
        out['c3_forms'] = c3_forms(ctx)
    if args.config == 'c2' and args.c5_side == 'auto':
        # the n_ao = 1500 configuration the metric is quoted on: one timed evaluation per form
        del handles, m
        ctx.close()
        ctx = u.Context(0)
        if args.verbose:
            ctx.set_verbosity(1)
        k5, i5 = CONFIG_DEFAULTS['c5']
        sys5, h5, t5 = setup_config('c5', i5, ctx)
        f5 = 'mixed' if args.form in ('fp64', 'fp32') else args.form  # streamed integrals: no FP32 form
        m5 = measure('c5', sys5, h5, ctx, k5, f5, 1, 1, 'gate', 1, energy_only=False, label='c5')
        r5 = m5['roofline']
        r5['traffic'] = measured_traffic(r5['kernel'], 'c5', f5)
        out['c5_form'] = {
            'workload': workload('c5', sys5, k5, f5, i5), 'form': f5,
            'evals_per_s': round(m5['value'], 6), 'ms_per_eval': round(1e3 * m5['elapsed'], 3),
            'steps': 1, 'warmup': 1, 'fp64_form': m5['fp64'], 'accuracy_gate': m5['gate'], 'roofline': r5,
            'energy': m5['terms']['total'], 'executed_flops_per_eval': m5['gpu_flops'],
            'model_flops_per_eval': m5['model'], 'setup_s': round(t5, 2)}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
