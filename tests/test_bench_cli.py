"""bench.py command line on the CPU: one GPU only; the FLOP model and the profile reading are
host logic."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr
    assert r.stdout == ''


def test_roofline_picks_dominant_instantiation_and_its_peak():
    sys.path.insert(0, ROOT)
    import bench
    recs = [('3e_ind_fock_kt|gemm_f64_kernel<128,128,0,0,0,1>', 10.0, 1.2e12, 0.0),
            ('3e_ind_fock_l|gemm_f64_kernel<128,128,0,0,0,1>', 8.0, 1.0e12, 0.0),
            ('3e_ind_fock_f|gemm_f64_kernel<64,64,1,1,0,0>', 6.0, 3.6e11, 0.0)]
    roof, table = bench.roofline_from_profile(recs)
    assert roof['kernel'] == 'gemm_f64_kernel<128,128,0,0,0,1>'
    assert roof['peak'] == bench.PEAK_FP32_TFLOPS
    assert abs(roof['achieved'] - 2.2e12 / 18e-3 / 1e12) < 1e-3  # rounded to 3 decimals
    assert roof['launches'] == 2
    roof64, _ = bench.roofline_from_profile(recs[2:])
    assert roof64['peak'] == bench.PEAK_FP64_TFLOPS
