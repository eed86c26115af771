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


def test_roofline_peak_of_the_split_kernels():
    """The plane GEMMs are rated against the BF16 / FP16 dense peak over their products per
    multiply-add (three with two planes, six with three), and carry their compulsory bytes."""
    sys.path.insert(0, ROOT)
    import bench
    recs = [('3e_ind_fock_kt|gemm_planes_kernel<128,128,2,1>', 10.0, 3.0e12, 4.0e9),
            ('3e_ind_fock_l|gemm_planes_kernel<128,128,2,1>', 5.0, 1.5e12, 2.0e9)]
    roof, _ = bench.roofline_from_profile(recs)
    assert roof['kernel'] == 'gemm_planes_kernel<128,128,2,1>'
    assert roof['peak'] == round(bench.PEAK_BF16_TFLOPS / 3, 1)
    assert abs(roof['compulsory_bytes_per_launch'] - 3.0e9) < 1.0
    roof6, _ = bench.roofline_from_profile([('x|gemm_planes_kernel<128,128,3,0>', 1.0, 1e11, 0.0)])
    assert roof6['peak'] == round(bench.PEAK_BF16_TFLOPS / 6, 1)
    assert 'compulsory_bytes_per_launch' not in roof6
