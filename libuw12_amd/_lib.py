"""ctypes binding of libuw12g.so (the HIP library).  Fails loudly when the
library has not been built: there is no CPU fallback anywhere in the package."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libuw12g.so')

_lib = None

dp = ctypes.POINTER(ctypes.c_double)
szp = ctypes.POINTER(ctypes.c_size_t)
vp = ctypes.c_void_p


class Map2(ctypes.Structure):
    _fields_ = [('s0', ctypes.c_longlong), ('s1', ctypes.c_longlong), ('n0', ctypes.c_uint)]


# every symbol declared in include/uw12g.h and include/uw12g_testing.h
EXPORTS = (
    'uw12g_last_error', 'uw12g_context_create', 'uw12g_context_destroy', 'uw12g_setup_base_integrals',
    'uw12g_setup_base_integrals_device', 'uw12g_setup_base_integrals_fn', 'uw12g_setup_base_integrals_gen',
    'uw12g_base_integrals_destroy',
    'uw12g_base_integrals_kept',
    'uw12g_setup_abs_projectors', 'uw12g_abs_projectors_destroy', 'uw12g_abs_projectors_get', 'uw12g_energy',
    'uw12g_fock', 'uw12g_form_fock', 'uw12g_form_fock_noalloc', 'uw12g_forward', 'uw12g_adjoint', 'uw12g_context_set_profiling',
    'uw12g_context_set_verbosity',
    'uw12g_context_profile',
    'uw12g_test_gemm', 'uw12g_test_map2', 'uw12g_time_gemm', 'uw12g_test_set_stream_cap',
    'uw12g_test_set_exact_stages', 'uw12g_test_set_block_cap', 'uw12g_test_set_tape_cap', 'uw12g_test_force_tile',
    'uw12g_test_set_bf16_planes', 'uw12g_test_set_kept_gt',
)


# int (*)(void* user, size_t shell, double* out): one DF shell of three-index integrals
THREE_INDEX_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double))


class LibraryMissing(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f'{LIB_PATH} is not built; run `python -m libuw12_amd.build` (hipcc, gfx950)')
    L = ctypes.CDLL(LIB_PATH)
    L.uw12g_last_error.restype = ctypes.c_char_p
    L.uw12g_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.uw12g_context_destroy.argtypes = [vp]
    L.uw12g_context_destroy.restype = None
    L.uw12g_setup_base_integrals.argtypes = [vp, dp, dp, dp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                             ctypes.POINTER(vp)]
    L.uw12g_setup_base_integrals_device.argtypes = [vp, vp, dp, vp, ctypes.c_size_t, ctypes.c_size_t,
                                                    ctypes.c_size_t, ctypes.POINTER(vp)]
    L.uw12g_setup_base_integrals_fn.argtypes = [vp, dp, THREE_INDEX_FN, THREE_INDEX_FN, vp,
                                                ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t, ctypes.c_size_t,
                                                ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(vp)]
    # the generator is a raw C function pointer (uw12g_three_index_gen_fn), e.g. from libuw12_synth_dev.so
    L.uw12g_setup_base_integrals_gen.argtypes = [vp, dp, vp, vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                                 ctypes.c_int, ctypes.POINTER(vp)]
    L.uw12g_base_integrals_destroy.argtypes = [vp]
    L.uw12g_base_integrals_destroy.restype = None
    L.uw12g_base_integrals_kept.argtypes = [vp]
    L.uw12g_base_integrals_kept.restype = ctypes.c_long
    L.uw12g_setup_abs_projectors.argtypes = [vp, dp, ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(vp)]
    L.uw12g_abs_projectors_destroy.argtypes = [vp]
    L.uw12g_abs_projectors_destroy.restype = None
    L.uw12g_abs_projectors_get.argtypes = [vp, dp]
    common = [vp, vp, vp, vp, dp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, dp, szp, szp]
    L.uw12g_energy.argtypes = [vp] + common + [ctypes.c_double, ctypes.c_double, dp]
    L.uw12g_fock.argtypes = [vp, dp] + common + [ctypes.c_double, ctypes.c_double, dp]
    L.uw12g_form_fock.argtypes = [vp] + common + [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp,
                                                  dp, dp]
    L.uw12g_form_fock_noalloc.argtypes = L.uw12g_form_fock.argtypes
    L.uw12g_forward.argtypes = [vp] + common + [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp,
                                                ctypes.POINTER(ctypes.c_ulonglong)]
    L.uw12g_adjoint.argtypes = [vp, ctypes.c_ulonglong, dp]
    L.uw12g_context_set_profiling.argtypes = [vp, ctypes.c_int]
    L.uw12g_context_set_verbosity.argtypes = [vp, ctypes.c_int]
    L.uw12g_context_profile.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p), dp, dp, dp, ctypes.c_int]
    L.uw12g_test_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double,
                                  dp, ctypes.c_size_t, Map2, Map2, ctypes.c_longlong,
                                  dp, ctypes.c_size_t, Map2, Map2, ctypes.c_longlong,
                                  dp, ctypes.c_size_t, Map2, Map2, ctypes.c_longlong,
                                  dp, ctypes.c_size_t, Map2, ctypes.c_int]
    L.uw12g_time_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double,
                                  vp, ctypes.c_size_t, Map2, Map2, ctypes.c_longlong,
                                  vp, ctypes.c_size_t, Map2, Map2, ctypes.c_longlong,
                                  vp, ctypes.c_size_t, Map2, Map2, ctypes.c_longlong,
                                  vp, ctypes.c_size_t, Map2, ctypes.c_int, ctypes.c_int, dp]
    L.uw12g_test_set_stream_cap.argtypes = [vp, ctypes.c_size_t]
    L.uw12g_test_set_exact_stages.argtypes = [vp, ctypes.c_char_p]
    L.uw12g_test_set_block_cap.argtypes = [vp, ctypes.c_size_t]
    L.uw12g_test_set_tape_cap.argtypes = [vp, ctypes.c_size_t]
    L.uw12g_test_force_tile.argtypes = [ctypes.c_int, ctypes.c_int]
    L.uw12g_test_set_bf16_planes.argtypes = [ctypes.c_int, ctypes.c_longlong]
    L.uw12g_test_set_kept_gt.argtypes = [ctypes.c_int]
    L.uw12g_test_map2.argtypes = [Map2, ctypes.c_uint]
    L.uw12g_test_map2.restype = ctypes.c_longlong
    _lib = L
    return L


class UW12Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'[{code}] {msg}')
        self.code = code


def check(rc):
    if rc != 0:
        raise UW12Error(rc, load().uw12g_last_error().decode())
