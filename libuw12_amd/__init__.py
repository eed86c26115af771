"""MI355X-native UW12 energy + Fock evaluation (drop-in for libuw12's
uw12_energy / uw12_fock; C ABI in include/uw12g.h, HIP kernels in csrc/)."""
from .api import (FP32, FP64, MIXED, MIXED_BF16X3, MIXED_BF16X6, MIXED_FP16X3, ABSProjectors, BaseIntegrals, Context, TERMS, Tape, UW12Error, adjoint, form_fock, forward,
                  setup_abs_projectors, setup_base_integrals, setup_base_integrals_device, setup_base_integrals_fn,
                  setup_base_integrals_gen,
                  setup_occupations,
                  setup_orbitals, uw12_energy, uw12_fock)

__all__ = ['FP32', 'FP64', 'MIXED', 'MIXED_BF16X3', 'MIXED_BF16X6', 'MIXED_FP16X3', 'ABSProjectors', 'BaseIntegrals', 'Context', 'TERMS', 'Tape', 'UW12Error', 'adjoint', 'form_fock',
           'forward', 'setup_abs_projectors',
           'setup_base_integrals', 'setup_base_integrals_device', 'setup_base_integrals_fn',
           'setup_base_integrals_gen', 'setup_occupations', 'setup_orbitals',
           'uw12_energy', 'uw12_fock']
