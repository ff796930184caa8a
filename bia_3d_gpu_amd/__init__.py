"""bia_3d_gpu_amd -- MI355X (gfx950) library for the pairwise hot path of the Binary
Interaction Approximation N-body method.

The compute lives in ``libbia_hip.so`` (hand-written HIP kernels, C ABI in
``include/bia_hip.h``).  This package is the Python host side: ctypes bindings that mirror
the reference's Fortran call interface (:mod:`.api`), the reference's file formats
(:mod:`.io`), seeded synthetic initial conditions (:mod:`.ic`), the monitored multi-step
driver (:mod:`.driver`) and the learned-initialiser training (:mod:`.learn`).
"""
from . import io, ic  # noqa: F401
from .api import (  # noqa: F401
    METHOD_ALG, METHOD_RKF, MODE_GUESS, MODE_LEARN, MODE_PLAIN, MODE_PLAIN_LEARN, PAIR_RECORD_DTYPE, STATS_LEN, Context, FailureLog,
    find_dxyuv_kernel, flops_per_step, partial_doubles, segment_length, staging,
    invariance_gpu, reference_step, rkf_c,
)
from ._lib import COUPLING_COULOMB, COUPLING_GRAVITY, COUPLING_SCALAR, LIB_PATH, load  # noqa: F401

__version__ = "1.0.0"
