"""ffmm: an MI355X (gfx950) backend for the device passes of FlexibleFMM.jl.

The compute path is libffmm_hip.so (hand-written HIP kernels behind the C ABI of
include/ffmm.h); this package is the host side: the reference's Julia API restated
in Python over ctypes, and the host FP64 operator precompute.
"""
from ._lib import ArgumentError, DimensionMismatch, DeviceError  # noqa: F401
from .fmm import (FMM, CubeKernelIndependent, DevicePlan, FMMState, Gradient, Laplace3D, OperatorCache,  # noqa: F401
                  Potential, PotentialGradient, Pressure, SphericalHarmonics, StokesHarmonics, Stokeslet3D, Velocity, Yukawa3D,
                  build_expansions, build_expansions_, device_build_expansions_, device_matvec_, direct, direct_self,
                  evaluate, evaluate_at_sources, plan_data_check, square_bbox)
from .circle2d import CircleOperatorCache, CirclePlan, KernelIndependent, Laplace2D  # noqa: F401
from .stokes_api import (StokesFMMKernel, StokesFMMPlan, stokes_densities, stokes_fmm_eval_,  # noqa: F401
                         stokes_fmm_eval_src_trg_)
