"""ipeps_amd -- MI355X (gfx950) CTMRG engine for iPEPS with a host API that
mirrors yitan1/iPEPS.jl.

The numerics run in libipeps_hip.so (hand-written HIP kernels behind a C
ABI, include/ipeps_hip.h); this package is the ctypes host side.
"""
from ._lib import C128, F64, MIXED, LIB_PATH, IpepsError, lib
from .ctm import (ENV_OPTIONS, absorb, Context, CTMTensors, compute_gs_energy, default_context, get_gs_energy,
                  get_gs_norm, gemm, gs_energy_grad, orth, rg_step, run_ctm, run_graph, run_gs_vjp, step_adjoint, svd1,
                  svd_back,
                  svd_thin,
                  tcon, tune, tune_clear, tune_load, tune_save, tune_stats)
from . import models
from . import excited
from . import batch
from . import expect
from . import optim
from .optim import optim_gs
from .expect import correlation_spin, correlation_TM
from .batch import BatchedCTM, c4_instances
from .excited import compute_es, get_es_grad, make_es_path, optim_es

__all__ = ["C128", "F64", "MIXED", "LIB_PATH", "IpepsError", "lib", "ENV_OPTIONS", "Context", "CTMTensors",
           "compute_gs_energy", "default_context", "get_gs_energy", "get_gs_norm", "gs_energy_grad", "rg_step",
           "run_ctm", "run_graph", "run_gs_vjp", "optim_gs", "step_adjoint", "tune", "tune_stats", "tune_clear", "tune_save", "tune_load", "svd1", "svd_back", "svd_thin", "orth", "tcon", "gemm", "models", "BatchedCTM",
           "c4_instances", "correlation_spin", "correlation_TM"]
