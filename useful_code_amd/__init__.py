"""MI355X (gfx950) LMM Monte Carlo pricing pipeline.

The compute path is ``libkoo_amd.so`` (hand-written HIP kernels behind a C
ABI, declared in include/koo_c_abi.h).  ``useful_code_amd.koo`` binds it with
ctypes and gives torch-tensor wrappers that keep the reference's launcher
names and argument order; ``useful_code_amd.pipeline`` mirrors the reference's
host orchestration (LMMEvolutionRoutine*) and the fused pricing/adjoint path;
``useful_code_amd.configs`` builds the synthetic workloads.

Importing the package loads the library and raises if it has not been built
(``python -m useful_code_amd.build``): there is no CPU fallback.
"""
from .koo import KooError, lib  # noqa: F401

__all__ = ["KooError", "lib"]
