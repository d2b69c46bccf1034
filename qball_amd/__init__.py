"""qball_amd -- MI355X-native Hamiltonian-application (Hψ) engine for plane-wave DFT/TDDFT.

The product is the C-ABI library ``libqbh.so`` (HIP kernels for gfx950, declared in
``include/qbh.h``) and the C++ adaptor ``qball_amd/host/qbh_qball.h`` that presents it
behind qball-shaped classes.  This Python package is a thin ctypes binding of the same C
ABI, used by the tests and the benchmark; it refuses to run without the built library.
"""
from .capi import (QBH_ALL, QBH_ACCUMULATE, QBH_FP32, QBH_KINETIC, QBH_LOCAL, QBH_NONLOCAL,
                   Basis, Batch, Context, Hamiltonian, QbhError, Species, TDPropagator,
                   auto_rcps, ewald_esr, factorizable, lib, local_tables, rods_build)

__all__ = ["QBH_ALL", "QBH_ACCUMULATE", "QBH_FP32", "QBH_KINETIC", "QBH_LOCAL",
           "QBH_NONLOCAL", "Basis", "Batch", "Context", "Hamiltonian", "QbhError", "Species",
           "TDPropagator", "auto_rcps", "ewald_esr", "factorizable", "lib", "local_tables",
           "rods_build"]
