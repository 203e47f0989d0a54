"""MI355X-native combustion-space SIMPLER(R) solver (GFM v4 hot path).

The solver itself is the HIP library ``libgfm_hip.so`` behind the C-ABI in
``include/gfm_hip.h``; this package holds its ctypes host mirror
(``gfm_amd.solver``), the synthetic case generator (``gfm_amd.case``) and
the Fortran host layer (``gfm_amd/fortran``).
"""
__all__ = ["case", "solver", "abi"]
