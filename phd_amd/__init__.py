"""phd_amd: the ADER-WENO time step of the Godunov-Peshkov-Romenski system on
AMD Instinct MI355X (gfx950), as a drop-in for the reference's GPRpy solver
functions (haranjackson/phd, code/cpp/bindings/bindings.cpp).

Compute runs only in libgpr_mi355x.so (hand-written HIP kernels); see
phd_amd.gprpy for the GPRpy-compatible Python surface and include/gpr_mi355x.h
for the C ABI.
"""

from .params import Par, Params, material_params, Cvec  # noqa: F401

__all__ = ["Par", "Params", "material_params", "Cvec"]
