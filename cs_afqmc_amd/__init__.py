"""cs_afqmc_amd: ad_afqmc's phaseless AFQMC propagation step on MI355X (gfx950).

The hot path (force bias, VHS, Taylor/one-body propagation, overlaps, weights, local
energy, QR, stochastic reconfiguration) runs in hand-written HIP kernels behind the C
ABI of include/cs_afqmc_amd.h; this package mirrors the reference's propagator /
trial / hamiltonian / sampler API on top of it.
"""

from . import _lib  # noqa: F401

__all__ = ["hamiltonian", "propagation", "wavefunctions", "sampling", "problems", "rng"]
