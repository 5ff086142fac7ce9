"""Unrestricted walkers with spin-resolved Cholesky vectors (ad_afqmc/prop_unrestricted).

The reference's prop_unrestricted package takes ham_data["chol"] as (2, G, N*N): alpha
vectors L^a_g and beta vectors L^b_g (for example the Cholesky vectors of a UHF
reference rotated into each spin's own orbital basis).  ``wavefunctions_unrestricted.uhf``
and ``propagation.propagator_unrestricted`` here mirror that package's trial and
propagator; the step runs in the HIP library's AFQMC_UHF_SPIN_CHOL contexts.
"""

from . import propagation, wavefunctions_unrestricted  # noqa: F401
