"""fhs: MI355X-native Faceted HyperSARA primal-dual iteration.

Host-side mirror of the reference's hot-path functions (MATLAB names and
argument meaning) over the C ABI of libfhs.so (include/fhs.h).
"""
from ._lib import FhsError, lib  # noqa: F401
from .core import Context, GMatrix, NufftPlan, default_context  # noqa: F401
from .operators import (  # noqa: F401
    apply_adjoint_operator,
    apply_direct_operator,
    apply_scaled_fourier_transform,
    fhs_update_primal,
    proj_l2ball,
    update_epsilon,
)
from .data_fidelity import DataOperator, update_dual_data_fidelity  # noqa: F401,E402
from .lowrank import LowRankPrior, hs_update_dual_lowrankness_serial, nuclear_norm  # noqa: F401,E402
from .sara import SaraDictionary, SaraFacets, compute_sara_prior, hs_update_dual_sparsity_distributed  # noqa: F401,E402,E501
