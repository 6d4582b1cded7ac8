"""Learned encoder proposal: a BF16 MLP trained on the GPU by this package's
HIP kernels proposes the predictor and residual coders that the exhaustive
search would pick; the GPU scores the proposal and the exhaustive search
stays the fallback and the yardstick of bit-rate regret."""
from .mlp import ProposalMLP  # noqa: F401
