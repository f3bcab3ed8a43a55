"""Proposal record of the static accept/reject tail (geomjax/mcmc/proposal.py:22-40)."""
from typing import Any, NamedTuple

import torch


class Proposal(NamedTuple):
    """state: trajectory end state; energy: its total energy; weight: H0 - H
    (NaN mapped to -inf); sum_log_p_accept: min(weight, 0)."""

    state: Any
    energy: torch.Tensor
    weight: torch.Tensor
    sum_log_p_accept: torch.Tensor
