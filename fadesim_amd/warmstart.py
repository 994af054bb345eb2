"""Learned warm start for the batched solver (placeholder until trained
weights exist: load_default returns None)."""
from __future__ import annotations

import os

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def load_default(device):
    """The trained warm-start network, or None when no weights are present."""
    return None
