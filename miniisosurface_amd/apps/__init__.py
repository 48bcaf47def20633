"""Command-line drivers with the reference miniapp's flags (built by build())."""
from __future__ import annotations


def build() -> None:
    """Placeholder until the C++ drivers land."""
    return None
