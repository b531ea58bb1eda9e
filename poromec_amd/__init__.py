"""MI355X (gfx950) implementation of poroMec's mef Biot u-p solve path.

Compute lives in ``libpmhip.so`` (HIP kernels behind the C ABI of
``include/poromec_hip.h``); this package holds the synthetic decks and a thin
host-side mirror of the reference driver."""
from .problem import Deck, box_counts, column_deck, heterogeneous_box_deck  # noqa: F401

__all__ = ["Deck", "box_counts", "column_deck", "heterogeneous_box_deck"]
