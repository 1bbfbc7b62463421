"""Block-size / kernel-variant tuner (placeholder: exhaustive timing)."""
NAME = "exhaustive-timing"


def pick_block_size(A, candidates):
    """Return the chosen s, or None to use the best measured candidate."""
    return None
