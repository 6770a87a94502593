"""Image sequences that lose tracking and come back, for the SlamSystem
relocaliser (shared by the CPU and GPU tests).

The camera moves through one synthetic scene 2x faster than the default
SlamSystem workload (so that several keyframes are finished early) and the
initialisation phase is one keyframe long (so that a bad track counts as a
loss early).  Two black images (a covered lens: most points bad) make the
tracker report the loss; then either
  * "resume": the sequence continues where it was, or
  * "kidnap": the camera reappears at an early pose of the sequence, 0.9 m
    back along the path, where only the first keyframe overlaps it."""
import numpy as np

import synth

STEP = (0.008, 0.002, 0.024)
N = 46
LOST_AFTER = 39


def sequence():
    return synth.vo_sequence(7, N, step=STEP)


def order(kind):
    """(sequence index or None for a black image) per tracked image."""
    head = list(range(1, LOST_AFTER + 1)) + [None, None]
    if kind == "resume":
        return head + list(range(LOST_AFTER + 1, N))
    return head + [3, 4, 5, 6]


def image(seq, idx):
    return np.zeros_like(seq["images"][0]) if idx is None else seq["images"][idx]
