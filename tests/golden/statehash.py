"""Checksums of solver state for full-size parity fixtures.

A state (tests/helpers.state layout: pos.xy and prev.xy of every body, polygon
centres) is reduced to f32 bit patterns with every NaN mapped to one pattern
(IEEE 754 leaves a NaN's sign and payload open, and x86 and gfx950 differ),
then hashed with SHA-256: once whole, and once per block of bodies so that a
mismatch can be located without the oracle's full state."""
import hashlib

import numpy as np


def canonical_bits(x):
    a = np.array(np.asarray(x, np.float32), copy=True).view(np.uint32)
    a[np.isnan(a.view(np.float32))] = 0x7FC00000
    return a


def state_hash(x):
    return hashlib.sha256(canonical_bits(x).tobytes()).hexdigest()


def block_hashes(x, block_bodies):
    """SHA-256 (first 16 hex digits) per block of block_bodies bodies (4 values each)."""
    b = canonical_bits(x)
    step = 4 * block_bodies
    return [hashlib.sha256(b[i:i + step].tobytes()).hexdigest()[:16] for i in range(0, len(b), step)]
