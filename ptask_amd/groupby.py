"""Groupby scratch layout (include/ptk_groupby.h) for host-side sizing and tests."""
from __future__ import annotations

CHUNK = 2048          # rows per chunk (PTK_GB_CHUNK)
BUCKET = 256          # keys per bucket (PTK_GB_BUCKET)
GROUP = 256           # chunks per merge group (PTK_GB_GROUP)
PLACE_THREADS = 512   # threads of a GroupbyReshuffle block (PTK_GB_PLACE_THREADS)
MAX_BUCKETS = 2048    # PTK_GB_MAX_BUCKETS
MAX_KEYS = BUCKET * MAX_BUCKETS   # numKeys + 1 <= MAX_KEYS


def chunks(n: int) -> int:
    return (n + CHUNK - 1) // CHUNK


def buckets(num_keys: int) -> int:
    return (num_keys + 1 + BUCKET - 1) // BUCKET


def groups(n: int) -> int:
    return (chunks(n) + GROUP - 1) // GROUP


def _a16(b: int) -> int:
    return (b + 15) // 16 * 16


def scratch_offsets(n: int, num_keys: int) -> dict:
    """Byte offsets of btot, berr, boff, gsum, gcnt, entry and the total size."""
    B, NB, NG = chunks(n), buckets(num_keys), groups(n)
    KP = NB * BUCKET
    o = {"btot": 0}
    o["berr"] = _a16(16 * NB)
    o["boff"] = _a16(o["berr"] + 4 * B)
    o["gsum"] = _a16(o["boff"] + 4 * B * (NB + 1))
    o["gcnt"] = _a16(o["gsum"] + 8 * NG * KP)
    o["entry"] = _a16(o["gcnt"] + 4 * NG * KP)
    o["total"] = o["entry"] + 16 * B * CHUNK
    return o


def scratch_bytes(n: int, num_keys: int) -> int:
    return scratch_offsets(n, num_keys)["total"]
