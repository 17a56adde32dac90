"""Digests of partition results shared by the committed golden-digest script
(tests/golden/make_config_digests.py) and the GPU tests that compare against it."""
import hashlib
import struct

import numpy as np


def labels_digest(lab) -> str:
    return hashlib.sha256(np.ascontiguousarray(lab, dtype=np.int32).tobytes()).hexdigest()


def f64_hex(x) -> str:
    return struct.pack("<d", float(x)).hex()


def trace_digest(records) -> str:
    """sha256 over every trace record (kind, B, a, b, h1, h2, x, y) in order."""
    h = hashlib.sha256()
    for r in records:
        h.update(struct.pack("<iiiiQQdd", *r))
    return h.hexdigest()


def result_digest(lab, B, S, levels, sweeps) -> dict:
    return dict(B=int(B), S=f64_hex(S), S_value=float(S), labels=labels_digest(lab),
                levels=int(levels), sweeps=int(sweeps))
