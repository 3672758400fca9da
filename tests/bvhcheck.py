"""Structural checks of flattened trees downloaded from a device scene."""
import numpy as np

import hiprt

INVALID = 0x7FFFFFFF


def children(nodes, fmt):
    """(refs (n, k) int32, lo (n, k, 3), hi (n, k, 3)) of downloaded nodes."""
    if fmt == hiprt.BVH2:
        w = nodes.reshape(-1, 16)
        f = w.view(np.float32)
        refs = w[:, 12:14].view(np.int32)
        lo = np.stack([np.stack([f[:, 0], f[:, 2], f[:, 8]], 1), np.stack([f[:, 4], f[:, 6], f[:, 10]], 1)], 1)
        hi = np.stack([np.stack([f[:, 1], f[:, 3], f[:, 9]], 1), np.stack([f[:, 5], f[:, 7], f[:, 11]], 1)], 1)
    else:
        w = nodes.reshape(-1, 32)
        f = w.view(np.float32)
        refs = w[:, 24:28].view(np.int32)
        lo = np.stack([f[:, 0:4], f[:, 8:12], f[:, 16:20]], 2)
        hi = np.stack([f[:, 4:8], f[:, 12:16], f[:, 20:24]], 2)
    return refs, lo, hi


def subtree_boxes(refs, slot_ids, pos, idx):
    """Exact bounding box of every child subtree's triangles (whole
    triangles), by a post-order walk from the children up."""
    tri = pos[idx]                                 # (n_tris, 3, 3)
    t_lo, t_hi = tri.min(1), tri.max(1)
    n, k = refs.shape
    lo = np.full((n, k, 3), np.inf, np.float32)
    hi = np.full((n, k, 3), -np.inf, np.float32)
    done = np.zeros(n, bool)
    for i in range(n - 1, -1, -1):                 # children have larger ids than their parents
        for c in range(k):
            r = int(refs[i, c])
            if r == INVALID:
                continue
            if r < 0:
                v = (~r) & 0xFFFFFFFF
                slot, cnt = v >> 5, (v & 31) + 1
                ids = slot_ids[slot:slot + cnt]
                lo[i, c] = t_lo[ids].min(0)
                hi[i, c] = t_hi[ids].max(0)
            else:
                assert r > i and done[r]
                lo[i, c] = lo[r].min(0)
                hi[i, c] = hi[r].max(0)
        done[i] = True
    return lo, hi
