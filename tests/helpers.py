"""Shared helpers for the parity tests: random sorted, disjoint leaf sets and comparisons."""
import numpy as np


def random_tree(rng, dim, n_target, max_depth=12, flag_prob=0.5, flags_pool=(0x08, 0x01, 0x09, 0x00, 0x10)):
    """A random set of disjoint leaves (bits, depth, flags) obtained by random refinement of the
    root, returned in arbitrary order; both implementations sort on upload."""
    leaves = [(0, 0)]
    while len(leaves) < n_target:
        i = int(rng.integers(len(leaves)))
        b, d = leaves[i]
        if d >= max_depth:
            if all(dd >= max_depth for _, dd in leaves):
                break
            continue
        leaves[i] = (b << 1, d + 1)
        leaves.append(((b << 1) | 1, d + 1))
    # drop some leaves so the tree has holes
    keep = rng.random(len(leaves)) > 0.2
    if not keep.any():
        keep[0] = True
    leaves = [l for l, k in zip(leaves, keep) if k]
    bits = np.array([l[0] for l in leaves], np.uint64)
    depth = np.array([l[1] for l in leaves], np.uint8)
    flags = np.array([flags_pool[int(rng.integers(len(flags_pool)))] if rng.random() < flag_prob else 0
                      for _ in leaves], np.uint8)
    return bits, depth, flags


def same_tree(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))
