"""TEST INFRASTRUCTURE ONLY: CPU restatement of the tree metadata.

Restates src/warp_psr/tree_metadata.py:95-153 (build_tree_metadata) with
Python sets and dictionaries instead of sorts and binary searches: per level
(finest first) the union of the active support nodes, the broods of the finer
level's parents (tree.py:26-39) and the broods of the samples, density splat
records and normal records of that depth (tree_metadata.py:14-26); then two
padding levels, global coordinates (local + 2^(d+1), :29-39) and the flags,
ghost marks and links of :42-74.  Used by tests/ to check the library's
kernels; nothing in the product imports it.
"""
import numpy as np

OFF = 524288


def tree_key(x, y, z):
    return ((int(x) + OFF) << 40) | ((int(y) + OFF) << 20) | (int(z) + OFF)


def tree_coords(k):
    k = int(k)
    return ((k >> 40) & 1048575) - OFF, ((k >> 20) & 1048575) - OFF, (k & 1048575) - OFF


def basis_coords(k):
    k = int(k)
    return (k >> 32) & 0xFFFF, (k >> 16) & 0xFFFF, k & 0xFFFF


def _brood(x, y, z):
    px, py, pz = x // 2, y // 2, z // 2  # floor division, as tree.py:26-28
    return [tree_key(2 * px + (k & 1), 2 * py + ((k >> 1) & 1), 2 * pz + ((k >> 2) & 1)) for k in range(8)]


def build(depth, tree_keys, sample_keys, density_keys, normal_keys):
    """tree_keys[d], density_keys[d], normal_keys[d]: per-depth int64 arrays;
    sample_keys: the leaf sample keys.  Returns [(local_depth, keys, flags,
    ghost, parents, children)] for levels -2 .. depth."""
    raw = [None] * (depth + 1)
    for d in range(depth, -1, -1):
        cand = set(int(k) for k in tree_keys[d])
        if d < depth:
            parents = {tree_key(x // 2, y // 2, z // 2) for x, y, z in map(tree_coords, raw[d + 1])}
            for p in parents:
                cand.update(_brood(*tree_coords(p)))
        sources = [sample_keys] if d == depth else []
        sources += [density_keys[d], normal_keys[d]]
        for keys in sources:
            for k in keys:
                if int(k) >> 48 == d:
                    cand.update(_brood(*basis_coords(k)))
        raw[d] = sorted(cand)
    levels = [(-2, [tree_key(0, 0, 0)]), (-1, [tree_key(i // 4, (i // 2) % 2, i % 2) for i in range(8)])]
    for d in range(depth + 1):
        o = 1 << (d + 1)
        levels.append((d, sorted(tree_key(x + o, y + o, z + o) for x, y, z in map(tree_coords, raw[d]))))
    index = [{k: i for i, k in enumerate(keys)} for _, keys in levels]
    active = [set(int(k) for k in tree_keys[d]) for d in range(depth + 1)]
    out = []
    for g, (d, keys) in enumerate(levels):
        n = len(keys)
        flags = np.zeros(n, np.uint8)
        ghost = np.zeros(n, np.uint8)
        parents = np.full(n, -1, np.int32)
        children = np.full((n, 8), -1, np.int32)
        for i, key in enumerate(keys):
            x, y, z = tree_coords(key)
            if g > 0:
                parents[i] = index[g - 1].get(tree_key(x // 2, y // 2, z // 2), -1)
            is_ghost, bits = d == -2, 0
            if d >= 0:
                o = 1 << (d + 1)
                lx, ly, lz = x - o, y - o, z - o
                is_ghost = tree_key(lx, ly, lz) not in active[d]
                if not is_ghost and min(lx, ly, lz) >= 0:
                    if max(lx, ly, lz) < (1 << d):
                        bits |= 1 | 4
                    if max(lx, ly, lz) <= (1 << d):
                        bits |= 2
            has_child, child_active = False, d < -1
            for k in range(8):
                q = (2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + ((k >> 2) & 1))
                c = index[g + 1].get(tree_key(*q), -1) if g + 1 < len(levels) else -1
                children[i, k] = c
                if c >= 0:
                    has_child = True
                    if d >= -1:
                        o = 1 << (d + 2)
                        child_active = child_active or tree_key(q[0] - o, q[1] - o, q[2] - o) in active[d + 1]
            if has_child and not child_active:
                bits |= 64
            flags[i], ghost[i] = bits, int(is_ghost)
        out.append((d, np.array(keys, np.int64), flags, ghost, parents, children))
    return out
