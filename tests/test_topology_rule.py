"""The node-set rule the GPU topology is built on (gx_build.hip, topology
section), checked on the CPU against the reference's recursive split
(buildNode/split, internal/sim/bh/build.go:134-172): over sorted 63-bit
Morton keys, with delta_i the number of leading octal digits keys i and i+1
share (-1 between instances and past the ends), m(j) = min(delta_j ..
delta_{j+23}) and M(s) = max(m(s-24) .. m(s)), the nodes starting at star s
are exactly the levels delta_{s-1}+1 .. min(M(s)+1, 21), each ending at the
first star after s whose delta to its predecessor is below the level.
Random, clustered, deep, duplicated and multi-instance key sets, including
empty and tiny instances."""
import numpy as np
import pytest

LEAF, BITS = 24, 21


def split_tree(keys, off):
    """The reference's rule: a node of more than LEAF stars below level BITS
    splits into its non-empty octants at the next digit."""
    nodes = []
    for b in range(len(off) - 1):
        lo, hi = int(off[b]), int(off[b + 1])
        if hi == lo:
            continue
        level = [(0, lo, hi)]
        while level:
            nxt = []
            for L, a, e in level:
                nodes.append((L, a, e))
                if e - a > LEAF and L < BITS:
                    d = (keys[a:e] >> np.uint64(3 * (BITS - 1 - L))) & np.uint64(7)
                    cuts = np.flatnonzero(np.diff(d)) + 1
                    bounds = np.r_[0, cuts, e - a]
                    nxt += [(L + 1, a + int(bounds[k]), a + int(bounds[k + 1])) for k in range(len(bounds) - 1)]
            level = nxt
    return sorted(nodes)


def chain_tree(keys, off):
    n = len(keys)
    delta = np.full(n, -1, np.int64)  # delta[i]: stars i and i+1
    if n > 1:
        x = keys[:-1] ^ keys[1:]
        bl = np.array([int(v).bit_length() for v in x], np.int64)  # of 63-bit keys
        delta[:-1] = np.where(x == 0, BITS, (63 - bl) // 3)
    starts = set(int(v) for v in off[1:-1])
    for s in starts:
        if 0 < s <= n:
            delta[s - 1] = -1
    d = lambda i: -1 if i < 0 or i >= n - 1 else int(delta[i])
    m = np.array([min(d(j + t) for t in range(LEAF)) for j in range(n)], np.int64)
    nodes = []
    for s in range(n):
        M = max(int(m[j]) for j in range(max(0, s - LEAF), s + 1))
        lo = d(s - 1) + 1
        for L in range(lo, min(M + 1, BITS) + 1):
            e = s + 1
            while e < n and d(e - 1) >= L:
                e += 1
            nodes.append((L, s, e))
    return sorted(nodes)


def chain_first_levels(keys, off):
    """Per star, the level its chain of nodes starts at (delta_{s-1}+1), or
    22 when no node starts there: what the GPU keeps per star."""
    n = len(keys)
    out = np.full(n, BITS + 1, np.int64)
    for L, s, e in chain_tree(keys, off):
        out[s] = min(out[s], L)
    return out


def split_ends(keys, off):
    """The end of every split node by the rule of the links pass: the first
    star from start + LEAF + 1 on whose chain starts at the node's level or
    above (a smaller level number), or the instance's end."""
    first = chain_first_levels(keys, off)
    inst_end = {}
    for b in range(len(off) - 1):
        for s in range(int(off[b]), int(off[b + 1])):
            inst_end[s] = int(off[b + 1])
    got = []
    for L, s, e in split_tree(keys, off):
        if e - s > LEAF and L < BITS and L > 0:
            end = inst_end[s]
            j = s + LEAF + 1
            while j < end and first[j] > L:
                j += 1
            got.append((L, s, min(j, end), e))
    return got


def key_sets(seed):
    rng = np.random.default_rng(seed)
    for trial in range(12):
        parts = []
        for _ in range(int(rng.integers(1, 4))):
            size = int(rng.integers(0, 300))
            mode = trial % 4
            if mode == 0:
                k = rng.integers(0, 1 << 63, size=size, dtype=np.uint64)
            elif mode == 1:  # one deep clump
                k = rng.integers(0, 1 << 20, size=size, dtype=np.uint64) | np.uint64(5 << 40)
            elif mode == 2:  # runs of equal keys (level-21 leaves of more than 24 stars)
                k = np.repeat(rng.integers(0, 1 << 63, size=max(1, size // 30), dtype=np.uint64), 30)[:size]
            else:  # few top octants, tiny offsets
                k = (rng.integers(0, 8, size=size, dtype=np.uint64) << np.uint64(60)) | \
                    rng.integers(0, 64, size=size, dtype=np.uint64)
            parts.append(np.sort(k))
        off = np.r_[0, np.cumsum([len(p) for p in parts])].astype(np.int64)
        keys = np.concatenate(parts) if parts else np.zeros(0, np.uint64)
        yield keys, off


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_chain_rule_equals_the_recursive_split(seed):
    for keys, off in key_sets(seed):
        assert chain_tree(keys, off) == split_tree(keys, off)


def test_edge_cases():
    # one star, 24 and 25 equal keys, an empty instance between two others
    for keys, off in ((np.array([7], np.uint64), [0, 1]),
                      (np.full(24, 9, np.uint64), [0, 24]),
                      (np.full(25, 9, np.uint64), [0, 25]),
                      (np.sort(np.random.default_rng(4).integers(0, 1 << 63, 60, dtype=np.uint64)), [0, 30, 30, 60])):
        off = np.asarray(off, np.int64)
        assert chain_tree(keys, off) == split_tree(keys, off)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_split_node_ends_from_the_chains(seed):
    for keys, off in key_sets(seed):
        for L, s, got, want in split_ends(keys, off):
            assert got == want, (L, s, got, want)
