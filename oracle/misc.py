"""Seeded RNG and groupings (reference phaser/utils/misc.py).  Test infrastructure only."""
import hashlib
import json
import math

import numpy
from numpy.random import SeedSequence, PCG64, Generator, BitGenerator


def _proc_seed(seed, entropy=None):
    """SHA-256 of json(seed) + json(entropy) -> SeedSequence (misc.py:14-33)."""
    if seed is None:
        return SeedSequence()
    if isinstance(seed, SeedSequence):
        seed = seed.entropy
    state = hashlib.sha256()
    state.update(json.dumps(seed).encode('utf-8'))
    if entropy is not None:
        state.update(json.dumps(entropy).encode('utf-8'))
    return SeedSequence(numpy.frombuffer(state.digest(), dtype=numpy.uint32))


def create_rng(seed=None, entropy=None):
    """PCG64 Generator from a JSON-able seed plus entropy (misc.py:36-56)."""
    if isinstance(seed, Generator):
        return seed
    if isinstance(seed, BitGenerator):
        return Generator(seed)
    return Generator(PCG64(_proc_seed(seed, entropy)))


def shuffled(vals, seed=None, i=0):
    """Iterate `vals` in a seeded random order (misc.py:84-93)."""
    idxs = numpy.arange(len(vals))
    rng = create_rng(seed, f"shuffle_{i}")
    rng.shuffle(idxs)
    for idx in idxs:
        yield vals[int(idx)]


def create_sparse_groupings(shape, max_grouping=8, seed=None, i=0):
    """Random partition of the scan indices into groups (misc.py:96-115)."""
    if isinstance(shape, int):
        shape = (shape,)
    if not isinstance(shape, (tuple, list)):
        shape = shape.shape[:-1]
    idxs = numpy.indices(shape)
    idxs = idxs.reshape(idxs.shape[0], -1).T
    rng = create_rng(seed, f'groupings_{i}' if i != 0 else 'groupings')
    rng.shuffle(idxs)
    return numpy.array_split(idxs.T, numpy.ceil(idxs.shape[0] / max_grouping).astype(numpy.int64), axis=-1)


def create_compact_groupings(positions, max_grouping=8, seed=None, i=0):
    """k-means seeded, greedy size-balanced spatial groups (misc.py:118-196)."""
    from scipy.spatial import KDTree
    from scipy.cluster.vq import kmeans2

    pos_flat = positions.reshape(-1, 2)
    idxs = numpy.indices(positions.shape[:-1])
    idxs = idxs.reshape(idxs.shape[0], -1)

    n_pos = pos_flat.shape[0]
    n_groups = numpy.ceil(n_pos / max_grouping).astype(numpy.int64)
    if n_groups == 1:
        return [idxs]
    grouping = numpy.ceil(n_pos / n_groups).astype(numpy.int64)

    rng = create_rng(seed, f'groupings_{i}' if i != 0 else 'groupings')
    centroids, _labels = kmeans2(pos_flat, n_groups, minit='++', missing='raise', rng=rng)

    group_sizes = numpy.zeros(n_groups)
    assignments = numpy.full(n_pos, -1)
    unassigned = numpy.arange(n_pos)
    tree = KDTree(centroids)

    for (max_group_size, n_assigned) in ((grouping - 1, (grouping - 1) * n_groups), (grouping, n_pos)):
        n_neighbors = min(4, n_groups)
        while True:
            distances, indices = tree.query(pos_flat[unassigned], n_neighbors)
            scan_idxs = numpy.argsort(distances[..., 0])
            indices = indices[scan_idxs]
            scan_idxs = unassigned[scan_idxs]
            for k, group_idxs in zip(scan_idxs, indices):
                for group_idx in group_idxs:
                    if group_sizes[group_idx] < max_group_size:
                        assignments[k] = group_idx
                        group_sizes[group_idx] += 1
                        break
            unassigned = numpy.where(assignments == -1)[0]
            if n_pos - len(unassigned) >= n_assigned:
                break
            assert n_neighbors < n_groups
            n_neighbors = min(n_neighbors * 2, n_groups)

    group_sizes = numpy.bincount(assignments)
    assert numpy.min(group_sizes) in (grouping - 1, grouping)
    assert numpy.max(group_sizes) == grouping
    return [idxs[..., assignments == g] for g in range(n_groups)]


def mask_fraction_of_groups(n_groups, fraction):
    """Every floor(n/k)-th group from 0, k = ceil(n*fraction) (misc.py:199-208)."""
    n_required = max(1, math.ceil(n_groups * fraction))
    if n_required >= n_groups:
        return numpy.ones(n_groups, dtype=numpy.bool_)
    every = n_groups // n_required
    mask = numpy.zeros(n_groups, dtype=numpy.bool_)
    mask[::every] = 1
    return mask


class FloatKey(float):
    """Float hashed/compared at 5 decimals (misc.py:211-218)."""
    def __hash__(self):
        return float.__hash__(round(self, 5))

    def __eq__(self, other):
        return isinstance(other, float) and round(self, 5) == round(other, 5)


class GroupManager:
    """Group iterator (phaser/engines/common/simulation.py:22-53)."""

    def __init__(self, scan, grouping=None, compact=False, seed=None):
        self.grouping = grouping or 64
        self.compact = compact
        self.seed = seed
        self.groups = None
        self.n_groups = int(numpy.ceil(numpy.prod(scan.shape[:-1]) / self.grouping))

    def _make(self, scan, i=0):
        if self.compact:
            return create_compact_groupings(scan, self.grouping, seed=self.seed, i=i)
        return create_sparse_groupings(scan, self.grouping, seed=self.seed, i=i)

    def iter(self, scan, i=0, shuffle_groups=False):
        if shuffle_groups or self.groups is None:
            self.groups = self._make(scan, i)
            return iter(self.groups)
        return shuffled(self.groups, seed=self.seed, i=i)

    def __len__(self):
        return self.n_groups
