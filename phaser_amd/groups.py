"""Host scheduling of scan positions into groups, unchanged from the reference so that the
group order of a seeded run is identical: SHA-256 seeded PCG64 (phaser/utils/misc.py:14-56),
sparse / compact groupings (:96-196), shuffles (:84-93), error-group masks (:199-208) and
GroupManager (phaser/engines/common/simulation.py:22-53)."""
import hashlib
import json
import math

import numpy
from numpy.random import SeedSequence, PCG64, Generator, BitGenerator


def _proc_seed(seed, entropy=None):
    if seed is None:
        return SeedSequence()
    if isinstance(seed, SeedSequence):
        seed = seed.entropy
    h = hashlib.sha256()
    h.update(json.dumps(seed).encode('utf-8'))
    if entropy is not None:
        h.update(json.dumps(entropy).encode('utf-8'))
    return SeedSequence(numpy.frombuffer(h.digest(), dtype=numpy.uint32))


def create_rng(seed=None, entropy=None) -> Generator:
    if isinstance(seed, Generator):
        return seed
    if isinstance(seed, BitGenerator):
        return Generator(seed)
    return Generator(PCG64(_proc_seed(seed, entropy)))


def shuffled(vals, seed=None, i=0):
    idxs = numpy.arange(len(vals))
    create_rng(seed, f"shuffle_{i}").shuffle(idxs)
    for idx in idxs:
        yield vals[int(idx)]


def create_sparse_groupings(shape, max_grouping=8, seed=None, i=0):
    if isinstance(shape, int):
        shape = (shape,)
    if not isinstance(shape, (tuple, list)):
        shape = shape.shape[:-1]
    idxs = numpy.indices(shape)
    idxs = idxs.reshape(idxs.shape[0], -1).T
    create_rng(seed, f'groupings_{i}' if i != 0 else 'groupings').shuffle(idxs)
    return numpy.array_split(idxs.T, numpy.ceil(idxs.shape[0] / max_grouping).astype(numpy.int64), axis=-1)


def create_compact_groupings(positions, max_grouping=8, seed=None, i=0):
    """k-means seeded, greedily size-balanced spatial groups (misc.py:118-196)."""
    from scipy.spatial import KDTree
    from scipy.cluster.vq import kmeans2

    pos_flat = positions.reshape(-1, 2)
    idxs = numpy.indices(positions.shape[:-1]).reshape(positions.ndim - 1, -1)
    n_pos = pos_flat.shape[0]
    n_groups = numpy.ceil(n_pos / max_grouping).astype(numpy.int64)
    if n_groups == 1:
        return [idxs]
    grouping = numpy.ceil(n_pos / n_groups).astype(numpy.int64)
    rng = create_rng(seed, f'groupings_{i}' if i != 0 else 'groupings')
    centroids, _ = kmeans2(pos_flat, n_groups, minit='++', missing='raise', rng=rng)
    sizes = numpy.zeros(n_groups)
    assign = numpy.full(n_pos, -1)
    unassigned = numpy.arange(n_pos)
    tree = KDTree(centroids)
    for (max_size, n_assigned) in ((grouping - 1, (grouping - 1) * n_groups), (grouping, n_pos)):
        n_nb = min(4, n_groups)
        while True:
            dist, ind = tree.query(pos_flat[unassigned], n_nb)
            order = numpy.argsort(dist[..., 0])
            ind = ind[order]
            cand = unassigned[order]
            for k, gidx in zip(cand, ind):
                for g in gidx:
                    if sizes[g] < max_size:
                        assign[k] = g
                        sizes[g] += 1
                        break
            unassigned = numpy.where(assign == -1)[0]
            if n_pos - len(unassigned) >= n_assigned:
                break
            assert n_nb < n_groups
            n_nb = min(n_nb * 2, n_groups)
    return [idxs[..., assign == g] for g in range(n_groups)]


def mask_fraction_of_groups(n_groups, fraction):
    n_required = max(1, math.ceil(n_groups * fraction))
    if n_required >= n_groups:
        return numpy.ones(n_groups, dtype=numpy.bool_)
    mask = numpy.zeros(n_groups, dtype=numpy.bool_)
    mask[::n_groups // n_required] = 1
    return mask


class GroupManager:
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

    def max_size(self):
        """Largest group this manager can produce (array_split / compact sizes)."""
        return self.grouping
