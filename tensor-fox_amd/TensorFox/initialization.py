"""MLSVD-based starting points 'smart_random' and 'smart' (Initialization.py:97-358).

Both build the start from a few entries of the compressed core S (at most R entries per sample):
they are index selection on the core, read back once from the device (the core is the small
tensor of the pipeline), and their output goes through the same clean-up, equalization and
transform as the 'random' start.

smart_random (Initialization.py:97-141) draws 1 + int(sqrt(prod(dims))) samples of R core
coordinates from the triangular distribution of smart_sample (:144-176, 179-191) and keeps the
sample whose rank-R tensor sum_r S[c_r] e_{c_r} is closest to S.  The candidate's error is
evaluated from the R selected entries only: with m_c the number of times coordinate c is drawn,
|S - S_init|^2 = (|S|^2 - sum_c S_c^2) + sum_c ((1 - m_c) S_c)^2 over the distinct drawn c, which
is |S1 - S1_init| of the reference without forming S1_init (O(R) per sample instead of O(|S|)).

smart (Initialization.py:194-255) takes the R entries of largest magnitude, then merges and
extends them with join_rank1 (:258-324) and find_pivot (:327-358); these are restated step for
step, including the order in which join_rank1 visits and removes indexes, because that order
decides which entries end up in which rank-one term.  The entries are ranked by one stable sort
of |S| (descending; equal magnitudes in increasing flat index), the order the reference's
repeated argmax produces.
"""

import math

import numpy as np


def _draw_coordinates(dims, R, rng):
    """One smart_sample draw: the R core coordinates (R x L int array), Initialization.py:
    144-191.  For mode l the weight of index i is dims[l] - i; randint(sum of weights) is mapped
    to the index whose cumulative-weight interval contains it."""
    L = len(dims)
    out = np.empty((R, L), dtype=np.int64)
    for l in range(L):
        w = dims[l] - np.arange(dims[l], dtype=np.float64)
        edges = np.cumsum(w)
        c = rng.randint(int(edges[-1]), size=R)
        out[:, l] = np.searchsorted(edges, c, side='right')
    return out


def smart_random(S, dims, R, rng=None):
    """Best of 1 + int(sqrt(prod(dims))) sampled starts (Initialization.py:97-141).  S: host core
    (FP64, C order, shape dims).  Draws from the global NumPy generator in the reference's order
    (per sample, per mode, R integers)."""
    rng = np.random.mtrand._rand if rng is None else rng
    dims = [int(d) for d in dims]
    L = len(dims)
    S = np.asarray(S, dtype=np.float64).reshape(dims)
    samples = 1 + int(math.sqrt(int(np.prod(dims))))
    total = float(np.sum(S * S))
    Ssize = math.sqrt(total)
    best_error, best = np.inf, None
    for _ in range(samples):
        coords = _draw_coordinates(dims, R, rng)
        flat = np.ravel_multi_index(tuple(coords.T), dims)
        uniq, mult = np.unique(flat, return_counts=True)
        v = S.ravel()[uniq]
        err2 = (total - float(np.dot(v, v))) + float(np.sum(((1 - mult) * v) ** 2))
        rel = math.sqrt(max(err2, 0.0)) / Ssize
        if rel < best_error:
            best_error, best = rel, coords
    F = [np.zeros((dims[l], R)) for l in range(L)]
    for r in range(R):
        c = tuple(int(x) for x in best[r])
        for l in range(L):
            F[l][c[l], r] = 1.0
        F[0][c[0], r] = S[c]
    return F


def join_rank1(idx):
    """Initialization.py:258-324: groups of indexes that differ in one coordinate only (one
    rank-one term each).  Returns (number of indexes merged into a group, the groups)."""
    n = len(idx)
    L = len(idx[0])
    count = 0
    tmp = list(idx)
    groups = [[] for _ in range(L * n)]
    g = 0
    for l in range(L):
        keep = [j for j in range(L) if j != l]
        r = 0
        while r < len(tmp):
            key = tuple(tmp[r][j] for j in keep)
            rr = r + 1
            while rr < len(tmp):
                if tuple(tmp[rr][j] for j in keep) == key:
                    groups[g].append(tmp[rr])
                    del tmp[rr]
                    count += 1
                # the index that moved into position rr is not compared (as in the reference)
                rr += 1
            if groups[g]:
                groups[g].append(tmp[r])
                del tmp[r]
            g += 1
            r += 1
    return count, [grp for grp in groups if grp]


def find_pivot(x):
    """Initialization.py:327-358: the coordinate in which the indexes of one group differ."""
    if len(x) == 1:
        return 0
    for l in range(len(x[0])):
        if x[0][l] != x[1][l]:
            return l
    return None


def smart(S, dims, R):
    """Deterministic start from the largest core entries (Initialization.py:194-255)."""
    dims = [int(d) for d in dims]
    L = len(dims)
    S = np.asarray(S, dtype=np.float64).reshape(dims)
    order = np.argsort(-np.abs(S).ravel(), kind='stable')
    coords = np.stack(np.unravel_index(order, dims), axis=1)
    ranked = [tuple(int(v) for v in c) for c in coords]
    chosen, rest = ranked[:R], ranked[R:]
    prev = 0
    for _ in range(100):
        count, groups = join_rank1(chosen)
        if count - prev <= 0:
            break
        chosen = chosen + rest[:count - prev]
        rest = rest[count - prev:]
        prev = count
    terms = [[] for _ in range(R)]
    r = 0
    for grp in groups:
        for y in grp:
            terms[r].append(y)
            chosen.remove(y)
        r += 1
    for x in chosen:
        terms[r].append(x)
        r += 1
    F = [np.zeros((dims[l], R)) for l in range(L)]
    for r in range(R):
        pivot = find_pivot(terms[r])
        for x in terms[r]:
            for l in range(L):
                if l == pivot:
                    F[l][x[l], r] += S[x]
                else:
                    F[l][x[l], r] = 1.0
    return F
