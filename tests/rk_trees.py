"""Butcher rooted-tree order conditions (Hairer, Norsett & Wanner, Solving ODEs I,
Sec. II.2, Theorem 2.13): an explicit RK method has order p iff
    b^T Phi(t) = 1 / gamma(t)   for every rooted tree t with |t| <= p,
where Phi(t) = prod_i (A Phi(t_i)) over the root's subtrees t_i (elementwise),
Phi(single node) = 1, gamma(t) = |t| * prod_i gamma(t_i).

Trees are canonical sorted tuples of subtrees; () is the single node.
"""
from __future__ import annotations

from functools import lru_cache
from itertools import combinations_with_replacement

import numpy as np


@lru_cache(maxsize=None)
def trees(order: int):
    """All rooted trees with exactly ``order`` nodes (canonical form)."""
    if order == 1:
        return [()]
    out = set()
    # the root's children form a multiset of trees whose sizes sum to order-1
    for parts in _partitions(order - 1):
        pools = [trees(p) for p in parts]
        for combo in _product_multiset(parts, pools):
            out.add(tuple(sorted(combo)))
    return sorted(out)


def _partitions(n, maxpart=None):
    maxpart = n if maxpart is None else maxpart
    if n == 0:
        yield ()
        return
    for k in range(min(n, maxpart), 0, -1):
        for rest in _partitions(n - k, k):
            yield (k,) + rest


def _product_multiset(parts, pools):
    # group equal sizes so subtrees of equal size are chosen as multisets
    groups = {}
    for p, pool in zip(parts, pools):
        groups.setdefault(p, [pool, 0])
        groups[p][1] += 1
    choices = []
    for p, (pool, cnt) in groups.items():
        choices.append(list(combinations_with_replacement(pool, cnt)))
    def rec(i):
        if i == len(choices):
            yield ()
            return
        for c in choices[i]:
            for rest in rec(i + 1):
                yield tuple(c) + rest
    yield from rec(0)


def size(t) -> int:
    return 1 + sum(size(c) for c in t)


def gamma(t) -> int:
    g = size(t)
    for c in t:
        g *= gamma(c)
    return g


def phi(t, A):
    v = np.ones(A.shape[0])
    for c in t:
        v = v * (A @ phi(c, A))
    return v


def residuals(A, b, order):
    """max |b.Phi(t) - 1/gamma(t)| per order 1..order."""
    out = []
    for p in range(1, order + 1):
        out.append(max(abs(float(b @ phi(t, A)) - 1.0 / gamma(t)) for t in trees(p)))
    return out
