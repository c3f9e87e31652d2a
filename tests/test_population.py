"""CPU tests of the NSGA-II population driver (bqp_amd.population): ranking
against brute force, crowding, operator bounds, and convergence on the
ZDT1 benchmark whose Pareto front is known in closed form."""
import numpy as np
import pytest

from bqp_amd import population as P


def _brute_ranks(F):
    n = F.shape[0]
    dom = np.zeros((n, n), bool)
    for i in range(n):
        for j in range(n):
            dom[i, j] = np.all(F[i] <= F[j]) and np.any(F[i] < F[j])
    ranks = np.full(n, -1)
    left = set(range(n))
    r = 0
    while left:
        front = [j for j in left if not any(dom[i, j] for i in left)]
        for j in front:
            ranks[j] = r
        left -= set(front)
        r += 1
    return ranks


@pytest.mark.parametrize("m", [2, 3])
def test_ranks_match_brute_force(m):
    rng = np.random.default_rng(m)
    F = rng.integers(0, 6, (120, m)).astype(float)  # many ties and duplicates
    assert np.array_equal(P.nondominated_ranks(F), _brute_ranks(F))
    G = rng.random((200, m))
    assert np.array_equal(P.nondominated_ranks(G), _brute_ranks(G))


def test_crowding_boundaries_and_order():
    F = np.array([[0.0, 4.0], [1.0, 3.0], [2.0, 1.0], [4.0, 0.0]])
    r = P.nondominated_ranks(F)
    assert np.all(r == 0)
    d = P.crowding_distance(F, r)
    assert np.isinf(d[0]) and np.isinf(d[3])
    # interior: (2-0)/4 + (4-1)/4 = 1.25 and (4-1)/4 + (3-0)/4 = 1.5
    assert np.isclose(d[1], 1.25) and np.isclose(d[2], 1.5)


def test_operators_stay_in_bounds():
    rng = np.random.Generator(np.random.PCG64(1))
    lo, hi = np.array([0.0, -50.0, 0.5]), np.array([10.0, 50.0, 10.0])
    a = lo + rng.random((500, 3)) * (hi - lo)
    b = lo + rng.random((500, 3)) * (hi - lo)
    c1, c2 = P.sbx(a, b, lo, hi, 20.0, 0.9, rng)
    m = P.polynomial_mutation(c1, lo, hi, 20.0, 1.0, rng)
    for x in (c1, c2, m):
        assert np.all(x >= lo) and np.all(x <= hi)
    # crossover preserves the parents' mean per variable (symmetric spread)
    assert np.allclose((c1 + c2).mean(axis=0), (a + b).mean(axis=0), rtol=0.05)


def test_nsga2_converges_on_zdt1():
    n = 30

    def zdt1(X):
        f1 = X[:, 0]
        g = 1.0 + 9.0 * X[:, 1:].mean(axis=1)
        return np.column_stack([f1, g * (1.0 - np.sqrt(f1 / g))])

    opt = P.Nsga2(zdt1, np.zeros(n), np.ones(n), pop_size=100, seed=3, mutation_probability=1.0 / n)
    opt.run(250)
    _, F = opt.front()
    gap = F[:, 1] - (1.0 - np.sqrt(F[:, 0]))  # distance above the true front f2 = 1 - sqrt(f1)
    assert opt.evaluations == 100 * 251
    assert np.median(gap) < 0.02 and F[:, 0].max() - F[:, 0].min() > 0.8
