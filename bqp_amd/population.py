"""Population-evaluation driver: NSGA-II generations fed to the GPU batch
propagator.

The reference's transfer optimiser (crates/two_phase_transfer_rs/src/oxymoo/
nsga2.rs:19-120 state, :151-211 generation step, :354-420 offspring;
defaults oxymoo/types.rs:146-163) evaluates each generation's candidates
one by one on the CPU.  Here a generation is one device call: the whole
population goes to ``Propagator.propagate_batch_final`` as a ragged batch
(per-row durations), and the context's field (and surrogate, if any) stays
resident between generations.

NSGA-II itself is the published algorithm (Deb, Pratap, Agarwal & Meyarivan,
IEEE TEVC 6(2), 2002) with the reference's operator defaults: binary
tournament on (rank, crowding), simulated binary crossover (probability 0.9,
eta_c 20, bounded form of Deb & Agrawal 1995), polynomial mutation
(probability 1/n_var, eta_m 20), elitist (mu + lambda) survival by rank then
crowding distance.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

DAY = 86400.0
FAILED_MISS_KM = 1e9  # objective of a row whose propagation failed


# ------------------------------------------------------------ ranking
def nondominated_ranks(F: np.ndarray) -> np.ndarray:
    """Pareto rank (0 = non-dominated) of each row of F (minimisation)."""
    F = np.asarray(F, dtype=np.float64)
    n, m = F.shape
    if n == 0:
        return np.zeros(0, dtype=np.int64)
    if m == 2:
        return _ranks_2d(F)
    # general case: domination matrix, then peel fronts
    le = np.all(F[:, None, :] <= F[None, :, :], axis=2)
    lt = np.any(F[:, None, :] < F[None, :, :], axis=2)
    dom = le & lt  # dom[i, j]: i dominates j
    count = dom.sum(axis=0)
    ranks = np.full(n, -1, dtype=np.int64)
    r = 0
    front = np.nonzero(count == 0)[0]
    while front.size:
        ranks[front] = r
        count = count - dom[front].sum(axis=0)
        count[ranks >= 0] = -1
        front = np.nonzero(count == 0)[0]
        r += 1
    return ranks


def _ranks_2d(F: np.ndarray) -> np.ndarray:
    """Two objectives: sweep in (f1, f2) order; front k keeps the smallest
    f2 seen so far, and these minima increase with k, so a point's front is
    the first whose minimum exceeds its f2 (binary search).  Identical
    points share a rank (neither dominates the other)."""
    uniq, inv = np.unique(F, axis=0, return_inverse=True)  # lexicographic (f1, f2) order
    inv = inv.reshape(-1)
    mins: list[float] = []
    ranks_u = np.empty(uniq.shape[0], dtype=np.int64)
    import bisect
    for i in range(uniq.shape[0]):
        f2 = uniq[i, 1]
        k = bisect.bisect_right(mins, f2)  # first front whose minimum > f2
        if k == len(mins):
            mins.append(f2)
        else:
            mins[k] = f2
        ranks_u[i] = k
    return ranks_u[inv]


def crowding_distance(F: np.ndarray, ranks: np.ndarray) -> np.ndarray:
    """Crowding distance within each front (boundary points: +inf)."""
    F = np.asarray(F, dtype=np.float64)
    n, m = F.shape
    d = np.zeros(n)
    for r in np.unique(ranks):
        idx = np.nonzero(ranks == r)[0]
        if idx.size <= 2:
            d[idx] = np.inf
            continue
        for j in range(m):
            order = idx[np.argsort(F[idx, j], kind="stable")]
            fmin, fmax = F[order[0], j], F[order[-1], j]
            d[order[0]] = d[order[-1]] = np.inf
            if fmax > fmin:
                d[order[1:-1]] += (F[order[2:], j] - F[order[:-2], j]) / (fmax - fmin)
    return d


# ---------------------------------------------------------- operators
def sbx(p1: np.ndarray, p2: np.ndarray, lo: np.ndarray, hi: np.ndarray, eta: float, prob: float,
        rng: np.random.Generator):
    """Bounded simulated binary crossover of parent rows (k, n) -> two children."""
    c1, c2 = p1.copy(), p2.copy()
    k, nv = p1.shape
    do = rng.random(k) <= prob
    for j in range(nv):
        swap = do & (rng.random(k) <= 0.5) & (np.abs(p1[:, j] - p2[:, j]) > 1e-14)
        if not np.any(swap):
            continue
        y1 = np.minimum(p1[swap, j], p2[swap, j])
        y2 = np.maximum(p1[swap, j], p2[swap, j])
        span = y2 - y1
        u = rng.random(y1.size)
        inv = 1.0 / (eta + 1.0)

        def betaq(beta):
            alpha = 2.0 - beta ** (-(eta + 1.0))
            return np.where(u <= 1.0 / alpha, (u * alpha) ** inv, (1.0 / (2.0 - u * alpha)) ** inv)

        b1 = np.maximum(1.0 + 2.0 * (y1 - lo[j]) / span, 1e-14)
        b2 = np.maximum(1.0 + 2.0 * (hi[j] - y2) / span, 1e-14)
        ch1 = np.clip(0.5 * ((y1 + y2) - betaq(b1) * span), lo[j], hi[j])
        ch2 = np.clip(0.5 * ((y1 + y2) + betaq(b2) * span), lo[j], hi[j])
        flip = rng.random(y1.size) <= 0.5
        c1[swap, j] = np.where(flip, ch2, ch1)
        c2[swap, j] = np.where(flip, ch1, ch2)
    return c1, c2


def polynomial_mutation(x: np.ndarray, lo: np.ndarray, hi: np.ndarray, eta: float, prob: float,
                        rng: np.random.Generator) -> np.ndarray:
    """Bounded polynomial mutation of rows (k, n), each variable with `prob`."""
    y = x.copy()
    k, nv = x.shape
    for j in range(nv):
        sel = rng.random(k) <= prob
        if not np.any(sel):
            continue
        v = y[sel, j]
        span = hi[j] - lo[j]
        d1 = (v - lo[j]) / span
        d2 = (hi[j] - v) / span
        u = rng.random(v.size)
        pw = 1.0 / (eta + 1.0)
        lo_side = u < 0.5
        xy1 = 1.0 - d1
        val1 = 2.0 * u + (1.0 - 2.0 * u) * xy1 ** (eta + 1.0)
        dq1 = np.abs(val1) ** pw - 1.0
        xy2 = 1.0 - d2
        val2 = 2.0 * (1.0 - u) + 2.0 * (u - 0.5) * xy2 ** (eta + 1.0)
        dq2 = 1.0 - np.abs(val2) ** pw
        y[sel, j] = np.clip(v + np.where(lo_side, dq1, dq2) * span, lo[j], hi[j])
    return y


def tournament(ranks: np.ndarray, crowd: np.ndarray, k: int, rng: np.random.Generator) -> np.ndarray:
    """k binary-tournament winners by (lower rank, then larger crowding)."""
    n = ranks.size
    a = rng.integers(0, n, k)
    b = rng.integers(0, n, k)
    a_wins = (ranks[a] < ranks[b]) | ((ranks[a] == ranks[b]) & (crowd[a] >= crowd[b]))
    return np.where(a_wins, a, b)


def survivors(F: np.ndarray, size: int):
    """Elitist selection of `size` rows: whole fronts by rank, the split front
    by descending crowding distance.  Returns (rows, ranks, crowding) of the
    combined set."""
    ranks = nondominated_ranks(F)
    crowd = crowding_distance(F, ranks)
    order = np.lexsort((-crowd, ranks))
    return order[:size], ranks, crowd


@dataclass
class Nsga2:
    """NSGA-II over a batch evaluator ``evaluate(X) -> F`` (rows = candidates)."""

    evaluate: object
    lower: np.ndarray
    upper: np.ndarray
    pop_size: int = 4096
    seed: int = 0
    crossover_probability: float = 0.9
    eta_c: float = 20.0
    eta_m: float = 20.0
    mutation_probability: float | None = None
    initial: np.ndarray | None = None
    X: np.ndarray = field(init=False, default=None)
    F: np.ndarray = field(init=False, default=None)
    generation: int = field(init=False, default=0)
    evaluations: int = field(init=False, default=0)

    def __post_init__(self):
        self.lower = np.asarray(self.lower, dtype=np.float64)
        self.upper = np.asarray(self.upper, dtype=np.float64)
        if self.lower.shape != self.upper.shape or np.any(self.upper <= self.lower):
            raise ValueError("need lower < upper for every variable")
        self.rng = np.random.Generator(np.random.PCG64(self.seed))
        nv = self.lower.size
        if self.initial is not None:
            X = np.asarray(self.initial, dtype=np.float64).reshape(-1, nv)[: self.pop_size]
            if X.shape[0] < self.pop_size:
                fill = self.lower + self.rng.random((self.pop_size - X.shape[0], nv)) * (self.upper - self.lower)
                X = np.vstack([X, fill])
        else:
            X = self.lower + self.rng.random((self.pop_size, nv)) * (self.upper - self.lower)
        self.X = np.clip(X, self.lower, self.upper)
        self.F = np.asarray(self.evaluate(self.X), dtype=np.float64)
        self.evaluations = self.pop_size
        self.ranks = nondominated_ranks(self.F)
        self.crowd = crowding_distance(self.F, self.ranks)

    def offspring(self) -> np.ndarray:
        nv = self.lower.size
        pm = self.mutation_probability if self.mutation_probability is not None else 1.0 / nv
        half = (self.pop_size + 1) // 2
        a = self.X[tournament(self.ranks, self.crowd, half, self.rng)]
        b = self.X[tournament(self.ranks, self.crowd, half, self.rng)]
        c1, c2 = sbx(a, b, self.lower, self.upper, self.eta_c, self.crossover_probability, self.rng)
        kids = np.vstack([c1, c2])[: self.pop_size]
        return polynomial_mutation(kids, self.lower, self.upper, self.eta_m, pm, self.rng)

    def step(self):
        Xo = self.offspring()
        Fo = np.asarray(self.evaluate(Xo), dtype=np.float64)
        self.evaluations += Xo.shape[0]
        Xc = np.vstack([self.X, Xo])
        Fc = np.vstack([self.F, Fo])
        rows, _, _ = survivors(Fc, self.pop_size)
        self.X, self.F = Xc[rows], Fc[rows]
        self.ranks = nondominated_ranks(self.F)
        self.crowd = crowding_distance(self.F, self.ranks)
        self.generation += 1

    def run(self, generations: int, callback=None):
        """Step until `generations` offspring generations are done;
        callback(self) after each."""
        while self.generation < generations:
            self.step()
            if callback is not None:
                callback(self)
        return self

    def front(self):
        sel = self.ranks == 0
        return self.X[sel], self.F[sel]


# ------------------------------------------------- departure problem
class DepartureProblem:
    """Synthetic transfer-style problem over seeded LEO departure states.

    Decision vector: (departure index k in [0, K), dv_x, dv_y, dv_z in
    [-dv_max, dv_max] m/s (inertial, applied at t0), coast time T in
    [t_min, t_max] days).  Objectives (minimise): |dv| (m/s) and
    | |r(T)| - r_target | (km) -- reach a target radius at the end of the
    coast.  Each generation is one ragged batch on the GPU (rows = candidates,
    per-row tf = T, per-row drag coefficient of the departure state).  Rows
    that fail (with the reference's ground event on the propagator, a
    candidate that re-enters: FinalPropagationFailure::Ground,
    two_phase_transfer_rs/src/evaluate.rs:1044-1060) lose.
    """

    def __init__(self, prop, states: np.ndarray, bc: np.ndarray | None, r_target: float,
                 dv_max: float = 50.0, t_min_days: float = 0.5, t_max_days: float = 10.0):
        self.prop = prop
        self.states = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, 6)
        self.bc = None if bc is None else np.ascontiguousarray(bc, dtype=np.float64)
        self.r_target = float(r_target)
        self.lower = np.array([0.0, -dv_max, -dv_max, -dv_max, t_min_days])
        self.upper = np.array([self.states.shape[0] - 1e-9, dv_max, dv_max, dv_max, t_max_days])
        self.calls = 0
        self.orbit_days = 0.0
        self.device_s = 0.0
        self.evals = 0
        self.last = None

    def initial_population(self, n: int, rng: np.random.Generator) -> np.ndarray:
        """Seeded departure states (one each, in order) with U(-dv_max, dv_max)
        per axis and U(t_min, t_max) coast times."""
        k = np.arange(n) % self.states.shape[0]
        dv = rng.uniform(self.lower[1], self.upper[1], (n, 3))
        T = rng.uniform(self.lower[4], self.upper[4], n)
        return np.column_stack([k + 0.5, dv, T])

    def candidates(self, X: np.ndarray):
        k = np.floor(X[:, 0]).astype(np.int64).clip(0, self.states.shape[0] - 1)
        y0 = self.states[k].copy()
        y0[:, 3:6] += X[:, 1:4]
        tf = X[:, 4] * DAY
        bc = None if self.bc is None else self.bc[k]
        return y0, tf, bc

    def __call__(self, X: np.ndarray) -> np.ndarray:
        import time

        y0, tf, bc = self.candidates(X)
        t = time.perf_counter()
        yf, status, info = self.prop.propagate_batch_final(y0, 0.0, tf, bc)
        self.device_s += time.perf_counter() - t
        self.calls += 1
        # only arcs that reach their coast time count: a row that ends early
        # (the ground event) is not charged its requested span
        self.orbit_days += float(np.sum(tf[status == 0]) / DAY)
        self.evals += int(info["evals"])
        self.last = (yf, status, info)
        dv = np.linalg.norm(X[:, 1:4], axis=1)
        miss = np.abs(np.linalg.norm(yf[:, :3], axis=1) - self.r_target) / 1e3
        miss = np.where(status == 0, miss, FAILED_MISS_KM)  # failed rows (e.g. re-entry) lose
        return np.column_stack([dv, miss])
