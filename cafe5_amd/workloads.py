"""Seeded synthetic inputs for the likelihood hot path.

Trees are Yule (pure-birth) ultrametric trees rescaled to a fixed height.  Gene
families are simulated down the tree with the linear birth-death process of
CAFE (lambda = mu), sampled exactly per gene:

    P(1 -> 0) = a,  P(1 -> c) = (1 - a)^2 a^(c - 1),  a = lt / (1 + lt)

which is the one-gene row of the transition matrix in probability.cpp:104-167;
s genes evolve independently, so a parent of size s draws Binomial(s, 1 - a)
survivors and each survivor adds a Geometric number of extra copies.

Everything is deterministic for a given seed (numpy PCG64).  Nothing here reads
the CAFE5 sources or the oracle.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

# Size buckets of the uneven-size workload: a family whose largest count is m
# runs with bound B = the first bucket >= m + max(5, m // 5).
SIZE_BUCKETS = (20, 32, 48, 64, 96, 128, 160, 200)


@dataclasses.dataclass
class Tree:
    """A rooted tree in children-before-parent order (root last).

    parent[i]    parent node index, -1 for the root
    leaf_col[i]  column of node i in a family's count row, -1 if internal
    length[i]    branch length above node i (0 for the root)
    names[i]     taxon name; interior nodes carry the concatenation of their
                 sorted leaf names, as clade.cpp:161-173 names them
    """

    parent: np.ndarray
    leaf_col: np.ndarray
    length: np.ndarray
    names: List[str]

    @property
    def n_nodes(self) -> int:
        return len(self.parent)

    @property
    def n_leaves(self) -> int:
        return int((self.leaf_col >= 0).sum())

    @property
    def root(self) -> int:
        return int(np.flatnonzero(self.parent < 0)[0])

    def children(self, u: int) -> List[int]:
        return [int(i) for i in np.flatnonzero(self.parent == u)]

    def leaf_names(self) -> List[str]:
        """Leaf names ordered by their count column."""
        out = [""] * self.n_leaves
        for i in range(self.n_nodes):
            if self.leaf_col[i] >= 0:
                out[self.leaf_col[i]] = self.names[i]
        return out

    def newick(self, labels: Optional[Sequence[str]] = None, root_length: Optional[float] = None) -> str:
        """Newick text of the tree; labels[i] replaces the branch length of node i."""

        def fmt(i: int) -> str:
            if labels is not None:
                return str(labels[i])
            return repr(float(self.length[i]))

        def rec(u: int) -> str:
            kids = self.children(u)
            if not kids:
                return f"{self.names[u]}:{fmt(u)}"
            inner = ",".join(rec(k) for k in kids)
            if u == self.root:
                tail = ""
                if labels is not None:
                    tail = f":{labels[u]}"
                elif root_length is not None:
                    tail = f":{root_length!r}"
                return f"({inner}){tail}"
            return f"({inner}):{fmt(u)}"

        return rec(self.root) + ";"

    def height(self) -> float:
        depth = np.zeros(self.n_nodes)
        for i in range(self.n_nodes - 1, -1, -1):
            p = self.parent[i]
            if p >= 0:
                depth[i] = depth[p] + self.length[i]
        return float(depth[self.leaf_col >= 0].max())


def _finish_tree(parent: List[int], length: List[float], is_leaf: List[bool]) -> Tree:
    """Reorder to children-before-parent (reverse BFS) and name nodes."""
    n = len(parent)
    kids: List[List[int]] = [[] for _ in range(n)]
    root = parent.index(-1)
    for i, p in enumerate(parent):
        if p >= 0:
            kids[p].append(i)
    # reverse of a BFS that enqueues siblings last-to-first: children precede
    # parents and siblings keep their original (descendant) order
    bfs = [root]
    k = 0
    while k < len(bfs):
        bfs.extend(kids[bfs[k]][::-1])
        k += 1
    order = bfs[::-1]
    new_index = {old: new for new, old in enumerate(order)}
    par = np.array([new_index[parent[o]] if parent[o] >= 0 else -1 for o in order], dtype=np.int32)
    ln = np.array([length[o] for o in order], dtype=np.float64)
    leaf = [is_leaf[o] for o in order]
    leaf_col = np.full(n, -1, dtype=np.int32)
    names: List[str] = [""] * n
    col = 0
    # leaves numbered in the order they appear in a left-to-right traversal
    def dfs(u_old: int):
        nonlocal col
        if is_leaf[u_old]:
            leaf_col[new_index[u_old]] = col
            col += 1
            return
        for c in kids[u_old]:
            dfs(c)
    dfs(root)
    width = max(2, len(str(col - 1)))
    for i in range(n):
        if leaf[i]:
            names[i] = f"S{int(leaf_col[i]):0{width}d}"
    # interior names: sorted concatenation of leaf names (clade.cpp:161-173)
    leafsets: List[List[str]] = [[] for _ in range(n)]
    for i in range(n):
        if leaf[i]:
            leafsets[i] = [names[i]]
    for i in range(n):
        if par[i] >= 0:
            leafsets[par[i]].extend(leafsets[i])
        if not leaf[i]:
            names[i] = "".join(sorted(leafsets[i]))
    return Tree(par, leaf_col, ln, names)


def yule_tree(n_taxa: int, seed: int, height: float = 100.0) -> Tree:
    """Seeded Yule tree with n_taxa leaves, ultrametric, root-to-tip = height."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parent = [-1]
    birth = [0.0]          # time each node's lineage started
    is_leaf = [True]
    active = [0]
    t = 0.0
    # the root splits at time 0 into two lineages
    while len(active) < n_taxa:
        if len(parent) > 1:
            t += rng.exponential(1.0 / len(active))
        u = active.pop(int(rng.integers(len(active))))
        is_leaf[u] = False
        for _ in range(2):
            parent.append(u)
            birth.append(t)
            is_leaf.append(True)
            active.append(len(parent) - 1)
    # a node's branch ends when it splits (interior) or at t_end (leaf)
    t_end = t + rng.exponential(1.0 / len(active))
    split_time = {}
    for i, p in enumerate(parent):
        if p >= 0:
            split_time[p] = birth[i]
    length = []
    for i in range(len(parent)):
        end = t_end if is_leaf[i] else split_time[i]
        length.append(0.0 if parent[i] < 0 else end - birth[i])
    scale = height / t_end
    length = [x * scale for x in length]
    return _finish_tree(parent, length, is_leaf)


def parse_newick(text: str) -> Tree:
    """Minimal Newick reader (names and branch lengths) into a Tree.  Text
    without the outer parentheses is read as the reference reads it
    (clade.cpp:334-341): a comma at the top level makes a new root whose
    first child is the clade read so far, and the clades after it are its
    siblings."""
    text = text.strip().rstrip(";")
    parent: List[int] = []
    length: List[float] = []
    raw_names: List[str] = []
    pos = 0

    def node(par: int) -> int:
        nonlocal pos
        me = len(parent)
        parent.append(par)
        length.append(0.0)
        raw_names.append("")
        if pos < len(text) and text[pos] == "(":
            pos += 1
            while True:
                node(me)
                if text[pos] == ",":
                    pos += 1
                    continue
                if text[pos] == ")":
                    pos += 1
                    break
        start = pos
        while pos < len(text) and text[pos] not in ",():":
            pos += 1
        raw_names[me] = text[start:pos]
        if pos < len(text) and text[pos] == ":":
            pos += 1
            start = pos
            while pos < len(text) and text[pos] not in ",()":
                pos += 1
            length[me] = float(text[start:pos])
        return me

    root = node(-1)
    if pos < len(text) and text[pos] == ",":
        top = len(parent)
        parent.append(-1)
        length.append(0.0)
        raw_names.append("")
        parent[root] = top
        while pos < len(text) and text[pos] == ",":
            pos += 1
            node(top)
    if pos != len(text):
        raise ValueError(f"unexpected {text[pos]!r} at offset {pos} of the Newick text")
    n = len(parent)
    is_leaf = [True] * n
    for p in parent:
        if p >= 0:
            is_leaf[p] = False
    tree = _finish_tree(parent, length, is_leaf)
    # keep the file's leaf names, re-deriving interior names from them
    order_names = {}
    kids: List[List[int]] = [[] for _ in range(n)]
    for i, p in enumerate(parent):
        if p >= 0:
            kids[p].append(i)
    col = 0

    def dfs(u: int):
        nonlocal col
        if is_leaf[u]:
            order_names[col] = raw_names[u]
            col += 1
            return
        for c in kids[u]:
            dfs(c)

    dfs(parent.index(-1))
    names = list(tree.names)
    leafsets: List[List[str]] = [[] for _ in range(n)]
    for i in range(n):
        if tree.leaf_col[i] >= 0:
            names[i] = order_names[int(tree.leaf_col[i])]
            leafsets[i] = [names[i]]
    for i in range(n):
        if tree.parent[i] >= 0:
            leafsets[tree.parent[i]].extend(leafsets[i])
        if tree.leaf_col[i] < 0:
            names[i] = "".join(sorted(leafsets[i]))
    tree.names = names
    return tree


def birth_death_alpha(lam: float, t: float) -> float:
    lt = lam * t
    return lt / (1.0 + lt)


def simulate_families(tree: Tree, root_sizes: np.ndarray, node_lambda: np.ndarray, seed: int,
                      cap: Optional[int] = None) -> np.ndarray:
    """Counts (F x n_leaves, int32) simulated down the tree from root_sizes.

    node_lambda[i] is the birth-death rate on the branch above node i (it may be
    an (F, n_nodes) array for per-family rates, e.g. gamma categories).
    Counts are clipped to cap when given.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    F = len(root_sizes)
    n = tree.n_nodes
    size = np.zeros((n, F), dtype=np.int64)
    size[tree.root] = root_sizes
    per_family = np.ndim(node_lambda) == 2
    for u in range(n - 1, -1, -1):      # parents after children -> walk backwards
        p = tree.parent[u]
        if p < 0:
            continue
        lam = node_lambda[:, u] if per_family else node_lambda[u]
        a = birth_death_alpha(lam, tree.length[u])
        s = size[p]
        surv = rng.binomial(s, 1.0 - a)
        extra = np.where(surv > 0, rng.negative_binomial(np.maximum(surv, 1), 1.0 - a), 0)
        size[u] = surv + extra
    counts = np.zeros((F, tree.n_leaves), dtype=np.int32)
    for u in range(n):
        if tree.leaf_col[u] >= 0:
            counts[:, tree.leaf_col[u]] = size[u]
    if cap is not None:
        np.minimum(counts, cap, out=counts)
    return counts


def bucket_bound(max_count: int) -> int:
    need = max_count + max(5, max_count // 5)
    for b in SIZE_BUCKETS:
        if b >= need:
            return b
    return SIZE_BUCKETS[-1]


def lambda_grid(n: int = 64, lo: float = 1e-4, hi: float = 1e-2) -> np.ndarray:
    return lo * (hi / lo) ** (np.arange(n) / (n - 1))


@dataclasses.dataclass
class Workload:
    name: str
    tree: Tree
    counts: np.ndarray                 # F x n_leaves int32
    bounds: np.ndarray                 # per-family B (== R)
    lambdas: np.ndarray                # trial points (T,) or class values
    node_class: Optional[np.ndarray] = None   # lambda class per node (multi-lambda)
    n_cat: int = 0
    alpha: float = 0.0
    epsilon: float = 0.0
    alphas: Optional[np.ndarray] = None       # gamma shape of each trial point (optimizer runs)


def config_base(seed: int = 1, F: int = 2000) -> Workload:
    """12 taxa, F families, root ~ 1 + Poisson(4) (mean 5), lambda 0.002, B = R = 100."""
    tree = yule_tree(12, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    roots = 1 + rng.poisson(4.0, F)
    counts = simulate_families(tree, roots, np.full(tree.n_nodes, 0.002), seed + 2000, cap=100)
    return Workload("base-12taxa", tree, counts, np.full(F, 100, np.int32), np.array([0.002]))


def heavy_tail_roots(rng: np.random.Generator, F: int, body_max: int = 10, tail_max: int = 150) -> np.ndarray:
    """85% of roots in 1..body_max, the rest Pareto-tailed up to tail_max."""
    body = 1 + np.minimum(rng.poisson(3.0, F), body_max - 1)
    tail = np.minimum(body_max + 1 + np.floor(rng.pareto(1.2, F) * 8).astype(np.int64), tail_max)
    return np.where(rng.random(F) < 0.85, body, tail)


def config_uneven(seed: int = 2, F: int = 20000, T: int = 64) -> Workload:
    """34 taxa, F families with heavy-tailed roots, per-family bucket bounds
    20..200, T-point lambda grid 1e-4..1e-2."""
    tree = yule_tree(34, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    roots = heavy_tail_roots(rng, F)
    counts = simulate_families(tree, roots, np.full(tree.n_nodes, 0.003), seed + 2000, cap=190)
    bounds = np.array([bucket_bound(int(m)) for m in counts.max(axis=1)], dtype=np.int32)
    return Workload("uneven-34taxa", tree, counts, bounds, lambda_grid(T))


def config_gamma(seed: int = 3, F: int = 50000, K: int = 4, alpha: float = 0.5, lam: float = 0.003,
                 bound: int = 100) -> Workload:
    """34 taxa, F families, K gamma categories at shape alpha, lambda 0.003."""
    tree = yule_tree(34, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    roots = 1 + np.minimum(rng.poisson(4.0, F), 40)
    mult = rng.gamma(alpha, 1.0 / alpha, F)
    node_lam = np.outer(np.minimum(mult * lam, 0.009), np.ones(tree.n_nodes))
    counts = simulate_families(tree, roots, node_lam, seed + 2000, cap=bound - 1)
    return Workload("gamma-34taxa", tree, counts, np.full(F, bound, np.int32), np.array([lam]),
                    n_cat=K, alpha=alpha)


def clade_classes(tree: Tree, n_classes: int, seed: int) -> np.ndarray:
    """Seeded lambda-class rule: each of n_classes-1 random interior clades
    (and everything under it) takes its own class; the rest is class 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cls = np.zeros(tree.n_nodes, dtype=np.int32)
    interior = [i for i in range(tree.n_nodes) if tree.leaf_col[i] < 0 and i != tree.root]
    picks = rng.choice(interior, size=n_classes - 1, replace=False)
    for c, top in enumerate(sorted(picks, key=lambda i: -i), start=1):
        stack = [int(top)]
        while stack:
            u = stack.pop()
            cls[u] = c
            stack.extend(tree.children(u))
    return cls


def config_error_multi(seed: int = 4, F: int = 100000, bound: int = 150) -> Workload:
    """60 taxa, 3 lambda classes (0.001, 0.004, 0.01), error model eps = 0.05."""
    tree = yule_tree(60, seed)
    cls = clade_classes(tree, 3, seed + 7)
    lams = np.array([0.001, 0.004, 0.01])
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    roots = 1 + np.minimum(rng.poisson(5.0, F), 60)
    counts = simulate_families(tree, roots, lams[cls], seed + 2000, cap=bound - 1)
    return Workload("errmodel-60taxa", tree, counts, np.full(F, bound, np.int32), lams,
                    node_class=cls, epsilon=0.05)


def config_optimizer(seed: int = 5, F: int = 200000, K: int = 8, bound: int = 500, T: int = 200) -> Workload:
    """100 taxa, F families, bound 500, K = 8, T seeded (lambda, alpha) trial
    points: lambda log-uniform in [1e-3, 5e-3], alpha log-uniform in [0.2, 2]."""
    tree = yule_tree(100, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    roots = heavy_tail_roots(rng, F, body_max=10, tail_max=300)
    counts = simulate_families(tree, roots, np.full(tree.n_nodes, 0.002), seed + 2000, cap=bound - 1)
    lam = np.exp(rng.uniform(np.log(1e-3), np.log(5e-3), T))
    alphas = np.exp(rng.uniform(np.log(0.2), np.log(2.0), T))
    return Workload("optimizer-100taxa", tree, counts, np.full(F, bound, np.int32), lam, n_cat=K,
                    alpha=1.0, alphas=alphas)


def error_table(epsilon: float, B: int) -> np.ndarray:
    """Rows 0..B of the symmetric +-1 leaf error model: row 0 = {0, 1-eps, eps},
    rows >= 1 = {eps, 1-2eps, eps} (core.cpp:39-44 with eps = 0.05)."""
    tab = np.empty((B + 1, 3))
    tab[0] = (0.0, 1.0 - epsilon, epsilon)
    tab[1:] = (epsilon, 1.0 - 2 * epsilon, epsilon)
    return tab
