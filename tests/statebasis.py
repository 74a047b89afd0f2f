"""Helpers of the state-basis (non-reversible and mixture model) tests: random
rate matrices and their complex eigensystems, an op planner that knows the
root (children toward the root, the root leaf), and brute-force summation
over internal states with transition matrices from scipy's expm, which shares
nothing with the eigen route the kernels take."""
import itertools

import numpy as np
from scipy.linalg import expm

from iqtree2_amd.engine import OP_DTYPE, end as End


def stationary(Q):
    w, vl = np.linalg.eig(Q.T)
    v = np.real(vl[:, np.argmin(np.abs(w))])
    return v / v.sum()


def random_q(n, rng, reversible=False):
    """A rate matrix normalised to one substitution per unit time at its
    stationary distribution; reversible = GTR-shaped."""
    if reversible:
        pi = rng.dirichlet(np.full(n, 5.0))
        s = rng.uniform(0.2, 3.0, (n, n))
        s = (s + s.T) / 2
        Q = s * pi[None, :]
    else:
        Q = rng.uniform(0.05, 2.0, (n, n))
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    pi = stationary(Q)
    return Q / -(pi * np.diag(Q)).sum(), pi


def gamma_like_rates(nrates, rng):
    r = np.sort(rng.uniform(0.1, 2.5, nrates))
    p = rng.dirichlet(np.full(nrates, 4.0))
    return r / (r * p).sum(), p


def nonrev_model(n, nmix=1, nrates=4, fused=False, seed=0, reversible=False):
    rng = np.random.default_rng(seed)
    Qs, pis, ev, V, W = [], [], [], [], []
    for _ in range(nmix):
        Q, pi = random_q(n, rng, reversible)
        w, vec = np.linalg.eig(Q)
        Qs.append(Q)
        pis.append(pi)
        ev.append(w)
        V.append(vec)
        W.append(np.linalg.inv(vec))
    rates, props = gamma_like_rates(nrates, rng)
    mixw = rng.dirichlet(np.full(nmix, 3.0)) if nmix > 1 else np.ones(1)
    return dict(Q=np.array(Qs), pi=np.array(pis), ceval=np.array(ev), cevec=np.array(V),
                cinv_evec=np.array(W), mixw=mixw, rates=rates, props=props, fused=fused)


def categories(nm):
    """(rate, weight, component) per category, as ncat_mix orders them."""
    nr, nmix = len(nm["rates"]), len(nm["mixw"])
    C = nr if nm["fused"] else nr * nmix
    denom = 1 if nm["fused"] else nr
    return [(nm["rates"][c % nr], nm["props"][c % nr] * nm["mixw"][c // denom], c // denom)
            for c in range(C)]


def adjacency(edges):
    u, v, ln = edges
    adj = {}
    for a, b, t in zip(u.tolist(), v.tolist(), ln.tolist()):
        adj.setdefault(a, []).append((b, t))
        adj.setdefault(b, []).append((a, t))
    return adj


def parents(adj, root):
    par, stack = {root: None}, [root]
    while stack:
        x = stack.pop()
        for y, _ in adj[x]:
            if y not in par:
                par[y] = x
                stack.append(y)
    return par


def plan(ntaxa, edges, u, v, root=None, root_leaf=None):
    """Ops for the branch (u, v) with u internal: u's partial away from v and,
    when v is internal, v's partial away from u. With a root node, a child on
    the path to the root gets flag 2 and the root leaf flag 4. Slots are
    internal node - ntaxa. Returns ops, node end, dad end, length."""
    adj = adjacency(edges)
    par = parents(adj, root) if root is not None else None
    ops = []

    def visit(x, frm):
        kids = [(y, t) for y, t in adj[x] if y != frm]
        for y, _ in kids:
            if y >= ntaxa:
                visit(y, x)
        assert len(kids) == 2
        rec = [x - ntaxa, 2]
        for y, t in kids:
            flag = 1 if y < ntaxa else 0
            if par is not None and par.get(x) == y:
                flag |= 2
            if y == root_leaf:
                flag |= 4
            rec += [flag, y if y < ntaxa else y - ntaxa, t]
        ops.append(tuple(rec))

    visit(u, v)
    if v >= ntaxa:
        visit(v, u)
        dad = End(0, v - ntaxa)
    else:
        dad = End(1 | (4 if v == root_leaf else 0), v)
    length = [t for y, t in adj[u] if y == v][0]
    return np.array(ops, dtype=OP_DTYPE), End(0, u - ntaxa), dad, length


def brute_force_rooted(ntaxa, edges, root, states, nm, ind, skip_leaf=None):
    """Pattern likelihoods of the tree rooted at `root` with the component's
    frequencies as the root prior, every edge directed away from the root,
    P = expm(Q r t), summed over every internal-state assignment; skip_leaf
    (the root leaf) carries no data."""
    adj = adjacency(edges)
    par = parents(adj, root)
    nodes = sorted(adj)
    internal = [x for x in nodes if x >= ntaxa]
    n = nm["Q"].shape[1]
    P = states.shape[1]
    total = np.zeros(P)
    for r, w, m in categories(nm):
        Q, pi = nm["Q"][m], nm["pi"][m]
        length = {x: [t for y, t in adj[x] if y == par[x]][0] for x in nodes if par[x] is not None}
        mats = {x: expm(Q * r * length[x]) for x in length}
        lh = np.zeros(P)
        for assign in itertools.product(range(n), repeat=len(internal)):
            st = dict(zip(internal, assign))
            wgt = pi[st[root]]
            for x in internal:
                if par[x] is not None:
                    wgt = wgt * mats[x][st[par[x]], st[x]]
            if wgt == 0.0:
                continue
            term = np.full(P, wgt)
            for leaf in range(ntaxa):
                if leaf == skip_leaf:
                    continue
                term = term * (mats[leaf][st[par[leaf]]] @ ind[states[leaf]].T)
            lh += term
        total += w * lh
    return total
