"""Independent likelihood mathematics used to pin the oracle.

Nothing here shares code or basis with the oracle: transition matrices come
from scipy.linalg.expm(Q r t) in the real basis (no eigen-decomposition),
partials are pruned with per-pattern log scaling in numpy, and tiny trees are
summed over every assignment of internal states.
"""
import itertools

import numpy as np
from scipy.linalg import expm


def adjacency(u, v, t):
    adj = {}
    for a, b, ln in zip(u.tolist(), v.tolist(), t.tolist()):
        adj.setdefault(a, []).append((b, ln))
        adj.setdefault(b, []).append((a, ln))
    return adj


def invar_by_definition(states, indicators, pi, p_inv):
    """+I mass of every pattern from its definition: the probability that the
    column arose from an invariable site, p_inv * sum_x pi_x * prod_taxa
    [state x is in the taxon's observed set] (a gap or unknown character
    admits every state, an ambiguity code its set)."""
    compat = np.prod(indicators[states], axis=0)      # (P, n): x compatible with every taxon
    return p_inv * (compat @ np.asarray(pi))


class Pruner:
    """Real-basis Felsenstein pruning with expm transition matrices. The +I
    mass is formed from p_inv and the tip states (invar_by_definition)."""

    def __init__(self, ntaxa, edges, states, Q, pi, rates, props, p_inv, freq, indicators):
        self.N = ntaxa
        self.adj = adjacency(*edges)
        self.states = states
        self.Q = Q
        self.pi = pi
        self.rates = np.asarray(rates)
        self.props = np.asarray(props)
        self.invar = invar_by_definition(states, indicators, pi, p_inv)
        self.freq = freq
        self.ind = indicators

    def P(self, t):
        return np.stack([expm(self.Q * r * t) for r in self.rates])   # (C, n, n)

    def side(self, node, dad):
        """Partial of `node` seen from `dad`: (C, P, n) array and log scale (P,)."""
        C = len(self.rates)
        if node < self.N:
            v = self.ind[self.states[node]]                     # (P, n)
            return np.broadcast_to(v, (C,) + v.shape).copy(), np.zeros(v.shape[0])
        acc, ls = None, None
        for ch, t in self.adj[node]:
            if ch == dad:
                continue
            L, s = self.side(ch, node)
            Pt = self.P(t)
            M = np.einsum("cxy,cpy->cpx", Pt, L)
            acc = M if acc is None else acc * M
            ls = s if ls is None else ls + s
        m = acc.max(axis=(0, 2))
        m = np.where(m > 0, m, 1.0)
        return acc / m[None, :, None], ls + np.log(m)

    def branch_terms(self, a, b, t):
        """Per-pattern (log lh, d lh/dt / lh, d2 lh/dt2 / lh) on branch (a, b) at length t."""
        A, sa = self.side(a, b)
        B, sb = self.side(b, a)
        Pt = self.P(t)
        rQ = self.rates[:, None, None] * self.Q[None]
        P1 = np.einsum("cxz,czy->cxy", rQ, Pt)
        P2 = np.einsum("cxz,czy->cxy", rQ, P1)

        def site(Pm):
            return np.einsum("c,x,cpx,cxy,cpy->p", self.props, self.pi, A, Pm, B)

        s0, s1, s2 = site(Pt), site(P1), site(P2)
        ls = sa + sb
        # lh = exp(ls) s0 + invar, kept in log space
        with np.errstate(divide="ignore"):
            log_inv = np.log(self.invar)
        log_lh = np.logaddexp(ls + np.log(s0), log_inv)
        scale = np.exp(ls - log_lh)          # exp(ls) / lh
        return log_lh, s1 * scale, s2 * scale

    def lnl(self, a, b, t):
        log_lh, _, _ = self.branch_terms(a, b, t)
        return float(np.dot(self.freq, log_lh)), log_lh

    def derivs(self, a, b, t):
        log_lh, d1, d2 = self.branch_terms(a, b, t)
        df = float(np.dot(self.freq, d1))
        ddf = float(np.dot(self.freq, d2 - d1 * d1))
        return float(np.dot(self.freq, log_lh)), df, ddf


def brute_force_lnl(ntaxa, edges, states, Q, pi, rates, props, p_inv, freq, indicators):
    """Sum over every assignment of internal-node states (tiny trees only),
    plus the +I term formed from p_inv and the tip states."""
    invar = invar_by_definition(states, indicators, pi, p_inv)
    adj = adjacency(*edges)
    nodes = sorted(adj)
    internal = [x for x in nodes if x >= ntaxa]
    n = Q.shape[0]
    root = internal[0]
    # parent pointers from the root
    parent, order, stack = {root: None}, [], [root]
    while stack:
        x = stack.pop()
        order.append(x)
        for y, t in adj[x]:
            if y not in parent:
                parent[y] = (x, t)
                stack.append(y)
    P = states.shape[1]
    total = np.zeros(P)
    for c, (r, pc) in enumerate(zip(rates, props)):
        lh_c = np.zeros(P)
        mats = {x: expm(Q * r * parent[x][1]) for x in order if parent[x] is not None}
        for assign in itertools.product(range(n), repeat=len(internal)):
            st = dict(zip(internal, assign))
            w = pi[st[root]]
            for x in internal:
                if parent[x] is not None:
                    w = w * mats[x][st[parent[x][0]], st[x]]
            if w == 0.0:
                continue
            site = np.full(P, w)
            for leaf in range(ntaxa):
                px = st[parent[leaf][0]]
                site = site * (indicators[states[leaf]] @ mats[leaf][px])
            lh_c += site
        total += pc * lh_c
    lh = total + invar
    with np.errstate(divide="ignore"):
        log_lh = np.log(lh)
    return float(np.dot(freq, log_lh)), log_lh


def indicators(seq_type, n, state_unknown):
    """Observed-state indicator vectors [SU+1][n] from the alphabet definitions:
    DNA ambiguity = bitmask + 3 over A,C,G,T; protein B/Z/J = N|D, Q|E, I|L;
    STATE_UNKNOWN = every state."""
    ind = np.zeros((state_unknown + 1, n))
    for s in range(state_unknown + 1):
        if s < n:
            ind[s, s] = 1.0
        elif s == state_unknown:
            ind[s, :] = 1.0
        elif seq_type == 0:
            mask = s - 3
            for x in range(4):
                if mask & (1 << x):
                    ind[s, x] = 1.0
        elif seq_type == 1:
            pair = {20: (2, 3), 21: (5, 6), 22: (9, 10)}[s]
            ind[s, list(pair)] = 1.0
    return ind


def brute_force_parsimony(ntaxa, edges, states, indicators, branch=None):
    """Per pattern, the least number of state changes over every assignment of
    internal-node states (tiny trees only): an internal edge costs 1 when its
    ends differ, a tip edge when the internal end's state is not among the
    states the tip's character admits (its indicator row).

    With branch=(a, b), also whether branch (a, b) itself carries a change
    (0 or 1): 1 where forcing (a, b) to cost nothing raises the least cost of
    the other edges, i.e. every optimal labelling changes on (a, b), which
    happens exactly when the two sides' optimal state sets are disjoint
    (Fitch's test at a root placed on that branch)."""
    adj = adjacency(*edges)
    internal = sorted(x for x in adj if x >= ntaxa)
    n = indicators.shape[1]
    P = states.shape[1]
    admits = indicators[states] > 0                  # (ntaxa, P, n)
    big = np.iinfo(np.int64).max // 4
    best = np.full(P, big)
    free, forced = np.full(P, big), np.full(P, big)
    iedges = [(x, y) for x in internal for y, _ in adj[x] if y in internal and x < y]
    tedges = [(x, y) for x in internal for y, _ in adj[x] if y < ntaxa]
    key = None if branch is None else frozenset(branch)
    for assign in itertools.product(range(n), repeat=len(internal)):
        st = dict(zip(internal, assign))
        cost = np.zeros(P, np.int64)
        rest = np.zeros(P, np.int64)            # every edge but the branch
        on = np.zeros(P, bool)                  # the branch's cost
        for x, y in iedges:
            c = int(st[x] != st[y])
            cost += c
            if frozenset((x, y)) == key:
                on[:] = bool(c)
            else:
                rest += c
        for x, leaf in tedges:
            c = ~admits[leaf, :, st[x]]
            cost += c
            if frozenset((x, leaf)) == key:
                on = c
            else:
                rest += c
        best = np.minimum(best, cost)
        if key is not None:
            free = np.minimum(free, rest)
            forced = np.minimum(forced, np.where(on, big, rest))
    if key is None:
        return best
    return best, (forced > free).astype(np.int64)
