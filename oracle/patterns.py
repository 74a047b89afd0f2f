"""Pure-Python restatement of IQ-TREE's site-pattern compression.

TEST INFRASTRUCTURE ONLY. Follows alignment/alignment.cpp: convertState
(1415-1499), codon triplets (1849-1879), first-occurrence deduplication
(addPatternLazy 1092-1124), computeConst (925-1022), getAppearance and
PhyloTree::computePtnInvar (tree/phylotreesse.cpp:549-617).
"""
DNA, PROTEIN, CODON = 0, 1, 2
NUM_STATES = {DNA: 4, PROTEIN: 20, CODON: 61}
STATE_UNKNOWN = {DNA: 18, PROTEIN: 23, CODON: 61}
_PROTEIN = "ARNDCQEGHILKMFPSTWYVX"
_DNA = {"A": 0, "C": 1, "G": 2, "T": 3, "U": 3,
        "R": 1 + 4 + 3, "Y": 2 + 8 + 3, "W": 1 + 8 + 3, "S": 2 + 4 + 3, "M": 1 + 2 + 3,
        "K": 4 + 8 + 3, "B": 2 + 4 + 8 + 3, "H": 1 + 2 + 8 + 3, "D": 1 + 4 + 8 + 3,
        "V": 1 + 2 + 4 + 3}
# standard genetic code in TCAG order (NCBI table 1)
_CODE_TCAG = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG"


def _sense_codons():
    tcag = "TCAG"
    aa = {}
    for i, a in enumerate(tcag):
        for j, b in enumerate(tcag):
            for k, c in enumerate(tcag):
                aa[a + b + c] = _CODE_TCAG[16 * i + 4 * j + k]
    sense = [a + b + c for a in "ACGT" for b in "ACGT" for c in "ACGT" if aa[a + b + c] != "*"]
    return {cod: idx for idx, cod in enumerate(sense)}, aa


SENSE_INDEX, AMINO = _sense_codons()


def convert_state(seq_type, ch):
    ch = ch.upper()
    su = STATE_UNKNOWN[seq_type]
    if ch in "?-.~":
        return su
    if seq_type == DNA:
        if ch in "ONX":
            return su
        return _DNA.get(ch, -1)
    if ch == "B":
        return 20
    if ch == "Z":
        return 21
    if ch == "J":
        return 22
    if ch in "*UO":
        return su
    k = _PROTEIN.find(ch)
    if k < 0:
        return -1
    return k if k < 20 else su


def convert_codon(triplet):
    t = triplet.upper().replace("U", "T")
    if all(c in "ACGT" for c in t):
        return SENSE_INDEX.get(t, -2)
    for c in t:
        if c not in "ACGTRYWSMKBHDVONX?-.~":
            return -1
    return STATE_UNKNOWN[CODON]


def appearance(seq_type, state):
    n, su = NUM_STATES[seq_type], STATE_UNKNOWN[seq_type]
    if state == su:
        return set(range(n))
    if state < n:
        return {state}
    if seq_type == DNA:
        mask = state - (n - 1)
        return {i for i in range(n) if mask & (1 << i)}
    if seq_type == PROTEIN and state < 23:
        return set([(2, 3), (5, 6), (9, 10)][state - 20])
    return set()


def const_char(seq_type, column):
    n, su = NUM_STATES[seq_type], STATE_UNKNOWN[seq_type]
    common = set(range(n))
    for s in column:
        common &= appearance(seq_type, s)
    if not common:
        return su + 1
    if len(common) == n:
        return su
    if len(common) == 1:
        return next(iter(common))
    if seq_type == DNA:
        return n - 1 + sum(1 << j for j in common)
    if seq_type == PROTEIN:
        for k, pair in enumerate([(2, 3), (5, 6), (9, 10)]):
            if set(pair) <= common:
                return n + k
    return su + 1


def build_patterns(seq_type, rows):
    """rows: list of equal-length strings. Returns (columns, freq, site_pattern, const_chars)
    with columns as tuples of states in first-occurrence order."""
    step = 3 if seq_type == CODON else 1
    nsite = len(rows[0]) // step
    index, columns, freq, site_pattern = {}, [], [], []
    for site in range(nsite):
        if seq_type == CODON:
            col = tuple(convert_codon(r[3 * site:3 * site + 3]) for r in rows)
        else:
            col = tuple(convert_state(seq_type, r[site]) for r in rows)
        if any(s < 0 for s in col):
            raise ValueError(f"invalid character at site {site}")
        if col in index:
            freq[index[col]] += 1
        else:
            index[col] = len(columns)
            columns.append(col)
            freq.append(1)
        site_pattern.append(index[col])
    consts = [const_char(seq_type, c) for c in columns]
    return columns, freq, site_pattern, consts


def ptn_invar(seq_type, consts, pi, p_inv):
    n, su = NUM_STATES[seq_type], STATE_UNKNOWN[seq_type]
    out = []
    for cc in consts:
        if p_inv == 0.0 or cc > su:
            out.append(0.0)
        elif cc == su:
            out.append(p_inv)
        elif cc < n:
            out.append(p_inv * pi[cc])
        elif seq_type == DNA:
            mask = cc - n + 1
            out.append(sum(pi[x] for x in range(n) if mask & (1 << x)) * p_inv)
        else:
            pair = [(2, 3), (5, 6), (9, 10)][cc - n]
            out.append((pi[pair[0]] + pi[pair[1]]) * p_inv)
    return out
