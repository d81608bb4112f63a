"""Seeded synthetic workloads (the five benchmark configurations).

Every generator returns ``(queries, targets, settings)`` where queries and
targets are lists of ``(id, sequence)`` in file order (before the reader's
reverse-complement append and length sort) and settings is a dict of
command-line options.  ``numpy.random.default_rng(seed)`` drives everything.
"""
from __future__ import annotations

import math
from typing import Dict, List, Tuple

import numpy as np

Records = List[Tuple[str, str]]

DNA = np.array(list("ACGT"))
AMINO = np.array(list("ARNDCQEGHILKMFPSTWYV"))
# Robinson & Robinson background frequencies (the BLOSUM62 background), in AMINO order
AMINO_FREQ = np.array([0.07805, 0.05129, 0.04487, 0.05364, 0.01925, 0.04264, 0.06295, 0.07377, 0.02199, 0.05142,
                       0.09019, 0.05744, 0.02243, 0.03856, 0.05203, 0.07120, 0.05841, 0.01330, 0.03216, 0.06441])
AMINO_FREQ = AMINO_FREQ / AMINO_FREQ.sum()

# explicit filter settings: the defaults (relative_score 2.0, base_score 2.0,
# query_coverage/identity 0.2) reject every hit when match = +1
OPEN_FILTERS = {"relative_score": "0", "base_score": "0", "query_coverage": "0", "query_identity": "0"}


def _uniform_dna(rng, n):
    return "".join(rng.choice(DNA, size=n))


def _lognormal_lengths(rng, n, median, sigma, lo, hi):
    raw = np.exp(rng.normal(math.log(median), sigma, size=n))
    return np.clip(np.rint(raw), lo, hi).astype(int)


def _substitute(rng, s, rate, alphabet=DNA, p=None):
    arr = np.array(list(s))
    mask = rng.random(len(arr)) < rate
    arr[mask] = rng.choice(alphabet, size=int(mask.sum()), p=p)
    return "".join(arr)


def _mutate_indel(rng, s, identity, indel_rate, alphabet, p=None):
    out = []
    for c in s:
        r = rng.random()
        if r < indel_rate / 2:
            continue
        if r < indel_rate:
            out.append(str(rng.choice(alphabet, p=p)))
        out.append(c if rng.random() < identity else str(rng.choice(alphabet, p=p)))
    return "".join(out)


def _mutate_indel_np(rng, s, identity, indel_rate, alphabet, p=None):
    """_mutate_indel's per-position rule in array form: each residue is deleted
    with probability indel_rate / 2, gets a random residue inserted before it
    with probability indel_rate / 2, and (when kept) is replaced by a random one
    with probability 1 - identity."""
    a = np.frombuffer(s.encode(), dtype=np.uint8)
    n = len(a)
    alph = np.frombuffer("".join(alphabet).encode(), dtype=np.uint8)
    r = rng.random(n)
    keep = r >= indel_rate / 2
    ins = keep & (r < indel_rate)
    cur = np.where(rng.random(n) < identity, a, alph[rng.choice(len(alph), size=n, p=p)])
    pieces = np.stack([alph[rng.choice(len(alph), size=n, p=p)], cur], axis=1)
    return pieces[np.stack([ins, keep], axis=1)].tobytes().decode()


def config1(seed: int = 1):
    """64 reads of 150 nt (uniform ACGT), each with a 30-nt piece of one of 16
    adapters (30-60 nt) written over a uniform offset (the read stays 150 nt),
    then 5% substitutions; linear gap, +1/-3/-5."""
    rng = np.random.default_rng(seed)
    adapters = [("adapter%02d" % i, _uniform_dna(rng, int(rng.integers(30, 61)))) for i in range(16)]
    reads = []
    for i in range(64):
        read = list(_uniform_dna(rng, 150))
        a = adapters[int(rng.integers(0, 16))][1]
        start = int(rng.integers(0, len(a) - 30 + 1))
        piece = a[start:start + 30]
        off = int(rng.integers(0, 150 - 30 + 1))
        read[off:off + 30] = list(piece)
        reads.append(("read%03d" % i, _substitute(rng, "".join(read), 0.05)))
    settings = dict(matrix_name="DNA-RNA", match_score="1", mismatch_score="-3", gap_score="-5", gap_extension="0",
                    minimum_score="10", lower_limit_score="1.0", filter_factor="0.2", **OPEN_FILTERS)
    return reads, adapters, settings


def config2(seed: int = 2, n_reads: int = 20000, n_adapters: int = 64):
    """20,000 reads, lengths lognormal(median 250, sigma 0.5) clipped 50-1000,
    uniform ACGT with 1% N, vs 64 adapters of 20-80 nt; affine +1/-3, -5/-1."""
    rng = np.random.default_rng(seed)
    lengths = _lognormal_lengths(rng, n_reads, 250, 0.5, 50, 1000)
    reads = []
    for i, n in enumerate(lengths):
        s = np.array(list(_uniform_dna(rng, int(n))))
        s[rng.random(len(s)) < 0.01] = "N"
        reads.append(("read%05d" % i, "".join(s)))
    adapters = [("adapter%02d" % i, _uniform_dna(rng, int(rng.integers(20, 81)))) for i in range(n_adapters)]
    settings = dict(program="trimmer", matrix_name="DNA-RNA", match_score="1", mismatch_score="-3", gap_score="-5",
                    gap_extension="-1", minimum_score="10", lower_limit_score="1.0", filter_factor="0.2",
                    **OPEN_FILTERS)
    return reads, adapters, settings


def config3(seed: int = 3, n_queries: int = 2000, n_targets: int = 20000, families: int = 100):
    """Proteins, lengths lognormal(median 350, sigma 0.6) clipped 30-5000,
    residues i.i.d. from the background frequencies; BLOSUM62, gap -11/-1.
    Homologs: every sequence descends from one of 100 family ancestors at
    ~63% identity with 2% indels, so 1% of all pairs are same-family pairs at
    ~40% identity."""
    rng = np.random.default_rng(seed)
    anc_len = _lognormal_lengths(rng, families, 350, 0.6, 30, 5000)
    ancestors = ["".join(rng.choice(AMINO, size=int(n), p=AMINO_FREQ)) for n in anc_len]
    ident = math.sqrt(0.4)

    def member(prefix, i):
        f = int(rng.integers(0, families))
        s = _mutate_indel_np(rng, ancestors[f], ident, 0.02, AMINO, AMINO_FREQ)
        s = s[:5000] if len(s) >= 30 else s + "".join(rng.choice(AMINO, size=30 - len(s), p=AMINO_FREQ))
        return ("%s%05d_f%03d" % (prefix, i, f), s)

    queries = [member("q", i) for i in range(n_queries)]
    targets = [member("t", i) for i in range(n_targets)]
    settings = dict(matrix_name="BLOSUM62", gap_score="-11", gap_extension="-1", minimum_score="30",
                    lower_limit_score="1.0", filter_factor="0.05", **OPEN_FILTERS)
    return queries, targets, settings


def config4(seed: int = 4, n_pairs: int = 64, length: int = 10000):
    """64 DNA pairs of 10,000 x 10,000 (uniform ACGT) sharing a planted
    2,000-nt region at 85% identity with 2% indels; linear +1/-3/-5.  Pair i
    is query i vs target i (and its reverse complement): 64 one-query jobs
    (see config4_pair_mask).  The filter factor is 0.05 so that the planted
    pairs reach the traceback (at the default 0.2 the minimum start score,
    0.2 x 10,000, exceeds every pair maximum)."""
    rng = np.random.default_rng(seed)
    queries, targets = [], []
    for i in range(n_pairs):
        q = list(_uniform_dna(rng, length))
        t = list(_uniform_dna(rng, length))
        region = "".join(q[int(rng.integers(0, length - 2000)):][:2000])
        planted = _mutate_indel(rng, region, 0.85, 0.02, DNA)[:2000]
        off = int(rng.integers(0, length - len(planted)))
        t[off:off + len(planted)] = list(planted)
        queries.append(("long%02d_a" % i, "".join(q)))
        targets.append(("long%02d_b" % i, "".join(t)))
    settings = dict(matrix_name="DNA-RNA", match_score="1", mismatch_score="-3", gap_score="-5", gap_extension="0",
                    minimum_score="10", lower_limit_score="1.0", filter_factor="0.05", **OPEN_FILTERS)
    return queries, targets, settings


def config4_pair_mask(queries: Records, targets: Records, pq: np.ndarray, pt: np.ndarray) -> np.ndarray:
    """Pairs of the 64 config-4 jobs among all query x target pairs: query
    longNN_a with target longNN_b and its reverse complement longNN_b_RC."""
    qk = np.array([i[:6] for i, _ in queries])
    tk = np.array([i[:6] for i, _ in targets])
    return qk[pq] == tk[pt]


# Serving mix: scoring settings per request kind (settings_dict overrides)
SERVING_SETTINGS = {
    # short reads against an adapter panel: the trimmer scoring
    "dna": dict(program="trimmer", matrix_name="DNA-RNA", match_score="1", mismatch_score="-3", gap_score="-5",
                gap_extension="-1", minimum_score="10", lower_limit_score="1.0", filter_factor="0.2",
                **OPEN_FILTERS),
    "protein": dict(matrix_name="BLOSUM62", gap_score="-11", gap_extension="-1", minimum_score="30",
                    lower_limit_score="1.0", filter_factor="0.05", **OPEN_FILTERS),
    "long": dict(matrix_name="DNA-RNA", match_score="1", mismatch_score="-3", gap_score="-5", gap_extension="0",
                 minimum_score="10", lower_limit_score="1.0", filter_factor="0.2", **OPEN_FILTERS),
}
SERVING_KINDS = ("dna", "protein", "long")


def serving_jobs(seed: int = 5, n_jobs: int = 1000, panel: int = 8, long_range=(5000, 20000)):
    """Serving mix of requests, each a small pyPaSWAS job (records in reader order):
      * 70% "dna": one read of 100-300 nt (uniform ACGT) against a panel of
        `panel` adapters of 20-80 nt drawn for the request (plus their reverse
        complements, as the DNA reader adds them), trimmer scoring; half the
        reads carry a 20-nt adapter piece at 5% substitutions;
      * 25% "protein": one pair of 100-2,000 residues (background
        frequencies), BLOSUM62 -11/-1; half the pairs are homologs at 60%
        identity with 2% indels;
      * 5% "long": one DNA pair of 5,000-20,000 nt (long_range) at 90%
        identity with 1% indels (plus the reverse complement target), linear
        +1/-3/-5.
    Returns a list of (kind, queries, targets)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        r = rng.random()
        if r < 0.70:
            adapters = [("a%d_%d" % (j, k), _uniform_dna(rng, int(rng.integers(20, 81)))) for k in range(panel)]
            read = list(_uniform_dna(rng, int(rng.integers(100, 301))))
            if rng.random() < 0.5:
                a = adapters[int(rng.integers(0, panel))][1]
                piece = _substitute(rng, a[:20], 0.05)
                off = int(rng.integers(0, len(read) - len(piece) + 1))
                read[off:off + len(piece)] = list(piece)
            q, t = reference_order([("r%d" % j, "".join(read))], adapters, True)
            jobs.append(("dna", q, t))
        elif r < 0.95:
            n = int(rng.integers(100, 2001))
            qs = "".join(rng.choice(AMINO, size=n, p=AMINO_FREQ))
            ts = _mutate_indel(rng, qs, 0.6, 0.02, AMINO, AMINO_FREQ) if rng.random() < 0.5 else \
                "".join(rng.choice(AMINO, size=int(rng.integers(100, 2001)), p=AMINO_FREQ))
            jobs.append(("protein", [("p%d" % j, qs)], [("h%d" % j, ts)]))
        else:
            n = int(rng.integers(long_range[0], long_range[1] + 1))
            qs = _uniform_dna(rng, n)
            ts = _mutate_indel(rng, qs, 0.9, 0.01, DNA)
            q, t = reference_order([("l%d" % j, qs)], [("m%d" % j, ts)], True)
            jobs.append(("long", q, t))
    return jobs


def serving_schemes():
    """(Scheme per kind, in SERVING_KINDS order) for pypaswas_amd.serving."""
    import logging
    from pypaswas_amd.Core.Scores import Blosum62Score, DnaRnaScore
    from pypaswas_amd.serving import Scheme
    out = []
    for kind in SERVING_KINDS:
        st = settings_dict(SERVING_SETTINGS[kind])
        cls = Blosum62Score if st.matrix_name.upper() == "BLOSUM62" else DnaRnaScore
        out.append(Scheme.from_score(cls(logging.getLogger("serving"), st), float(st.filter_factor)))
    return out


def reverse_complement(s: str) -> str:
    from pypaswas_amd.Core.SWSeq import Seq
    return str(Seq(s).reverse_complement())


def reference_order(queries: Records, targets: Records, dna: bool) -> Tuple[Records, Records]:
    """Reader order: queries sorted by length (longest first, stable); DNA
    targets get their _RC copies appended, then the same sort."""
    q = sorted(queries, key=lambda r: len(r[1]), reverse=True)
    t = list(targets)
    if dna:
        t = t + [(i + "_RC", reverse_complement(s)) for i, s in t]
    t = sorted(t, key=lambda r: len(r[1]), reverse=True)
    return q, t


def write_fasta(path: str, records: Records, width: int = 60):
    with open(path, "w") as f:
        for i, s in records:
            f.write(">" + i + "\n")
            for k in range(0, len(s), width):
                f.write(s[k:k + width] + "\n")


def settings_dict(overrides: Dict[str, str]):
    """Full settings namespace: the defaults file plus overrides."""
    from pypaswas_amd.cli import parse_cli
    argv = []
    for k, v in overrides.items():
        argv.append("--" + {"matrix_name": "matrixname", "match_score": "match_score",
                            "mismatch_score": "mismatch_score", "gap_score": "gap_score",
                            "gap_extension": "gap_extension", "minimum_score": "minimum",
                            "lower_limit_score": "llimit", "filter_factor": "filter_factor",
                            "relative_score": "relative_score", "base_score": "base_score",
                            "query_coverage": "query_coverage", "query_identity": "query_identity",
                            "program": "program", "framework": "framework", "loglevel": "loglevel",
                            "out_format": "outputformat", "any_score": "any", "other_score": "other"}[k] + "=" + v)
    settings, _ = parse_cli(argv=argv + ["-", "-"])
    return settings
