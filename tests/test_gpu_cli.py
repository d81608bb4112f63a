"""The program end to end on the GPU: FASTA files in, `--framework HIP`, records out.

Configuration 1 (64 reads of 150 nt with an adapter piece, 16 adapters + their
reverse complements) through `pypaswas_amd.pypaswasall.Pypaswas` — the
reference's pypaswas.py flow (readers, engine dispatch, hit filtering,
formatter) — compared byte for byte with the oracle's records of the same
reader-ordered inputs, in every output format (TXT, SAM, trimmedFasta,
FASTA), through both plugins (aligner, trimmer) and under both kernel
semantics (the OpenCL-CPU recurrence and the GPU-OpenCL / CUDA one).
"""
import pytest

pytestmark = pytest.mark.gpu


def run_program(tmp_path, reads, adapters, extra):
    from pypaswas_amd import workloads as W
    from pypaswas_amd.pypaswasall import Pypaswas
    qf, tf, out = tmp_path / "reads.fa", tmp_path / "adapters.fa", tmp_path / "out.txt"
    W.write_fasta(str(qf), reads)
    W.write_fasta(str(tf), adapters)
    argv = ["-M", "DNA-RNA", "-r", "1", "-q", "-3", "-G", "-5", "--minimum", "10", "--relative_score", "0",
            "--base_score", "0", "--query_coverage", "0", "--query_identity", "0", "-o", str(out), "-O", "T",
            "--framework", "HIP"] + extra + [str(qf), str(tf)]
    Pypaswas(argv=argv).run()
    return out.read_text()


@pytest.mark.parametrize("affine", [False, True])
def test_aligner_config1_matches_oracle(tmp_path, oracle, affine):
    from pypaswas_amd import workloads as W
    print("program config1 affine", affine, flush=True)
    reads, adapters, _ = W.config1()
    got = run_program(tmp_path, reads, adapters, ["-g", "-1" if affine else "0"])
    q, t = oracle.prepare_records(reads, adapters, True)
    p = oracle.Params(matrix=oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0,
                      gap_ext=(-1.0 if affine else None), lower_limit=1.0, minimum=10.0, filter_factor=0.2,
                      highest=1.0, relative_score=0.0, base_score=0.0, query_coverage=0.0, query_identity=0.0,
                      dna=True)
    hl, _, _ = oracle.align_all(q, t, p)
    want = oracle.txt_records(hl)
    assert got == want
    assert len(hl.real_hits) >= 32  # most reads carry an adapter piece


def config1_oracle(oracle, affine):
    from pypaswas_amd import workloads as W
    reads, adapters, _ = W.config1()
    q, t = oracle.prepare_records(reads, adapters, True)
    p = oracle.Params(matrix=oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=-5.0,
                      gap_ext=(-1.0 if affine else None), lower_limit=1.0, minimum=10.0, filter_factor=0.2,
                      highest=1.0, relative_score=0.0, base_score=0.0, query_coverage=0.0, query_identity=0.0,
                      dna=True)
    hl, _, _ = oracle.align_all(q, t, p)
    return reads, adapters, hl


@pytest.mark.parametrize("affine", [False, True])
@pytest.mark.parametrize("fmt", ["SAM", "trimmedFasta", "FASTA"])
def test_aligner_config1_records_match_oracle(tmp_path, oracle, fmt, affine):
    """SAM, trimmed FASTA and FASTA records of configuration 1 through
    `--framework HIP`, byte for byte against the oracle's restatements of
    SamFormatter, TrimmerFormatter and FASTA (Formatters.py:51-191,
    Hit.py:222-307) on the oracle's hits."""
    print("program config1", fmt, "affine", affine, flush=True)
    reads, adapters, hl = config1_oracle(oracle, affine)
    got = run_program(tmp_path, reads, adapters, ["-g", "-1" if affine else "0", "--outputformat", fmt])
    want = {"SAM": oracle.sam_records, "trimmedFasta": oracle.trimmed_records, "FASTA": oracle.fasta_records}[fmt](hl)
    assert got == want
    assert len(hl.real_hits) >= 32
    if fmt == "SAM":
        # reverse-complement hits are part of the comparison (FLAG 16, reversed strings)
        assert any(l.split("\t")[1] == "16" for l in got.splitlines() if not l.startswith("@"))


def test_trimmer_program_trimmed_fasta(tmp_path, oracle):
    """The trimmer plugin on the same engine, trimmedFasta output, byte for byte."""
    reads, adapters, hl = config1_oracle(oracle, False)
    got = run_program(tmp_path, reads, adapters, ["-g", "0", "-p", "trimmer", "--outputformat", "trimmedFasta"])
    assert got == oracle.trimmed_records(hl)
    assert len(hl.real_hits) >= 32


def indel_set(seed=9):
    """Reads carrying a 40-nt adapter with one or two residues inserted (gapped
    best alignments), plus uniform reads."""
    import numpy as np
    rng = np.random.default_rng(seed)
    dna = np.array(list("ACGT"))
    adapters = [("adapter%d" % i, "".join(rng.choice(dna, size=40))) for i in range(4)]
    reads = []
    for i in range(24):
        a = adapters[i % 4][1]
        cut = int(rng.integers(12, 28))
        ins = "".join(rng.choice(dna, size=1 + i % 2))
        piece = a[:cut] + ins + a[cut:]
        left, right = "".join(rng.choice(dna, size=30)), "".join(rng.choice(dna, size=30))
        reads.append(("read%02d" % i, left + piece + right))
    return reads, adapters


@pytest.mark.parametrize("semantics", ["CPU", "GPU"])
def test_aligner_kernel_semantics_match_oracle(tmp_path, oracle, semantics):
    """`--kernel_semantics` CPU (the OpenCL-CPU recurrence, gap open G) and GPU
    (the GPU-OpenCL / CUDA recurrence, gap open G + E) on reads with gapped
    adapter hits equal the oracle run with gap open -5 or -5 + -1; the two
    settings give different records on this data."""
    print("program indel set, kernel semantics", semantics, flush=True)
    reads, adapters = indel_set()
    got = run_program(tmp_path, reads, adapters, ["-g", "-1", "--kernel_semantics", semantics])
    q, t = oracle.prepare_records(reads, adapters, True)
    want = {}
    for name, g in (("CPU", -5.0), ("GPU", -6.0)):
        p = oracle.Params(matrix=oracle.dna_rna_matrix(1.0, -3.0, -1.0, 0.0), gap=g, gap_ext=-1.0, lower_limit=1.0,
                          minimum=10.0, filter_factor=0.2, highest=1.0, relative_score=0.0, base_score=0.0,
                          query_coverage=0.0, query_identity=0.0, dna=True)
        hl, _, _ = oracle.align_all(q, t, p)
        want[name] = oracle.txt_records(hl)
    assert got == want[semantics]
    assert want["CPU"] != want["GPU"]
    assert got.count("\n") >= 24
