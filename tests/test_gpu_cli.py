"""The program end to end on the GPU: FASTA files in, `--framework HIP`, records out.

Configuration 1 (64 reads of 150 nt with an adapter piece, 16 adapters + their
reverse complements) through `pypaswas_amd.pypaswasall.Pypaswas` — the
reference's pypaswas.py flow (readers, engine dispatch, hit filtering,
formatter) — compared byte for byte with the oracle's records of the same
reader-ordered inputs, in every output format (TXT, SAM, trimmedFasta,
FASTA) and through both plugins (aligner, trimmer).
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
