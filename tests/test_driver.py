"""The application driver on the CPU: option tables, output path and the
window rules of pyPaSWAS/pypaswasall.py:173-245, with a recording program in
place of the aligner (no device is opened).

Expected windows are worked out by hand from the reference's loop:
query windows [start, start+step) until the reader finds a window empty or a
window's end passes end_query; for each, target windows from start_target,
the loop stopping after a window that holds no more records (reverse
complements included) than its end index.
"""
import pytest

from pypaswas_amd.Core.Exceptions import InvalidOptionException


class Recorder(object):
    calls = []

    def __init__(self, logger, score, settings):
        self.logger = logger

    def process(self, queries, targets, pypaswas):
        from pypaswas_amd.Core.HitList import HitList
        Recorder.calls.append((sorted(r.id for r in queries), sorted(r.id for r in targets)))
        return HitList(self.logger)


@pytest.fixture
def driver(tmp_path, monkeypatch):
    from pypaswas_amd import pypaswasall
    from pypaswas_amd import workloads as W
    monkeypatch.setitem(pypaswasall.PROGRAMS, "aligner", Recorder)
    Recorder.calls = []

    def run(n_queries, n_targets, extra, protein=False):
        res = "ACDEFGHIKL" if protein else "ACGT"
        W.write_fasta(str(tmp_path / "q.fa"), [("q%d" % i, res * 3) for i in range(n_queries)])
        W.write_fasta(str(tmp_path / "t.fa"), [("t%d" % i, res * 2) for i in range(n_targets)])
        matrix = ["-M", "BLOSUM62"] if protein else ["-M", "DNA-RNA"]
        argv = matrix + ["-o", str(tmp_path / "out.txt"), "-O", "T"] + extra + [str(tmp_path / "q.fa"),
                                                                               str(tmp_path / "t.fa")]
        pypaswasall.Pypaswas(argv=argv).run()
        return Recorder.calls
    return run


def ids(prefix, idx, rc=False):
    out = ["%s%d" % (prefix, i) for i in idx]
    return sorted(out + (["%s%d_RC" % (prefix, i) for i in idx] if rc else []))


def test_query_windows_page_through_the_file(driver):
    calls = driver(5, 3, ["--query_step", "2"])
    # [0,2) [2,4) [4,6); [6,8) is empty.  Targets: [0,1000) holds 6 records <= 1000, the only window
    t = ids("t", range(3), rc=True)
    assert calls == [(ids("q", [0, 1]), t), (ids("q", [2, 3]), t), (ids("q", [4]), t)]


def test_dna_target_windows_continue_while_complements_outnumber_the_end(driver):
    calls = driver(1, 5, ["--sequence_step", "2"])
    # [0,2) holds t0, t1 + complements = 4 > 2: go on; [2,4) holds 4 <= 4: last.  t4 is never read.
    assert calls == [(["q0"], ids("t", [0, 1], rc=True)), (["q0"], ids("t", [2, 3], rc=True))]


def test_protein_target_windows_stop_after_the_first(driver):
    calls = driver(1, 5, ["--sequence_step", "2"], protein=True)
    assert calls == [(["q0"], ids("t", [0, 1]))]


def test_end_query_ends_the_loop_after_the_window_that_passes_it(driver):
    calls = driver(6, 1, ["--query_step", "2", "--end_query", "3"])
    # [0,2), then [2,4) passes 3 and is the last
    t = ids("t", [0], rc=True)
    assert calls == [(ids("q", [0, 1]), t), (ids("q", [2, 3]), t)]


def test_start_equal_to_end_is_not_a_bound_on_the_first_window(driver):
    calls = driver(6, 1, ["--query_step", "2", "--start_query", "1", "--end_query", "1"])
    assert calls == [(ids("q", [1, 2]), ids("t", [0], rc=True))]


def test_start_target_and_end_target(driver):
    calls = driver(1, 6, ["--start_target", "1", "--sequence_step", "2", "--end_target", "4"])
    # [1,3): 4 records > 3, go on; next [3,5) ends past 4: stop before reading it
    assert calls == [(["q0"], ids("t", [1, 2], rc=True))]


def test_first_window():
    from pypaswas_amd.pypaswasall import first_window
    assert first_window(0, 1000, 0) == (0, 1000)
    assert first_window(0, 10, 4) == (0, 4)
    assert first_window(3, 10, 3) == (3, 13)
    assert first_window(2, 2, 9) == (2, 4)


@pytest.mark.parametrize("args,message", [
    (["-M", "NOPE"], "NOPE is not a valid substitution matrix"),
    (["--outputformat", "XML"], "Invalid output format XML"),
    (["--outputformat", "GRAPH"], "GRAPH output"),
    (["-p", "palindrome"], "Invalid program selected palindrome"),
])
def test_unknown_names_are_rejected(tmp_path, args, message):
    from pypaswas_amd import workloads as W
    from pypaswas_amd.pypaswasall import Pypaswas
    W.write_fasta(str(tmp_path / "q.fa"), [("q0", "ACGT")])
    argv = ["-o", str(tmp_path / "out.txt"), "-O", "T"] + args + [str(tmp_path / "q.fa"), str(tmp_path / "q.fa")]
    with pytest.raises(InvalidOptionException, match=message):
        Pypaswas(argv=argv).run()


def test_output_path_rules(tmp_path):
    from types import SimpleNamespace
    from pypaswas_amd.pypaswasall import output_path
    existing = tmp_path / "there.txt"
    existing.write_text("x")
    with pytest.raises(IOError, match="already exists"):
        output_path(SimpleNamespace(out_file=str(existing), override_output="F"))
    assert output_path(SimpleNamespace(out_file=str(existing), override_output="T")) == str(existing)
    nested = tmp_path / "a" / "b" / "out.sam"
    assert output_path(SimpleNamespace(out_file=str(nested), override_output="F")) == str(nested)
    assert nested.parent.is_dir()
