"""The autotuner's table writer (host logic, no GPU): retuning some buckets keeps
every other bucket and the ragged lines of the existing table."""
from tridiagonalsolver_amd import autotune


def _entry(M):
    return dict(M=M, NW=4, S=1, vec=1, lds_alpha=0, nontemporal=1, transposed=1)


def test_partial_retune_keeps_other_buckets(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("# header\n"
                 "R 0 32 0 1\n"
                 "0 0 0 512 8 4 4 1 0 1 1\n"
                 "0 0 0 1023 16 4 4 1 0 1 1\n"
                 "0 0 0 1024 16 4 4 1 0 1 1\n"
                 "1 0 0 1024 8 4 2 1 0 1 1\n"
                 "0 2 0 4096 64 8 8 1 0 0 1\n")
    autotune.write_table(str(p), [(0, 0, 1023, _entry(8))])
    t = autotune.read_table(str(p))
    assert t[(0, 0, 1023, 0)] == (8, 4, 1, 1, 0, 1, 1)          # retuned
    assert t[(0, 0, 512, 0)] == (8, 4, 4, 1, 0, 1, 1)           # kept
    assert t[(0, 0, 1024, 0)] == (16, 4, 4, 1, 0, 1, 1)
    assert t[(1, 0, 1024, 0)] == (8, 4, 2, 1, 0, 1, 1)
    assert t[(0, 2, 4096, 0)] == (64, 8, 8, 1, 0, 0, 1)
    assert "R 0 32 0 1" in p.read_text().splitlines()
    # a row of another interface-solver family (PCR, 2) is a bucket of its own
    autotune.write_table(str(p), [(0, 0, 1024, dict(_entry(4), family=2))])
    t = autotune.read_table(str(p))
    assert t[(0, 0, 1024, 2)] == (4, 4, 1, 1, 0, 1, 1)
    assert t[(0, 0, 1024, 0)] == (16, 4, 4, 1, 0, 1, 1)
    assert t[(0, 0, 1023, 0)] == (8, 4, 1, 1, 0, 1, 1)
    # buckets stay sorted by family, then n_max, within (precision, layout)
    lines = [ln.split() for ln in p.read_text().splitlines() if ln[:1].isdigit()]
    keys = [(int(f[0]), int(f[1]), int(f[2]), int(f[3])) for f in lines]
    assert keys == sorted(keys)
