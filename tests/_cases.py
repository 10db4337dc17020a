"""Shared helpers of the tests: the library's environment switches set for
a block, and the inputs of the writer-format checks
(tests/golden/writers_expected.json)."""
import contextlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


@contextlib.contextmanager
def env(**kw):
    """Environment variables set for the block (None: unset), as the library
    reads them when a context or batch is created; the previous values are
    restored afterwards, also when the block raises."""
    old = {k: os.environ.get(k) for k in kw}
    try:
        for k, v in kw.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def writers_fixture():
    with open(os.path.join(HERE, "golden", "writers_expected.json")) as f:
        return json.load(f)


def writers_case(abi):
    """(params, w, p, fixture) of the hand-solved writer case."""
    g = writers_fixture()
    q = g["params"]
    p = abi.make_params(lx=q["lx"], ly=q["ly"], lz=q["lz"], dr_target=q["dr_target"], ns=q["ns"], mask=q["mask"],
                        g1=q["g1"], chi=q["chi"], vol_homo=tuple(q["vol_homo"]), len_homo=tuple(q["len_homo"]))
    nx, ny, nz = abi.grid_of(p)
    V = nx * ny * nz
    w = np.zeros((13, V))
    w[:5] = g["w"]
    return p, w, np.full(V, g["p"]), g


def check_writer_files(d, g):
    """Compare a run directory's output with the hand-written lines."""
    def norm(line, padded=True):
        # p - mean(p) is zero up to rounding: a negative zero keeps the field's width
        return line.replace("-0.000000", " 0.000000" if padded else "0.000000")
    for fn, lines in g["files"].items():
        got = open(os.path.join(d, fn)).read()
        want = "\n".join(lines) + ("\n" if fn != "counter.dat" else "")
        assert norm(got) == want, (fn, got, want)
    snap = open(os.path.join(d, g["snapshot_file"])).read().split("\n")
    assert len(snap) == 17 and snap[16] == ""
    for i, line in g["snapshot_lines"].items():
        assert norm(snap[int(i)]) == line, (i, snap[int(i)], line)
    rest = open(os.path.join(d, "restart_fields.dat")).read().split("\n")
    assert len(rest) == 17 and rest[16] == ""
    for line in rest[:16]:
        assert norm(line, padded=False) == g["restart_line"], line
