"""Host side of the served-pool checkpoint format (no GPU)."""
import numpy as np


def test_restart_file_round_trip(tmp_path):
    import serve
    rng = np.random.default_rng(3)
    w, pf = rng.standard_normal((13, 40)), rng.standard_normal(40)
    path = str(tmp_path / "x.bin")
    serve._write_rst(path, 17, w, pf)
    it, w2, p2 = serve._read_rst(path)
    assert it == 17 and np.array_equal(w, w2) and np.array_equal(pf, p2)
    raw = open(path, "rb").read()
    assert raw[:8] == b"SCFTRST1" and len(raw) == 24 + 14 * 40 * 8


def test_request_ids_survive_json():
    import json
    import serve
    for rid in (3, "a", ("trial", 2, (0, 1, 1), 4), ("polish", "t", 0, 1)):
        assert serve._from_json(json.loads(json.dumps(serve._jsonable(rid)))) == rid
