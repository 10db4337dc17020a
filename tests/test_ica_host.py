"""Host logic of the ICA workflow (no GPU): neighbours, loop order, plane
parameters, seeded trial perturbations and the per-point records."""
import json
import os

import numpy as np


def _ica(tmp_path):
    import ica
    return ica.ICA(None, (10.0, 12.0, 14.0), (20.0, 25.0), str(tmp_path), delta=8.0,
                   base=dict(lx=8.0, lz=6.0, dr=1.0, ns=11))


def test_neighbours_in_plane_and_cross_plane(tmp_path):
    w = _ica(tmp_path)
    pt = w.points[(0, 1, 0)]
    keys = [n.key for n in w.neighbours(pt)]
    # in-plane i-1, i+1, (j-1 missing), j+1, then previous and next plane
    assert keys == [(0, 0, 0), (0, 2, 0), (0, 1, 1), (2, 1, 0), (1, 1, 0)]
    corner = w.points[(2, 2, 1)]
    assert [n.key for n in w.neighbours(corner)] == [(2, 1, 1), (2, 2, 0), (1, 2, 1), (0, 2, 1)]
    # every relation is symmetric
    for p in w.points.values():
        for n in w.neighbours(p):
            assert p in w.neighbours(n)


def test_order_and_plane_parameters(tmp_path):
    import abi
    w = _ica(tmp_path)
    assert [p.key for p in w.order()][:4] == [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1)]
    a, c, s1 = abi.field_index("A"), abi.field_index("C"), abi.field_index("S1")
    for pl, d in ((0, 0.0), (1, 8.0), (2, -8.0)):
        p = w.points[(pl, 2, 1)].params
        assert p.chi[a][abi.field_index("B")] == 14.0 and p.chi[a][c] == 25.0
        assert p.chi[s1][a] == 25.0 + d and p.chi[s1][abi.field_index("B")] == 25.0 + d


def test_trial_seed_is_deterministic_and_only_on_active_fields(tmp_path):
    w = _ica(tmp_path)
    nb, pt = w.points[(0, 0, 0)], w.points[(0, 1, 0)]
    V = 48
    wf = np.zeros((13, V))
    wf[0] = 1.0
    wf[2] = -2.0
    nb.field = w._save_field("nb", (wf, np.zeros(V)))
    a = w._trial_seed(pt, nb)[0]
    b = w._trial_seed(pt, nb)[0]
    assert np.array_equal(a, b)
    assert np.all(a[1] == 0.0) and np.all(a[3:] == 0.0)
    assert 0 < np.abs(a[0] - 1.0).max() <= 1e-3 and 0 < np.abs(a[2] + 2.0).max() <= 1e-3
    nb.version += 1
    assert not np.array_equal(w._trial_seed(pt, nb)[0], a)


def test_records_hold_state_and_history(tmp_path):
    w = _ica(tmp_path)
    pt = w.points[(1, 2, 0)]
    pt.F, pt.field, pt.status = 3.5, "fields/x.npz", "converged"
    pt.history.append({"iteration": 0, "F": 3.6, "field": "fields/y.npz"})
    pt.tried.add((0, 2, 0, 1))
    w.write_records()
    files = [f for f in os.listdir(str(tmp_path)) if f.endswith(".json")]
    assert len(files) == 18
    r = json.load(open(os.path.join(str(tmp_path), "P1_2_0.json")))
    assert r["F"] == 3.5 and r["chi_AB_N"] == 14.0 and r["chi_AC_N"] == 20.0
    assert r["history"] == [{"iteration": 0, "F": 3.6, "field": "fields/y.npz"}]
    assert r["tried"] == ["P0-2-0@1"]
