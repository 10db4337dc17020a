"""Host logic of the serving front (no GPU)."""


def test_structure_key_groups_by_chain_structure(configs):
    import serve
    a = configs.chem_droplet(lx=24.0, chi_ab=15.0)
    b = configs.chem_droplet(lx=40.0, chi_ab=30.0, chi_ac=50.0)
    c = configs.params_c5_3d()
    assert serve.structure_key(a) == serve.structure_key(b)      # LX and chi may differ
    assert serve.structure_key(a) != serve.structure_key(c)


def test_iteration_bytes_match_the_accounting_table(abi, configs, lib):
    """bench.iteration_bytes: 2 b V S + 2 b V P + 8 V (3 n + 2) with S, P, n
    from the derived chain structure; C2 and C3 in FP32-split give the
    per-instance-iteration bytes of the hot-path accounting (142.1 MB with
    S = 316, P = 212, n = 4; 10.65 GB with S = 734, P = 521)."""
    import bench
    b, t = bench.iteration_bytes(abi, lib, configs.params_c2(), 4.0)
    assert t == {"V": 32768, "S": 316, "P": 212, "n": 4}
    assert b == 142082048.0
    b3, t3 = bench.iteration_bytes(abi, lib, configs.params_c3(), 4.0)
    assert (t3["S"], t3["P"], t3["n"]) == (734, 521, 4)
    assert abs(b3 - 10.645e9) < 1e6
