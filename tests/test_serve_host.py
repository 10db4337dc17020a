"""Host logic of the serving front (no GPU)."""


def test_structure_key_groups_by_chain_structure(configs):
    import serve
    a = configs.chem_droplet(lx=24.0, chi_ab=15.0)
    b = configs.chem_droplet(lx=40.0, chi_ab=30.0, chi_ac=50.0)
    c = configs.params_c5_3d()
    assert serve.structure_key(a) == serve.structure_key(b)      # LX and chi may differ
    assert serve.structure_key(a) != serve.structure_key(c)
