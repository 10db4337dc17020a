"""Serving front over several chain structures (2D and 3D requests in one
stream, the config-5 pattern): requests arrive while others run, each is
admitted into the pool of its structure, and each result is the one the
same request gives alone."""
import pytest

pytestmark = pytest.mark.gpu


def _requests(configs, serve):
    p2 = lambda chi: configs.chem_droplet(lx=12.0, lz=10.0, dr=1.0, ns=21, vol_dbk=0.2, vol_s1=0.05, len_s1=0.15,
                                          g1=0.02, chi_ab=chi)
    p3 = lambda chi: configs.chem_droplet(lx=6.0, ly=5.0, lz=4.0, dr=1.0, ns=11, vol_dbk=0.2, vol_s1=0.05,
                                          vol_homo_a=0.1, len_homo_a=0.5, len_s1=0.3, g1=0.02, chi_ab=chi)
    out = []
    for k, (mk, chi, t) in enumerate([(p2, 12.0, 0.0), (p3, 12.0, 0.0), (p2, 14.0, 0.02), (p2, 16.0, 0.05),
                                      (p3, 15.0, 0.08), (p2, 13.0, 0.3)]):
        p = mk(chi)
        r = serve.Request(k, p, (serve.SEED_DROPLET, {"radius": 3.0, "seed": k}), arrival=t)
        out.append(r)
    return out


def test_mixed_structures_stream_matches_alone(abi, configs, lib):
    import serve
    srv = serve.Server(lib, precision=abi.SCFT_FP64, default_slots=4, max_iters=30000)
    reqs = _requests(configs, serve)
    srv.run_stream(reqs)
    assert len(srv.pools) == 2                       # one pool per structure
    done = {r.rid: r for r in srv.done()}
    assert sorted(done) == list(range(len(reqs)))
    srv.close()
    for r in _requests(configs, serve):
        alone = serve.Pool(lib, r.params, precision=abi.SCFT_FP64, max_slots=1, max_iters=30000)
        r.arrival = None
        alone.submit(r)
        alone.run()
        a = alone.done[0]
        m = done[r.rid]
        print(r.rid, m.status, m.iters, m.F)
        assert a.status == m.status and a.iters == m.iters
        assert abs(a.F - m.F) <= 1e-12 * abs(a.F)
        alone.close()
