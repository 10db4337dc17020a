"""Pins for the CPU oracle (oracle/scft_oracle.c), run without a GPU.

The reference ships no tests or golden vectors, and SCFT.c cannot be built
here (it needs FFTW3, which the image lacks), so every check below comes
from mathematics the reference's algorithm must satisfy: brute-force and
numpy DFTs, closed forms for uniform fields and sine modes, contour
invariants of the symmetric split-operator scheme, glibc's own drand48.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest


# ------------------------------------------------------------------- FFT

@pytest.mark.parametrize("shape", [(8, 1, 8), (4, 3, 258), (6, 5, 130), (1, 1, 43), (2, 7, 98), (5, 1, 64),
                                   (3, 1, 129), (1, 1, 1), (2, 2, 2)])
def test_fft3d_matches_numpy(orc, shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    a = orc.fft3d(x, -1)
    b = np.fft.fftn(x)
    assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())
    c = orc.fft3d(x, +1)
    d = np.fft.ifftn(x) * x.size
    assert np.abs(c - d).max() <= 1e-13 * max(1.0, np.abs(d).max())


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 12, 13, 16, 43, 60, 129, 258])
def test_fft_line_matches_brute_force(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        a = orc.fft3d(x.reshape(1, 1, n), sign).reshape(n)
        b = orc.dft_brute(x, sign)
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_fft_parseval_and_delta(orc):
    n = (4, 3, 258)
    x = np.zeros(n, complex)
    x[0, 0, 0] = 1.0
    assert np.allclose(orc.fft3d(x, -1), 1.0, atol=1e-15)   # delta <-> constant
    rng = np.random.default_rng(5)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Y = orc.fft3d(y, -1)
    assert abs(np.vdot(Y, Y).real / y.size - np.vdot(y, y).real) <= 1e-11 * np.vdot(y, y).real


# ------------------------------------------------------------ setup rules

def _derive_both(orc, lib, abi, p):
    rc1, d1 = orc.derive(p)
    d2 = abi.ScftDerived()
    rc2 = lib.scft_derive(C.byref(p), C.byref(d2))
    return rc1, d1, rc2, d2


def _param_sets(abi, configs):
    out = [configs.params_c1(), configs.params_c2(), configs.params_c3(), configs.params_c5_3d(),
           configs.t_bulk_small(), configs.t_film_small(), configs.t_3d_small(),
           configs.t_all_species(1), configs.t_all_species(0)]
    rng = np.random.default_rng(7)
    for _ in range(20):
        vd = rng.dirichlet(np.ones(14)) * (rng.random(14) < 0.5)
        vd = vd / vd.sum()
        chi = {}
        names = list("ABCDEFGHIJ") + ["S1", "S2", "S3"]
        for a in range(13):
            for b in range(a + 1, 13):
                if rng.random() < 0.4:
                    chi[names[a] + names[b]] = float(rng.uniform(0, 50))
        out.append(abi.make_params(lx=float(rng.choice([1.0, 4.0, 6.0])), ly=float(rng.choice([1.0, 3.0])),
                                   lz=float(rng.choice([2.0, 5.0])), dr_target=0.5, ns=int(rng.choice([21, 51, 101, 100])),
                                   mask=int(rng.integers(0, 2)), chi=chi,
                                   f_dbk=tuple(rng.uniform(0.2, 0.8, 4)), len_dbk=tuple(rng.uniform(0.5, 1.0, 4)),
                                   vol_dbk=tuple(vd[:4]), vol_homo=tuple(vd[4:]),
                                   len_homo=tuple(rng.uniform(0.3, 1.2, 10)),
                                   vol_s=tuple(rng.uniform(0, 0.05, 3) * (rng.random(3) < 0.5)),
                                   len_s=tuple(rng.uniform(0.1, 0.5, 3))))
    return out


def test_product_setup_matches_oracle(orc, lib, abi, configs):
    """libscft's scft_derive (host code, no device) against the oracle's
    restatement of SCFT.c:84-396, 718-734, 2001-2042."""
    for p in _param_sets(abi, configs):
        rc1, d1, rc2, d2 = _derive_both(orc, lib, abi, p)
        if rc2 == -1:
            # the product rejects chains with fewer than 3 contour points
            # (the reference would divide by zero in DS_EFF)
            assert any(0 < n < 3 for n in d1.ns_sp) or any(
                d1.ns_sp[b] and (d1.ns_x[b] < 1 or d1.ns_y[b] < 1) for b in range(4))
            continue
        assert rc1 == rc2 == 0
        for f in ("nx", "ny", "nz", "nzext", "coating_vol", "n_active"):
            assert getattr(d1, f) == getattr(d2, f), f
        assert list(d1.ns_sp) == list(d2.ns_sp)
        for b in range(4):
            if d1.ns_sp[b]:
                assert (d1.ns_x[b], d1.ns_y[b]) == (d2.ns_x[b], d2.ns_y[b])
        assert list(d1.alpha) == list(d2.alpha)
        assert np.allclose(list(d1.phi_avg), list(d2.phi_avg), rtol=0, atol=1e-15)
        assert [list(r) for r in d1.chi_eff] == [list(r) for r in d2.chi_eff]
        assert list(d1.active) == list(d2.active)


def test_setup_known_values(orc, abi, configs):
    # NS convention: NS 101 -> dbk (101, 51, 50); a literal 100 -> (99, 49, 50)
    rc, d = orc.derive(configs.params_c1())
    assert (d.nx, d.ny, d.nz) == (64, 1, 64)
    assert (d.ns_sp[0], d.ns_x[0], d.ns_y[0]) == (101, 51, 50)
    p = configs.params_c1()
    p.ns = 100
    rc, d = orc.derive(p)
    assert (d.ns_sp[0], d.ns_x[0], d.ns_y[0]) == (99, 49, 50)
    rc, d = orc.derive(configs.params_c2())
    assert (d.nx, d.nz, d.nzext) == (256, 128, 258)
    assert d.ns_sp[14] == 9 and d.ns_sp[6] == 101
    assert abs(d.phi_avg[2] - 0.85 * 0.95) < 1e-15     # coating rescale
    assert d.phi_avg[10] == 0.05
    rc, d = orc.derive(configs.params_c3())
    assert d.ns_sp[4] == 99 and d.ns_sp[14] == 19 and d.nzext == 130
    # volume check (SCFT.c:84-101)
    p = configs.params_c1()
    p.vol_dbk[0] = 0.9
    rc, d = orc.derive(p)
    assert rc == -4
    # brushes with MASK 0 are active in the update but carry no chain
    rc, d = orc.derive(configs.t_all_species(0))
    assert d.active[10] == d.active[11] == d.active[12] == 1
    assert d.ns_sp[14] == 0 and d.chi_eff[10][0] == 0.0


# ---------------------------------------------------------------- drand48

def test_drand48_matches_glibc(orc, abi, configs):
    libc = C.CDLL("libc.so.6")
    libc.drand48.restype = C.c_double
    libc.seed48.restype = C.POINTER(C.c_ushort)
    libc.seed48.argtypes = [C.POINTER(C.c_ushort)]
    st = orc.state(configs.t_bulk_small())
    for seed in [(0, 0, 0), (1, 2, 3), (65535, 12345, 999)]:
        arr = (C.c_ushort * 3)(*seed)
        libc.seed48(arr)
        st.seed48(seed)
        a = [libc.drand48() for _ in range(1000)]
        b = [st.drand48() for _ in range(1000)]
        assert a == b
        # the state the reference writes to restart_random.dat (SCFT.c:1652-1663)
        cur = (C.c_ushort * 3)(0, 0, 0)
        old = libc.seed48(cur)
        assert tuple(old[i] for i in range(3)) == st.get_seed()


def test_gauss_is_standard_normal(orc, configs):
    from scipy import stats
    st = orc.state(configs.t_bulk_small())
    st.seed48((4, 5, 6))
    x = np.array([st.gauss() for _ in range(40000)])
    assert abs(x.mean()) < 0.02 and abs(x.std() - 1.0) < 0.02
    assert stats.kstest(x, "norm").pvalue > 1e-3


def _glibc_drand48():
    libc = C.CDLL("libc.so.6")
    libc.drand48.restype = C.c_double
    libc.seed48.restype = C.POINTER(C.c_ushort)
    libc.seed48.argtypes = [C.POINTER(C.c_ushort)]
    return libc


def _polar(drand48):
    """funGaussDev as SCFT.c:1769-1782 states it: uniform pairs on
    (-1, 1)^2 until x1^2 + x2^2 < 1, then x1 * sqrt(-2 ln w / w)."""
    while True:
        x1 = 2.0 * drand48() - 1.0
        x2 = 2.0 * drand48() - 1.0
        w = x1 * x1 + x2 * x2
        if not w >= 1.0:
            break
    return x1 * math.sqrt((-2.0 * math.log(w)) / w)


@pytest.mark.parametrize("seed", [(0, 0, 0), (7, 8, 9), (65535, 4242, 31337)])
def test_gauss_matches_polar_method_on_glibc_drand48(orc, configs, seed):
    """The oracle's deviates equal the reference's polar method driven by
    glibc's own drand48, bit for bit, and leave the same generator state
    (what restart_random.dat records, SCFT.c:1652-1663)."""
    libc = _glibc_drand48()
    libc.seed48((C.c_ushort * 3)(*seed))
    want = [_polar(libc.drand48) for _ in range(10000)]
    st = orc.state(configs.t_bulk_small())
    st.seed48(seed)
    got = st.gauss_n(10000)
    assert [float(x) for x in got] == want
    old = libc.seed48((C.c_ushort * 3)(0, 0, 0))
    assert tuple(old[i] for i in range(3)) == st.get_seed()


def test_main_writer_lines_match_hand_fixture(orc, abi, tmp_path):
    """orc_main's output files on a hand-solved state equal the lines
    written out by hand from the reference's printf formats
    (tests/golden/writers_expected.json: column order, widths, separators,
    i-fastest voxel order).  restart_random.dat holds the state glibc's
    drand48 reaches after the 5 x 16 deviates of the iteration."""
    import paramsfile
    from _cases import writers_case, check_writer_files
    p, w, pf, g = writers_case(abi)
    rc, d = orc.derive(p)
    assert [round(d.alpha[4 + h], 14) for h in range(5)] == g["alpha"]
    paramsfile.write_inputs(str(tmp_path), p, w, pf, None, counter=1, seed=(0, 0, 0))
    assert orc.main(p, 1, 2, 1, str(tmp_path)) == 0
    check_writer_files(str(tmp_path), g)
    libc = _glibc_drand48()
    libc.seed48((C.c_ushort * 3)(0, 0, 0))
    for _ in range(5 * 16):
        _polar(libc.drand48)
    old = libc.seed48((C.c_ushort * 3)(0, 0, 0))
    assert open(str(tmp_path / "restart_random.dat")).read() == "%d %d %d" % tuple(old[i] for i in range(3))


# ------------------------------------------------------------- propagator

def test_uniform_field_closed_form(orc, abi):
    p = abi.make_params(lx=8.0, ly=1.0, lz=6.0, dr_target=0.5, ns=21, mask=0,
                        vol_homo=(0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0))
    st = orc.state(p)
    w = np.zeros((13, st.V))
    w[2] = 0.7
    st.set_fields(w, np.zeros(st.V))
    o = st.iterate(2)
    for s in range(21):
        assert np.abs(st.get_propagator(6, 0, s) - math.exp(-0.7 * s / 20)).max() <= 1e-14
    assert abs(o.chempot[6] - 0.7) <= 1e-13                    # -ln Q = w alpha
    phi, tot = st.get_densities()
    assert np.abs(phi[2] - 1.0).max() <= 1e-13


def test_dirichlet_sine_mode_decay(orc, abi, configs):
    """With MASK 1 and w = 0 the brush propagator started from
    sin(pi m (k+1)/(NZ+1)) cos(2 pi mx i / NX) decays by Kz(m) Kx(mx) per
    step with the extended grid N = 2(NZ+1) (SCFT.c:680, 1000-1004)."""
    p = configs.chem_droplet(lx=8.0, ly=1.0, lz=5.0, dr=0.5, ns=21, vol_dbk=0.1, vol_s1=0.05, len_s1=0.5)
    st = orc.state(p)
    nx, ny, nz = st.shape
    st.set_fields(np.zeros((13, st.V)), np.zeros(st.V))
    m, mx = 2, 3
    i = np.arange(nx)[:, None]
    k = np.arange(nz)[None, :]
    q0 = (np.sin(np.pi * m * (k + 1) / (nz + 1)) * np.cos(2 * np.pi * mx * i / nx)).reshape(-1)
    mask = np.zeros((3, st.V))
    mask[0] = q0
    st.set_mask(mask)
    st.iterate(2)
    ds = 1.0 / 20
    fac = math.exp(-ds / (p.lz / nz) ** 2 * (2 * math.pi * m / (2 * (nz + 1))) ** 2) * \
        math.exp(-ds / (p.lx / nx) ** 2 * (2 * math.pi * mx / nx) ** 2)
    for s in range(st.derived.ns_sp[14]):
        assert np.abs(st.get_propagator(14, 0, s) - fac ** s * q0).max() <= 1e-14


def test_periodic_mode_decay(orc, abi):
    """MASK 0: q = 1 + eps cos(2 pi m k / NZ) cos(2 pi mx i / NX) under w = 0
    would need an injected start; instead use the block semigroup: two
    half blocks of the same field equal one full block."""
    p = abi.make_params(lx=6.0, ly=1.0, lz=4.0, dr_target=0.5, ns=21, mask=0,
                        vol_homo=(0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0))
    st = orc.state(p)
    rng = np.random.default_rng(2)
    w = rng.uniform(-1, 1, st.V)
    q = np.zeros((21, st.V))
    q[0] = rng.uniform(0.5, 1.5, st.V)
    full = st.pde_block(0, 20, 21, q.copy(), w, 0)
    half = st.pde_block(0, 10, 21, q.copy(), w, 0)
    half = st.pde_block(10, 20, 21, half, w, 0)
    assert np.abs(full - half).max() <= 1e-13
    assert (full > 0).all()


def test_operator_symmetry_q_equals_qdagger(orc, abi, configs):
    """L = E M E is symmetric, so Q from q and from q-dagger agree, and
    sum_r q(s) q-dagger(NS-1-s) = V Q for every s (any w)."""
    for p in (configs.t_all_species(1), configs.t_3d_small(), configs.params_c1()):
        st = orc.state(p)
        d = st.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        import configs as cf
        w, pf = cf.random_fields(p, seed=1, amp=0.8, fields=act)
        st.set_fields(w, pf)
        if p.mask:
            st.set_mask(cf.brush_mask(p))
        o = st.iterate(2)
        for sp in range(17):
            ns = d.ns_sp[sp]
            if ns <= 0:
                continue
            q = [st.get_propagator(sp, 0, s) for s in range(ns)]
            qd = q if 4 <= sp < 14 else [st.get_propagator(sp, 1, s) for s in range(ns)]
            Q = o.qsum[sp]
            for s in range(ns):
                assert abs(np.dot(q[s], qd[ns - 1 - s]) / st.V - Q) <= 1e-12 * Q, (sp, s)


def test_mean_density_equals_volume(orc, abi, configs):
    """Simpson with DS_EFF = 1/(NS_p-1) over the whole chain integrates the
    constant V Q exactly, so mean(phi_p) = vol_p (1 - coating) (brushes:
    vol_s) for any field; per block the share is (NS_X-1)/(NS_p-1)."""
    for p in (configs.t_all_species(1), configs.t_all_species(0), configs.t_film_small()):
        st = orc.state(p)
        d = st.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        import configs as cf
        w, pf = cf.random_fields(p, seed=2, amp=1.0, fields=act)
        st.set_fields(w, pf)
        if p.mask:
            st.set_mask(cf.brush_mask(p))
        st.iterate(2)
        phi, tot = st.get_densities()
        cv = d.coating_vol
        exp = np.zeros(13)
        pairs = [(0, 1), (3, 4), (5, 6), (7, 8)]
        for b in range(4):
            if d.ns_sp[b]:
                share = (d.ns_x[b] - 1) / (d.ns_sp[b] - 1)
                exp[pairs[b][0]] += p.vol_dbk[b] * (1 - cv) * share
                exp[pairs[b][1]] += p.vol_dbk[b] * (1 - cv) * (1 - share)
        for h in range(10):
            if d.ns_sp[4 + h]:
                exp[h] += p.vol_homo[h] * (1 - cv)
        for s in range(3):
            if d.ns_sp[14 + s]:
                exp[10 + s] += p.vol_s[s]
        assert np.abs(phi.mean(axis=1) - exp).max() <= 1e-12
        assert abs(tot.mean() - exp.sum()) <= 1e-12


# ------------------------------------------------------- free energy, update

def test_free_energy_closed_form_uniform(orc, abi, configs):
    """Periodic, w = p = 0: q = 1, Q = 1, phi = phi_avg, so
    F = sum_{f<g} chi_fg phi_f phi_g; with uniform w_f = c_f,
    ln Q_p = -sum_blocks c alpha_block."""
    p = configs.t_all_species(0)
    st = orc.state(p)
    d = st.derived
    chi = np.array([list(r) for r in d.chi_eff])
    ds = 1.0 / (p.ns - 1)
    pairs = [(0, 1), (3, 4), (5, 6), (7, 8)]
    for c, pres in ((np.zeros(13), 0.0), (np.linspace(-0.5, 0.7, 13), 0.3)):
        st.set_fields(np.repeat(c[:, None], st.V, axis=1), np.full(st.V, pres))
        o = st.iterate(2)
        lnq = 0.0
        phi = np.zeros(13)
        for b in range(4):
            ns, nx = d.ns_sp[b], d.ns_x[b]
            a, bb = pairs[b]
            lq = -(c[a] * (nx - 1) + c[bb] * (ns - nx)) * ds
            assert abs(math.log(o.qsum[b]) - lq) <= 1e-12
            lnq += p.vol_dbk[b] / d.alpha[b] * lq
            # block shares of a discrete chain: (NS_X - 1) / (NS_p - 1)
            phi[a] += p.vol_dbk[b] * (nx - 1) / (ns - 1)
            phi[bb] += p.vol_dbk[b] * (ns - nx) / (ns - 1)
        for h in range(10):
            ns = d.ns_sp[4 + h]
            lq = -c[h] * (ns - 1) * ds
            lnq += p.vol_homo[h] / d.alpha[4 + h] * lq
            phi[h] += p.vol_homo[h]
        F = 0.5 * phi @ chi @ phi - c @ phi - pres * (1 - phi.sum()) - lnq
        assert abs(o.free_energy - F) <= 1e-11, (o.free_energy, F)


def test_update_matrix_form(orc, abi, configs):
    """funUpdate on a uniform state, against its matrix form:
    dW = -w + p + X r (X = chi restricted to the active set, zero diagonal),
    w' = w + (G1 I + G1/6 (J - I)) dW, p' = (sum w' + corr)/n - mean,
    corr = sum_{f<g} chi_fg (R - r_f - r_g), R = sum r."""
    p = configs.t_all_species(0)
    st = orc.state(p)
    d = st.derived
    act = [f for f in range(13) if d.active[f]]
    n = len(act)
    rng = np.random.default_rng(3)
    c = rng.uniform(-1, 1, 13)
    w = np.repeat(c[:, None], st.V, axis=1)
    st.set_fields(w, np.full(st.V, 0.2))
    st.iterate(2)
    w1, p1 = st.get_fields()
    phi, _ = st.get_densities()
    r = phi[act, 0] - np.array([d.phi_avg[f] for f in act])
    X = np.array([[d.chi_eff[f][g] for g in act] for f in act])
    dW = -c[act] + 0.2 + X @ r
    G = p.g1 / 6.0 * np.ones((n, n)) + (p.g1 - p.g1 / 6.0) * np.eye(n)
    wn = c[act] + G @ dW
    R = r.sum()
    corr = sum(X[i, j] * (R - r[i] - r[j]) for i in range(n) for j in range(i + 1, n))
    pn = (wn.sum() + corr) / n
    assert np.abs(w1[act, 0] - wn).max() <= 1e-12
    assert np.abs(p1).max() <= 1e-12             # uniform p minus its mean
    # non-uniform: mean(p) = 0 after the update (SCFT.c:2119-2123)
    import configs as cf
    w, pf = cf.random_fields(p, seed=4, amp=0.5, fields=[abi.FIELD_NAMES[f] for f in act])
    st.set_fields(w, pf)
    st.iterate(3)
    _, p2 = st.get_fields()
    assert abs(p2.mean()) <= 1e-14
    del pn


def test_update_no_change_at_mean_field(orc, abi, configs):
    """phi = phi_avg and w = p = 0 is a fixed point of the update."""
    p = configs.params_c1()
    st = orc.state(p)
    st.set_fields(np.zeros((13, st.V)), np.zeros(st.V))
    o = st.iterate(2)
    w, pf = st.get_fields()
    assert np.abs(w).max() <= 1e-14 and np.abs(pf).max() <= 1e-14 and o.dw_rms <= 1e-14


# ------------------------------------------------------------- file I/O

def _write_restart(path, w, pf, shape):
    nx, ny, nz = shape
    with open(path, "w") as f:
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    idx = k + nz * (j + ny * i)
                    f.write(" ".join("%.17g" % w[c, idx] for c in range(13)) + " %.17g \n" % pf[idx])


def test_main_writes_reference_formats(orc, abi, configs, tmp_path):
    p = configs.t_film_small()
    st = orc.state(p)
    shape = st.shape
    w, pf = configs.droplet_fields(p, radius=3.0)
    d = str(tmp_path)
    _write_restart(os.path.join(d, "restart_fields.dat"), w, pf, shape)
    open(os.path.join(d, "restart_random.dat"), "w").write("0 0 0")
    open(os.path.join(d, "counter.dat"), "w").write("1")
    m = configs.brush_mask(p)
    with open(os.path.join(d, "wMask.dat"), "w") as f:
        nx, ny, nz = shape
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    idx = k + nz * (j + ny * i)
                    f.write("%g %g %g\n" % (m[0, idx], m[1, idx], m[2, idx]))
    assert orc.main(p, 4, 2, 1, d) == 0
    # iterations 2..5 (SCFT.c:1088-1090: runs until iIter = ITER + 1)
    fe = open(os.path.join(d, "free_energy.dat")).read().splitlines()
    assert [int(l.split()[0]) for l in fe] == [2, 3, 4, 5]
    assert fe[0].endswith(" ") and "  " in fe[0]
    en = open(os.path.join(d, "energies.dat")).read().splitlines()
    assert len(en[0].split()) == 18
    dv = open(os.path.join(d, "deviations.dat")).read().splitlines()
    assert len(dv[0].split()) == 5 and float(dv[0].split()[3]) == 40.0
    snaps = sorted(f for f in os.listdir(d) if f.startswith("densities_dbk_"))
    assert snaps == ["densities_dbk_100000002.dat", "densities_dbk_100000004.dat"]
    lines = open(os.path.join(d, snaps[0])).read().splitlines()
    assert len(lines) == st.V and len(lines[0].split()) == 30
    assert open(os.path.join(d, "counter.dat")).read() == "4"
    rr = open(os.path.join(d, "restart_random.dat")).read().split()
    assert len(rr) == 3
    run = open(os.path.join(d, "run.txt")).read()
    assert "NX, NY, NZ  = 12, 1, 10" in run and "MASK = 1, vol S[1+2+3]: 0.050000" in run


def test_main_volume_check_error_file(orc, configs, tmp_path):
    p = configs.params_c1()
    p.vol_dbk[0] = 0.5
    assert orc.main(p, 2, 1, 1, str(tmp_path)) == -4
    assert open(os.path.join(str(tmp_path), "DENSITY_CHECK_FAIL.error")).read() == "0.500000"


# ------------------------------------------------------- ABI exports (no GPU)

def test_library_exports_every_declared_symbol(lib):
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    names = set()
    for fn in os.listdir(os.path.join(root, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(root, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char \*)\s*(scft_\w+)\s*\(", txt, re.M))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(lib, n), n
