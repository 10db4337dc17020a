"""Workloads: parameter sets and initial states.

C1..C5 are the benchmark configurations of this project (2D bulk lamellae,
2D thin-film nanodroplet, 3D thin-film droplet, the batched phase-diagram
sweep and the served stream); the small cases t_* are parity-test sizes.

Initial states are produced by this module's own documented generator
(numpy PCG64 with an explicit seed) and injected the way the reference
accepts injected fields: as restart fields (SCFT.c:844-866).
"""
import numpy as np

from abi import make_params, grid_of, NF

# ---------------------------------------------------------------- chemistry


def chem_droplet(chi_ab=25.0, chi_ac=40.0, chi_bc=None, chi_s1a=None, chi_s1b=None,
                 vol_dbk=0.15, vol_s1=0.05, vol_homo_a=0.0, len_homo_a=0.5, ns=101,
                 lx=32.0, ly=1.0, lz=16.0, dr=0.125, g1=0.01, len_s1=0.1):
    """Thin-film nanodroplet chemistry: symmetric AB diblock in a homopolymer
    C matrix over an S1 brush grafted at z = 0 whose chemistry equals C
    (chi_S1X = chi_CX, chi_S1C = 0).  Optional short homopolymer A taken
    from the C volume."""
    chi_bc = chi_ac if chi_bc is None else chi_bc
    chi_s1a = chi_ac if chi_s1a is None else chi_s1a
    chi_s1b = chi_bc if chi_s1b is None else chi_s1b
    vol_c = 1.0 - vol_dbk - vol_homo_a
    vh = [0.0] * 10
    lh = [1.0] * 10
    vh[0] = vol_homo_a
    lh[0] = len_homo_a
    vh[2] = vol_c
    return make_params(lx=lx, ly=ly, lz=lz, dr_target=dr, ns=ns, mask=1, g1=g1,
                       chi={"AB": chi_ab, "AC": chi_ac, "BC": chi_bc,
                            "S1A": chi_s1a, "S1B": chi_s1b, "S1C": 0.0},
                       f_dbk=(0.5, 0.5, 0.3, 0.5), vol_dbk=(vol_dbk, 0, 0, 0),
                       vol_homo=tuple(vh), len_homo=tuple(lh),
                       vol_s=(vol_s1, 0.0, 0.0), len_s=(len_s1, 0.1, 0.1))


def params_c1():
    """2D bulk symmetric AB diblock melt, 64x1x64, chi_AB N = 20, G1 = 0.1."""
    return make_params(lx=8.0, ly=1.0, lz=8.0, dr_target=0.125, ns=101, mask=0, g1=0.1,
                       chi={"AB": 20.0}, vol_dbk=(1.0, 0, 0, 0))


def params_c2(**kw):
    """2D thin-film nanodroplet, 256x1x128, NS 101."""
    return chem_droplet(**kw)


def params_c3(**kw):
    """3D thin-film droplet, 128x128x64 (LX=LY=16, LZ=8), NS 201, plus homo A."""
    args = dict(lx=16.0, ly=16.0, lz=8.0, ns=201, vol_homo_a=0.05, len_homo_a=0.5)
    args.update(kw)
    return chem_droplet(**args)


def params_c5_3d(**kw):
    """3D served request, 96x96x48 (LX=LY=12, LZ=6), C3 chemistry."""
    args = dict(lx=12.0, ly=12.0, lz=6.0, ns=201, vol_homo_a=0.05, len_homo_a=0.5)
    args.update(kw)
    return chem_droplet(**args)


# ------------------------------------------------------------ initial states


def z_coords(p):
    """Voxel centres along z: the reference's Dirichlet extension puts the
    walls at z = 0 and z = (NZ+1)*DZ, so voxel k sits at (k+1)*DZ (MASK 1);
    for MASK 0 voxel k sits at k*DZ."""
    nx, ny, nz = grid_of(p)
    dz = p.lz / nz
    return (np.arange(nz) + (1.0 if p.mask else 0.0)) * dz


def random_fields(p, seed=0, amp=0.5, fields=("A", "B")):
    """w of the listed fields i.i.d. uniform in [-amp, amp) (PCG64(seed)),
    every other column 0, p = 0."""
    nx, ny, nz = grid_of(p)
    V = nx * ny * nz
    rng = np.random.Generator(np.random.PCG64(seed))
    w = np.zeros((NF, V))
    from abi import field_index
    for f in fields:
        w[field_index(f)] = rng.uniform(-amp, amp, V)
    return w, np.zeros(V)


def droplet_fields(p, radius=6.0, seed=0, noise=0.1, centre=None, spherical=None):
    """Hemispherical (2D: half-disc; 3D: spherical-cap) droplet of diblock on
    the z = 0 wall, phi_dbk = 1 inside + U(-noise, noise) (PCG64(seed)),
    clipped to [0, 1]; A and B take half each, C (and short homo A, if
    present, in proportion to its volume) fill the rest.  Converted to w by
    w_f = sum_g chi_fg phi_g over the active fields; p = 0."""
    nx, ny, nz = grid_of(p)
    V = nx * ny * nz
    dx, dy = p.lx / nx, p.ly / ny
    x = np.arange(nx) * dx
    y = np.arange(ny) * dy
    z = z_coords(p)
    cx = p.lx / 2.0 if centre is None else centre[0]
    cy = p.ly / 2.0 if (centre is None or len(centre) < 2) else centre[1]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    r2 = (X - cx) ** 2 + (Z) ** 2
    if ny > 1:
        r2 = r2 + (Y - cy) ** 2
    inside = (r2 < radius * radius).astype(np.float64)
    rng = np.random.Generator(np.random.PCG64(seed))
    phid = np.clip(inside + rng.uniform(-noise, noise, inside.shape), 0.0, 1.0)
    coat = p.vol_s[0] + p.vol_s[1] + p.vol_s[2] if p.mask else 0.0
    free = 1.0 - coat
    phi = np.zeros((NF, V))
    phi[0] = 0.5 * phid.reshape(-1) * free
    phi[1] = 0.5 * phid.reshape(-1) * free
    va, vc = p.vol_homo[0], p.vol_homo[2]
    rest = (1.0 - phid.reshape(-1)) * free
    if va + vc > 0:
        phi[0] += rest * va / (va + vc)
        phi[2] = rest * vc / (va + vc)
    chi = np.array([[p.chi[i][j] for j in range(NF)] for i in range(NF)])
    np.fill_diagonal(chi, 0.0)
    w = chi @ phi
    active = np.zeros(NF, bool)
    active[[0, 1]] = p.vol_dbk[0] > 0
    for h in range(10):
        active[h] |= p.vol_homo[h] > 0
    for s in range(3):
        active[10 + s] |= p.vol_s[s] > 0
    w[~active] = 0.0
    return w, np.zeros(V)


def lamella_fields(p, periods=1.0, amp=2.0, axis="x"):
    """A lamellar seed: w_A = amp*cos(2*pi*periods*x/L), w_B = -w_A."""
    nx, ny, nz = grid_of(p)
    V = nx * ny * nz
    if axis == "x":
        c = np.cos(2 * np.pi * periods * np.arange(nx) / nx)[:, None, None] * np.ones((nx, ny, nz))
    else:
        c = np.cos(2 * np.pi * periods * np.arange(nz) / nz)[None, None, :] * np.ones((nx, ny, nz))
    w = np.zeros((NF, V))
    w[0] = amp * c.reshape(-1)
    w[1] = -amp * c.reshape(-1)
    return w, np.zeros(V)


def brush_mask(p, layers=1):
    """wMask.dat content: S1 start = 1 on the first `layers` z-layers, 0 elsewhere;
    S2, S3 the same when present."""
    nx, ny, nz = grid_of(p)
    m = np.zeros((3, nx, ny, nz))
    for s in range(3):
        if p.vol_s[s] > 0:
            m[s, :, :, :layers] = 1.0
    return m.reshape(3, -1)


# --------------------------------------------------------------- small cases


def t_bulk_small(ns=11):
    """8x1x8 periodic AB diblock melt."""
    return make_params(lx=8.0, ly=1.0, lz=8.0, dr_target=1.0, ns=ns, mask=0, g1=0.1,
                       chi={"AB": 20.0}, vol_dbk=(1.0, 0, 0, 0))


def t_film_small(ns=21):
    """12x1x10 Dirichlet film, dbk1 + homo C + brush S1."""
    return chem_droplet(lx=12.0, ly=1.0, lz=10.0, dr=1.0, ns=ns, vol_dbk=0.2, vol_s1=0.05, len_s1=0.15)


def t_3d_small(ns=11):
    """6x5x4 3D film: dbk1, homo A (short), homo C, brush S1."""
    return chem_droplet(lx=6.0, ly=5.0, lz=4.0, dr=1.0, ns=ns, vol_dbk=0.2, vol_s1=0.05,
                        vol_homo_a=0.1, len_homo_a=0.5, len_s1=0.3)


def t_all_species(mask=1, ns=11):
    """Every species the reference supports on a 6x1x6 grid: dbk1..4,
    homopolymers A..J, brushes S1..S3 (MASK 1) with a dense chi matrix."""
    chi = {}
    names = list("ABCDEFGHIJ") + ["S1", "S2", "S3"]
    k = 0
    for a in range(13):
        for b in range(a + 1, 13):
            k += 1
            chi[names[a] + names[b]] = 2.0 + (k * 7 % 11)
    return make_params(lx=6.0, ly=1.0, lz=6.0, dr_target=1.0, ns=ns, mask=mask, g1=0.02,
                       chi=chi, f_dbk=(0.5, 0.4, 0.3, 0.6), len_dbk=(1.0, 0.8, 1.0, 0.6),
                       vol_dbk=(0.1, 0.1, 0.1, 0.1),
                       vol_homo=(0.06,) * 10, len_homo=(1.0, 0.5, 1.0, 0.7, 1.0, 1.0, 0.4, 1.0, 1.0, 0.9),
                       vol_s=(0.03, 0.02, 0.01), len_s=(0.3, 0.5, 0.3))
