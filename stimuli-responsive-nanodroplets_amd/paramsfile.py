"""Write a parameters.h (the reference's compile-time run description,
parameters.h:1-190) for an ScftParams, so the drop-in driver scft_run can
be built for any workload, and the matching input files of a run directory."""
import os

from abi import grid_of

HOMO = "ABCDEFGHIJ"


def _d(x):
    """A C double literal (always with a decimal point, as parameters.h has)."""
    t = "%.17g" % x
    if not any(ch in t for ch in ".eEn"):
        t += ".0"
    return t


def render(p, iter_max, d_wrt, d_wrt_energies):
    L = []
    a = L.append
    a("#define LX %s" % _d(p.lx))
    a("#define LY %s" % _d(p.ly))
    a("#define LZ %s" % _d(p.lz))
    a("#define ITER %d" % iter_max)
    a("#define D_WRT %d" % d_wrt)
    a("#define RELAXATION %d" % p.relaxation)
    a("#define D_WRT_ENERGIES %d" % d_wrt_energies)
    a("#define DT_NOISE %s" % _d(p.dt_noise))
    a("#define MASK %d" % p.mask)
    a("#define ENABLE_SEED %d" % p.enable_seed)
    a("#define SEED_STEPS %d" % p.seed_steps)
    a("#define SEED_AREA_DELTA %d" % p.seed_area_delta)
    a("#define NS %d" % p.ns)
    a("#define DR_TARGET %s" % _d(p.dr_target))
    a("#define DS (1.0 / (NS - 1))")
    a("#define G1 %s" % _d(p.g1))
    for i in range(10):
        for j in range(i + 1, 10):
            a("#define CHI_N_%s%s %s" % (HOMO[i], HOMO[j], _d(p.chi[i][j])))
    for b in range(4):
        a("#define F_INPUT_DBK_%d %s" % (b + 1, _d(p.f_dbk[b])))
        a("#define LEN_DBK_%d %s" % (b + 1, _d(p.len_dbk[b])))
    for b in range(4):
        a("#define VOL_DBK_%d %s" % (b + 1, _d(p.vol_dbk[b])))
    for h in range(10):
        a("#define VOL_HOMOPOLYMER_%s %s" % (HOMO[h], _d(p.vol_homo[h])))
    for h in range(10):
        a("#define LEN_HOMO_%s %s" % (HOMO[h], _d(p.len_homo[h])))
    for s in range(3):
        a("#define VOL_S%d %s" % (s + 1, _d(p.vol_s[s])))
        a("#define LEN_S%d %s" % (s + 1, _d(p.len_s[s])))
    a("#define CHI_N_S1_S2 %s" % _d(p.chi[10][11]))
    a("#define CHI_N_S1_S3 %s" % _d(p.chi[10][12]))
    a("#define CHI_N_S2_S3 %s" % _d(p.chi[11][12]))
    for s in range(3):
        for h in range(10):
            a("#define CHI_N_S%d_%s %s" % (s + 1, HOMO[h], _d(p.chi[10 + s][h])))
    a("#define DX  (LX/NX)")
    a("#define DY  (LY/NY)")
    a("#define DZ  (LZ/NZ)")
    for name, fn in (("COUNTER", "counter.dat"), ("DENSITY", "densities_dbk_"), ("DEVIATIONS", "deviations.dat"),
                     ("ENERGY", "energies.dat"), ("RESTART_FIELDS", "restart_fields.dat"),
                     ("RESTART_RANDOM", "restart_random.dat"), ("RUN_PARAMS", "run.txt"), ("W_MASK", "wMask.dat"),
                     ("FREE_ENERGY", "free_energy.dat")):
        a('#define %s "%s"' % (name, fn))
    return "\n".join(L) + "\n"


def write(p, path, iter_max, d_wrt=1000, d_wrt_energies=1):
    with open(path, "w") as f:
        f.write(render(p, iter_max, d_wrt, d_wrt_energies))


def write_inputs(directory, p, w, pf, mask=None, counter=1, seed=(0, 0, 0)):
    """counter.dat, restart_fields.dat (full precision; fscanf %lf reads it,
    SCFT.c:861-862), restart_random.dat and wMask.dat for a run directory."""
    nx, ny, nz = grid_of(p)
    with open(os.path.join(directory, "counter.dat"), "w") as f:
        f.write("%d" % counter)
    with open(os.path.join(directory, "restart_fields.dat"), "w") as f:
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    idx = k + nz * (j + ny * i)
                    f.write(" ".join("%.17g" % w[c][idx] for c in range(13)) + " %.17g \n" % pf[idx])
    with open(os.path.join(directory, "restart_random.dat"), "w") as f:
        f.write("%d %d %d" % tuple(seed))
    if mask is not None:
        with open(os.path.join(directory, "wMask.dat"), "w") as f:
            for k in range(nz):
                for j in range(ny):
                    for i in range(nx):
                        idx = k + nz * (j + ny * i)
                        f.write("%.17g %.17g %.17g\n" % (mask[0][idx], mask[1][idx], mask[2][idx]))
