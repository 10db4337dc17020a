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


# ------------------------------------------------------------- reading

class ParamsError(ValueError):
    pass


def _strip_comments(text):
    out, i, n = [], 0, len(text)
    while i < n:
        if text.startswith("/*", i):
            j = text.find("*/", i + 2)
            if j < 0:
                raise ParamsError("unterminated /* comment")
            i = j + 2
            out.append(" ")
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text[i] == '"':
            j = text.find('"', i + 1)
            if j < 0:
                raise ParamsError("unterminated string")
            out.append(text[i:j + 1])
            i = j + 1
        else:
            out.append(text[i])
            i += 1
    return "".join(out)


def macros(text):
    """The object-like #define NAME value lines of a header (comments
    removed, line continuations joined), as NAME -> value text."""
    text = _strip_comments(text.replace("\\\n", " "))
    out = {}
    for line in text.splitlines():
        s = line.strip()
        if not s.startswith("#"):
            continue
        parts = s[1:].strip().split(None, 2)
        if not parts or parts[0] != "define" or len(parts) < 2:
            continue
        name = parts[1]
        if "(" in name:          # function-like macro: not a run parameter
            continue
        out[name] = parts[2].strip() if len(parts) > 2 else ""
    return out


def _evaluate(name, defs, seen=()):
    """Value of macro `name` as C would compute it: numbers, other macros,
    + - * / and parentheses (integer division truncates, as in C)."""
    import ast
    if name not in defs:
        raise ParamsError("%s is not defined" % name)
    if name in seen:
        raise ParamsError("%s is defined in terms of itself" % name)
    src = defs[name]
    try:
        tree = ast.parse(src.replace("\t", " "), mode="eval")
    except SyntaxError:
        raise ParamsError("%s = %r is not an arithmetic expression" % (name, src))

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) and not isinstance(node.value, bool):
            return node.value
        if isinstance(node, ast.Name):
            return _evaluate(node.id, defs, seen + (name,))
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.UAdd, ast.USub)):
            v = ev(node.operand)
            return v if isinstance(node.op, ast.UAdd) else -v
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            a, b = ev(node.left), ev(node.right)
            if isinstance(node.op, ast.Add):
                return a + b
            if isinstance(node.op, ast.Sub):
                return a - b
            if isinstance(node.op, ast.Mult):
                return a * b
            if isinstance(a, int) and isinstance(b, int):
                if b == 0:
                    raise ParamsError("%s divides by zero" % name)
                q = abs(a) // abs(b)
                return q if (a >= 0) == (b >= 0) else -q
            return a / b
        raise ParamsError("%s = %r uses something other than numbers, macros and + - * /" % (name, src))

    return ev(tree)


def parse(text):
    """A parameters.h (the reference's run description, parameters.h:1-190)
    as (scft_params, run), run = {"iter": ITER, "d_wrt": D_WRT,
    "d_wrt_energies": D_WRT_ENERGIES}: the same values the drop-in driver's
    params_from_macros takes from the compiled header, so one server
    process accepts any run description without a rebuild.  DS must be
    1/(NS-1) (the library derives it from NS, as parameters.h:20 defines
    it); a missing or non-arithmetic parameter raises ParamsError."""
    from abi import ScftParams, NF
    defs = macros(text)

    def num(k):
        return _evaluate(k, defs)

    p = ScftParams()
    p.lx, p.ly, p.lz, p.dr_target = float(num("LX")), float(num("LY")), float(num("LZ")), float(num("DR_TARGET"))
    p.ns, p.mask = int(num("NS")), int(num("MASK"))
    p.g1, p.dt_noise = float(num("G1")), float(num("DT_NOISE"))
    p.relaxation = int(num("RELAXATION"))
    p.enable_seed, p.seed_steps, p.seed_area_delta = int(num("ENABLE_SEED")), int(num("SEED_STEPS")), \
        int(num("SEED_AREA_DELTA"))
    if "DS" in defs and abs(float(num("DS")) - 1.0 / (p.ns - 1)) > 1e-15:
        raise ParamsError("DS = %r: the library takes DS = 1/(NS-1)" % defs["DS"])
    for i in range(NF):
        for j in range(NF):
            p.chi[i][j] = 0.0
    names = list(HOMO)
    for i in range(10):
        for j in range(i + 1, 10):
            v = float(num("CHI_N_%s%s" % (names[i], names[j])))
            p.chi[i][j] = p.chi[j][i] = v
    for s in range(3):
        for h in range(10):
            v = float(num("CHI_N_S%d_%s" % (s + 1, names[h])))
            p.chi[10 + s][h] = p.chi[h][10 + s] = v
    for a, b in ((1, 2), (1, 3), (2, 3)):
        v = float(num("CHI_N_S%d_S%d" % (a, b)))
        p.chi[9 + a][9 + b] = p.chi[9 + b][9 + a] = v
    for b in range(4):
        p.f_dbk[b] = float(num("F_INPUT_DBK_%d" % (b + 1)))
        p.len_dbk[b] = float(num("LEN_DBK_%d" % (b + 1)))
        p.vol_dbk[b] = float(num("VOL_DBK_%d" % (b + 1)))
    for h in range(10):
        p.vol_homo[h] = float(num("VOL_HOMOPOLYMER_%s" % names[h]))
        p.len_homo[h] = float(num("LEN_HOMO_%s" % names[h]))
    for s in range(3):
        p.vol_s[s] = float(num("VOL_S%d" % (s + 1)))
        p.len_s[s] = float(num("LEN_S%d" % (s + 1)))
    run = {"iter": int(num("ITER")), "d_wrt": int(num("D_WRT")), "d_wrt_energies": int(num("D_WRT_ENERGIES"))}
    return p, run


def load(path):
    with open(path) as f:
        return parse(f.read())


def read_inputs(directory, p):
    """The fields (restart_fields.dat) and, for MASK 1, the brush masks
    (wMask.dat) of a run directory, as the reference reads them
    (SCFT.c:756-781, 844-866; voxels in k -> j -> i order): (w [13][V],
    p [V], mask [3][V] or None).  A short file raises ParamsError."""
    import numpy as np
    nx, ny, nz = grid_of(p)
    V = nx * ny * nz
    order = [k + nz * (j + ny * i) for k in range(nz) for j in range(ny) for i in range(nx)]

    def table(name, cols):
        path = os.path.join(directory, name)
        vals = np.array(open(path).read().split(), dtype=np.float64)
        if vals.size < cols * V:
            raise ParamsError("%s holds %d values, %d x %d expected" % (path, vals.size, cols, V))
        t = vals[:cols * V].reshape(V, cols)
        out = np.zeros((cols, V))
        out[:, order] = t.T
        return out

    f = table("restart_fields.dat", 14)
    mask = table("wMask.dat", 3) if p.mask else None
    return f[:13].copy(), f[13].copy(), mask
