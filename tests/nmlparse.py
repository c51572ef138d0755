"""Minimal Fortran-namelist reader used by the tests.

It is deliberately independent of the product's C++ reader (the two are
checked against each other).  It handles what SNaC's inputs use
(param.f90:88-106 and the example ``dns.nml``/``blocks.nml`` files):
``&group ... /`` blocks, scalar and sliced array assignments such as
``block_cbcvel(0:1,1:3,2,1) = 'F', 'F', ...`` (filled in Fortran order),
numbers, quoted strings, logicals (T/F/.true./.false.), ``n*value`` repeats
and ``!`` comments.
"""
import re
import numpy as np


def _tokens(text):
    out = []
    i, n = 0, len(text)
    while i < n:
        ch = text[i]
        if ch in " \t\r\n,":
            i += 1
        elif ch == "!":
            while i < n and text[i] != "\n":
                i += 1
        elif ch in "'\"":
            j = text.index(ch, i + 1)
            out.append(("str", text[i + 1:j]))
            i = j + 1
        elif ch in "&/=()":
            out.append(("sym", ch))
            i += 1
        else:
            j = i
            while j < n and text[j] not in " \t\r\n,=()!&/'\"":
                j += 1
            # a '/' inside a number is not possible; keep ':' and '*' inside words
            out.append(("word", text[i:j]))
            i = j
    return out


def _value(tok):
    kind, s = tok
    if kind == "str":
        return s
    low = s.lower()
    if low in ("t", ".t.", ".true.", "true"):
        return True
    if low in ("f", ".f.", ".false.", "false"):
        return False
    try:
        if re.fullmatch(r"[+-]?\d+", s):
            return int(s)
        return float(low.replace("d", "e"))
    except ValueError:
        return s


def _expand_index(spec, nvals=1):
    """'0:1,1:3,2,1' -> list of index tuples in Fortran order; an open bound
    (':' as in alphai(:)) starts at 1 and extends over the values given."""
    ranges = []
    for part in spec.split(","):
        part = part.strip()
        if ":" in part:
            a, b = part.split(":")
            a = int(a) if a.strip() else 1
            b = int(b) if b.strip() else a + nvals - 1
            ranges.append(range(a, b + 1))
        else:
            ranges.append(range(int(part), int(part) + 1))
    combos = []  # first index fastest

    def rec(level, cur):
        if level < 0:
            combos.append(tuple(cur))
            return
        for v in ranges[level]:
            cur[level] = v
            rec(level - 1, cur)
    rec(len(ranges) - 1, [0] * len(ranges))
    return combos


def parse(path):
    text = open(path).read()
    toks = _tokens(text)
    groups = {}
    i = 0
    cur = None
    while i < len(toks):
        kind, s = toks[i]
        if kind == "sym" and s == "&":
            cur = groups.setdefault(toks[i + 1][1].lower(), {})
            i += 2
            continue
        if kind == "sym" and s == "/":
            cur = None
            i += 1
            continue
        if cur is None:
            i += 1
            continue
        name = s.lower()
        i += 1
        index = None
        if i < len(toks) and toks[i] == ("sym", "("):
            j = i + 1
            parts = []
            while toks[j] != ("sym", ")"):
                parts.append(toks[j][1])
                j += 1
            index = ",".join(parts)
            i = j + 1
        assert toks[i] == ("sym", "="), f"expected '=' after {name} in {path}"
        i += 1
        vals = []
        while i < len(toks):
            k2, s2 = toks[i]
            if k2 == "sym":
                break
            # next assignment starts with a word followed by '=' or '('
            if k2 == "word" and i + 1 < len(toks) and toks[i + 1][0] == "sym" and toks[i + 1][1] in "=(":
                if not re.fullmatch(r"[+-]?[\d.eEdD+-]+", s2):
                    break
            if k2 == "word" and "*" in s2 and re.fullmatch(r"\d+\*.+", s2):
                cnt, v = s2.split("*", 1)
                vals.extend([_value(("word", v))] * int(cnt))
            else:
                vals.append(_value(toks[i]))
            i += 1
        entry = cur.setdefault(name, {})
        if index is None:
            if len(vals) == 1:
                entry[()] = vals[0]
            else:
                for q, v in enumerate(vals):
                    entry[(q + 1,)] = v
        else:
            combos = _expand_index(index, len(vals))
            for q, v in enumerate(vals[:len(combos)]):
                entry[combos[q]] = v
    return groups


def get(group, name, default=None, index=()):
    e = group.get(name.lower())
    if e is None:
        return default
    if index in e:
        return e[index]
    if index == () and (1,) in e:
        return e[(1,)]
    return default


def load_case(dns_path, blocks_path, fft_axis):
    """Inputs of param.f90:68-277 as flat arrays (block-major, Fortran order
    inside a block) plus the dns scalars needed by the hot path."""
    dns = parse(dns_path)
    blk = parse(blocks_path)["blocks"]
    d = dns["dns"]
    nb = int(get(blk, "nblocks"))
    ng = np.ones((nb, 3), dtype=np.int32)
    lmin = np.zeros((nb, 3))
    lmax = np.ones((nb, 3))
    gt = np.zeros((nb, 3), dtype=np.int32)
    gr = np.zeros((nb, 3))
    cbcvel = np.full((nb, 3, 3, 2), "D")
    bcvel = np.zeros((nb, 3, 3, 2))
    cbcpre = np.full((nb, 3, 2), "N")
    bcpre = np.zeros((nb, 3, 2))
    inflow = np.zeros((nb, 3, 2), dtype=np.int32)
    for b in range(nb):
        for q in range(3):
            ng[b, q] = get(blk, "block_ng", 1, (q + 1, b + 1))
            lmin[b, q] = get(blk, "block_lmin", 0.0, (q + 1, b + 1))
            lmax[b, q] = get(blk, "block_lmax", 1.0, (q + 1, b + 1))
            gt[b, q] = get(blk, "block_gt", 0, (q + 1, b + 1))
            gr[b, q] = get(blk, "block_gr", 0.0, (q + 1, b + 1))
            for s in range(2):
                cbcpre[b, q, s] = get(blk, "block_cbcpre", "N", (s, q + 1, b + 1))
                bcpre[b, q, s] = get(blk, "block_bcpre", 0.0, (s, q + 1, b + 1))
                inflow[b, q, s] = get(blk, "block_inflow_type", 0, (s, q + 1, b + 1))
                for c in range(3):
                    cbcvel[b, c, q, s] = get(blk, "block_cbcvel", "D", (s, q + 1, c + 1, b + 1))
                    bcvel[b, c, q, s] = get(blk, "block_bcvel", 0.0, (s, q + 1, c + 1, b + 1))
    bforce = np.array([get(d, "bforce", 0.0, (q + 1,)) for q in range(3)], dtype=float)
    gacc = np.array([get(d, "gacc", 0.0, (q + 1,)) for q in range(3)], dtype=float)
    hyp = dns.get("hypre", {})
    # scalars: &scalar (param.f90:96-100,148-163), block_cbcscal/bcscal
    nscal = int(get(d, "nscal", 0))
    sc = dns.get("scalar", {})
    if nscal > 0 and "scalar" not in dns:
        raise ValueError("Error reading scalar namelist")
    alphai = np.array([float(get(sc, "alphai", 1.7976931348623157e308, (i + 1,))) for i in range(nscal)])
    ssource = np.array([float(get(sc, "ssource", 0.0, (i + 1,))) for i in range(nscal)])
    is_sforced = np.array([int(bool(get(sc, "is_sforced", False, (i + 1,)))) for i in range(nscal)], np.int32)
    scalf = np.array([float(get(sc, "scalf", 0.0, (i + 1,))) for i in range(nscal)])
    cbcscal = np.full((nb, max(nscal, 1), 3, 2), "N")
    bcscal = np.zeros((nb, max(nscal, 1), 3, 2))
    for b in range(nb):
        for i in range(nscal):
            for q in range(3):
                for s in range(2):
                    cbcscal[b, i, q, s] = get(blk, "block_cbcscal", "N", (s, q + 1, i + 1, b + 1))
                    bcscal[b, i, q, s] = get(blk, "block_bcscal", 0.0, (s, q + 1, i + 1, b + 1))
    return dict(
        nblocks=nb, ng=ng, lmin=lmin, lmax=lmax, gt=gt, gr=gr, cbcvel=cbcvel, bcvel=bcvel,
        cbcpre=cbcpre, bcpre=bcpre, inflow_type=inflow, fft_axis=fft_axis,
        visci=float(get(d, "visci", 1.0)), bforce=bforce,
        cfl=float(get(d, "cfl")), dtmax=float(get(d, "dtmax")), dt_f=float(get(d, "dt_f", -1.0)),
        nstep=int(get(d, "nstep")), icheck=int(get(d, "icheck")),
        hypre_tol=float(get(hyp, "hypre_tol", 1e-4)), hypre_maxiter=int(get(hyp, "hypre_maxiter", 50)),
        nscal=nscal, gacc=gacc, beta=float(get(sc, "beta", 0.0)), alphai=alphai, ssource=ssource,
        is_sforced=is_sforced, scalf=scalf, cbcscal=cbcscal[:, :nscal], bcscal=bcscal[:, :nscal],
    )
