"""A reader of AMReX plotfiles (HyperCLaw-V1.1 Header, VisMF Cell_H, FAB
data files), written from the format's description, to check what
bx_write_plotfile produces."""
import os
import re

import numpy as np

IEEE_LE = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"


def _ivec(s):
    return tuple(int(x) for x in s.strip("()").split(","))


def _box(s):
    m = re.fullmatch(r"\((\([-\d,]+\)) (\([-\d,]+\)) (\([-\d,]+\))\)", s.strip())
    assert m, s
    return _ivec(m.group(1)), _ivec(m.group(2)), _ivec(m.group(3))


def read_plotfile(path):
    lines = open(os.path.join(path, "Header")).read().split("\n")
    it = iter(lines)
    out = {"version": next(it)}
    nvar = int(next(it))
    out["names"] = [next(it) for _ in range(nvar)]
    dim = int(next(it))
    out["dim"] = dim
    out["time"] = float(next(it))
    finest = int(next(it))
    out["prob_lo"] = [float(x) for x in next(it).split()]
    out["prob_hi"] = [float(x) for x in next(it).split()]
    out["ref_ratio"] = [int(x) for x in next(it).split()]
    doms = re.findall(r"\(\([-\d,]+\) \([-\d,]+\) \([-\d,]+\)\)", next(it))
    out["domains"] = [_box(d) for d in doms]
    out["steps"] = [int(x) for x in next(it).split()]
    out["dx"] = [[float(x) for x in next(it).split()] for _ in range(finest + 1)]
    out["coord"] = int(next(it))
    out["bwidth"] = int(next(it))
    levels = []
    for l in range(finest + 1):
        lev, nbox, t = next(it).split()
        assert int(lev) == l
        lsteps = int(next(it))
        phys = []
        for _ in range(int(nbox)):
            phys.append([tuple(float(x) for x in next(it).split()) for _ in range(dim)])
        mf = next(it)
        levels.append(dict(nbox=int(nbox), time=float(t), steps=lsteps, phys=phys,
                           fabs=_read_multifab(os.path.join(path, mf), dim, nvar)))
    out["levels"] = levels
    return out


def _read_multifab(prefix, dim, nvar):
    lines = open(prefix + "_H").read().split("\n")
    it = iter(lines)
    assert next(it) == "1" and next(it) == "1"
    assert int(next(it)) == nvar
    assert int(next(it)) == 0
    head = next(it)
    nbox = int(head.strip("(").split()[0])
    boxes = [_box(next(it)) for _ in range(nbox)]
    assert next(it) == ")"
    assert int(next(it)) == nbox
    fod = []
    for _ in range(nbox):
        tag, fname, off = next(it).split()
        assert tag == "FabOnDisk:"
        fod.append((fname, int(off)))
    assert next(it) == ""
    n, c = next(it).split(",")
    assert int(n) == nbox and int(c) == nvar
    mins = [[float(x) for x in next(it).rstrip(",").split(",")] for _ in range(nbox)]
    assert next(it) == ""
    next(it)
    maxs = [[float(x) for x in next(it).rstrip(",").split(",")] for _ in range(nbox)]
    fabs = []
    base = os.path.dirname(prefix)
    for (lo, hi, _t), (fname, off), mn, mx in zip(boxes, fod, mins, maxs):
        with open(os.path.join(base, fname), "rb") as f:
            f.seek(off)
            line = f.readline().decode()
            assert line.startswith("FAB " + IEEE_LE), line
            rest = line[len("FAB " + IEEE_LE):].strip()
            bstr, ncomp = rest.rsplit(" ", 1)
            assert _box(bstr)[:2] == (lo, hi) and int(ncomp) == nvar
            shape = tuple(hi[d] - lo[d] + 1 for d in reversed(range(dim)))
            npts = int(np.prod(shape))
            data = np.frombuffer(f.read(8 * npts * nvar), dtype="<f8").reshape((nvar,) + shape)
        fabs.append(dict(lo=lo, hi=hi, data=data, min=mn, max=mx))
    return fabs
