"""ParmParse-compatible reader for the reference's input files.

Accepts the same `key = value value ...  # comment` files and command-line
overrides (`key=value`) that amrex::ParmParse does for the keys on the hot
path, and maps them onto the library's parameters exactly as the reference
does:
  adv.num_method            NumericalMethods.H:21-44
  adv.cfl, stop_time, max_step   AmrLevelAdv.cpp:837-870, main.cpp:30-42
  prob.adiabatic, prob.epsilon   FluidSystem.cpp:10-25
  limiting.{limiter,density,mom,energy}  FluidSystem.cpp:27-47, 91-128
  bc.{x,y,z}                BCs.cpp:35-74 (default transmissive)
  geometry.{prob_lo,prob_hi,is_periodic}, amr.{n_cell,max_grid_size}
"""
import shlex

from ._lib import BC, LIMITER, METHOD

VARIABLE_INDEX = {  # FluidSystem::VariableIndex / VariableIsPrimitive
    "density": (0, False), "mom_x": (1, False), "mom_y": (2, False), "mom_z": (3, False),
    "u_x": (1, True), "u_y": (2, True), "u_z": (3, True),
}


class ParmParse:
    def __init__(self, table=None):
        self.table = dict(table or {})

    @classmethod
    def from_file(cls, path, overrides=()):
        with open(path) as fh:
            text = fh.read()
        t = parse_text(text)
        for ov in overrides:
            t.update(parse_text(ov))
        return cls(t)

    def contains(self, key):
        return key in self.table

    def get(self, key, typ=str):
        if key not in self.table:
            raise KeyError("ParmParse: required key %s not found" % key)
        return typ(self.table[key][0])

    def query(self, key, default=None, typ=str):
        if key not in self.table:
            return default
        return typ(self.table[key][0])

    def getarr(self, key, typ=str):
        if key not in self.table:
            raise KeyError("ParmParse: required key %s not found" % key)
        return [typ(v) for v in self.table[key]]

    def queryarr(self, key, default=None, typ=str):
        if key not in self.table:
            return default
        return [typ(v) for v in self.table[key]]


def parse_text(text):
    t = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line or "=" not in line:
            continue
        k, v = line.split("=", 1)
        toks = shlex.split(v)
        if toks:
            t[k.strip()] = toks
    return t


def read_inputs(path, overrides=()):
    return ParmParse.from_file(path, overrides)


def limiting_from_pp(pp, dim):
    """FluidSystem::ReadParameters limiting block -> (limiter, indices, prim)."""
    limiter = pp.query("limiting.limiter", "van_leer")
    if limiter not in LIMITER:
        raise ValueError("Invalid limiter %s" % limiter)
    idx = [dim + 1] * (dim + 2)
    prim = [False] * (dim + 2)

    def var(name):
        if name == "energy":
            return dim + 1, False
        if name == "p":
            return dim + 1, True
        if name not in VARIABLE_INDEX:
            raise ValueError("Invalid variable name " + name)
        i, pr = VARIABLE_INDEX[name]
        if i > dim:
            raise ValueError("Invalid variable name " + name)
        return i, pr

    d = pp.query("limiting.density")
    if d:
        idx[0], prim[0] = var(d)
    e = pp.query("limiting.energy")
    if e:
        idx[dim + 1], prim[dim + 1] = var(e)
    moms = pp.queryarr("limiting.mom", [])
    for i in range(min(dim, len(moms))):
        if moms[i]:
            idx[1 + i], prim[1 + i] = var(moms[i])
    return limiter, idx, prim


def config_from_inputs(pp, dim):
    """Everything the hot path needs from an input file."""
    method = pp.get("adv.num_method")
    if method not in METHOD:
        raise ValueError("Invalid numerical method \"%s\" for the explicit path" % method)
    limiter, idx, prim = limiting_from_pp(pp, dim)
    n_cell = pp.getarr("amr.n_cell", int)[:dim]
    prob_lo = pp.getarr("geometry.prob_lo", float)[:dim]
    prob_hi = pp.getarr("geometry.prob_hi", float)[:dim]
    periodic = pp.queryarr("geometry.is_periodic", [0, 0, 0], int)[:dim]
    bc_lo, bc_hi = [], []
    for d, ax in enumerate("xyz"[:dim]):
        if periodic[d]:
            bc_lo.append("periodic")
            bc_hi.append("periodic")
            continue
        b = pp.queryarr("bc." + ax, ["transmissive", "transmissive"])
        for s in b:
            if s not in BC:
                raise ValueError("Boundary type not implemented!")
        bc_lo.append(b[0])
        bc_hi.append(b[1])
    return dict(
        num_method=method, limiter=limiter, limiter_indices=idx,
        limit_on_primv=[int(x) for x in prim],
        adiabatic=pp.get("prob.adiabatic", float), epsilon=pp.query("prob.epsilon", 1.0, float),
        cfl=pp.query("adv.cfl", 0.9, float), stop_time=pp.query("stop_time", -1.0, float),
        max_step=pp.query("max_step", -1, int), n_cell=n_cell, prob_lo=prob_lo,
        prob_hi=prob_hi, bc_lo=bc_lo, bc_hi=bc_hi,
        max_grid_size=pp.query("amr.max_grid_size", 128 if dim < 3 else 32, int),
    )
