"""Expected ghost cells of a single level that covers its whole domain, from
numpy's own padding routines.  The reference maps its boundary types to
AMReX BCTypes in Source/BoundaryConditions/BCs.cpp:226-262 (periodic ->
int_dir, transmissive -> foextrap, reflective -> reflect_even with
reflect_odd for the normal momentum); these are np.pad's "wrap", "edge"
and "symmetric" modes, with the momentum normal to a reflecting wall
negated.  Padding one axis after another gives the
corner and edge ghosts the value the dimension-by-dimension fill gives them
(the image of the corner under each dimension's map in turn)."""
import numpy as np

MODE = {"periodic": "wrap", "transmissive": "edge", "reflective": "symmetric"}

# (dim, bc_lo, bc_hi) cases: each type alone, and mixed per dimension and side
GHOST_BCS = [
    (2, ["transmissive"] * 2, ["transmissive"] * 2),
    (2, ["reflective"] * 2, ["reflective"] * 2),
    (2, ["periodic"] * 2, ["periodic"] * 2),
    (2, ["periodic", "reflective"], ["periodic", "transmissive"]),
    (2, ["reflective", "periodic"], ["transmissive", "periodic"]),
    (3, ["reflective", "transmissive", "periodic"], ["transmissive", "reflective", "periodic"]),
    (3, ["reflective"] * 3, ["reflective"] * 3),
]


def padded(field, ng, bc_lo, bc_hi):
    """field: (nc, [z,] [y,] x) valid cells of the whole domain; returns the
    same array with ng ghost cells on every side of every spatial axis."""
    dim = field.ndim - 1
    out = np.asarray(field, dtype=np.float64)
    for d in range(dim):
        axis = dim - d  # arrays are (n, z, y, x) ordered
        for side, bc in ((0, bc_lo[d]), (1, bc_hi[d])):
            if bc == "periodic":
                assert bc_lo[d] == bc_hi[d] == "periodic"
            pw = [(0, 0)] * out.ndim
            pw[axis] = (ng, 0) if side == 0 else (0, ng)
            if bc == "periodic":
                # wrap against the unpadded extent: pad both sides at once
                if side == 1:
                    continue
                pw[axis] = (ng, ng)
            out = np.pad(out, pw, mode=MODE[bc])
            if bc == "reflective":
                sl = [slice(None)] * out.ndim
                sl[0] = 1 + d  # the momentum normal to the wall
                sl[axis] = slice(0, ng) if side == 0 else slice(out.shape[axis] - ng, None)
                out[tuple(sl)] *= -1.0
    return out


def expected_fab(full, lay, b):
    """Box b's FAB (ghosts included) cut from padded(...)."""
    lo, hi = lay.boxes[b]
    g = lay.ng
    sl = (slice(None),) + tuple(slice(lo[d], hi[d] + 1 + 2 * g) for d in reversed(range(lay.dim)))
    return full[sl]
