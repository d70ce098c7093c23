"""Per-field parity report between two states in the reference layout."""
from __future__ import annotations

import numpy as np

SRMHD_CONS = ["D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz", "phi"]
SRMHD_PRIMS = ["rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz"]
SRMHD_AUX = ["h", "W", "e", "c", "b0", "bx", "by", "bz", "bsq", "vsq", "BS", "Bsq", "Ssq"]

SRRMHD_CONS = ["D", "Sx", "Sy", "Sz", "tau", "Bx", "By", "Bz", "Ex", "Ey", "Ez", "psi", "phi", "qch"]
SRRMHD_PRIMS = ["rho", "vx", "vy", "vz", "p", "Bx", "By", "Bz", "Ex", "Ey", "Ez"]

# Components of one vector share that vector's scale; divergence-cleaning
# scalars take the magnetic scale and the charge density (div E) the electric one.
_GROUPS = [("Sx", "Sy", "Sz"), ("vx", "vy", "vz"), ("Bx", "By", "Bz", "phi", "psi"), ("bx", "by", "bz"),
           ("Ex", "Ey", "Ez", "qch")]


def interior(a, Ng, dims):
    sl = [slice(None), slice(Ng, -Ng), slice(Ng, -Ng) if dims >= 2 else slice(None),
          slice(Ng, -Ng) if dims == 3 else slice(None)]
    return a[tuple(sl)]


def group_scales(ref, names):
    """(field max |ref|, group scale) per field name: a field's group scale is
    the largest max |ref| over the vector group it belongs to."""
    scales = {n: float(np.abs(ref[i]).max()) for i, n in enumerate(names)}
    gscale = dict(scales)
    for grp in _GROUPS:
        m = max([scales[n] for n in grp if n in scales] or [0.0])
        for n in grp:
            if n in gscale:
                gscale[n] = m
    return scales, gscale


def field_errors(got, ref, names):
    """Rows of (name, max|diff|, max|ref|, diff/max|ref|, diff/group scale, max pointwise rel)."""
    scales, gscale = group_scales(ref, names)
    rows = []
    for i, n in enumerate(names):
        d = np.abs(got[i] - ref[i])
        dm = float(d.max())
        fm = scales[n]
        rel_f = dm / fm if fm > 0 else (0.0 if dm == 0 else np.inf)
        rel_g = dm / gscale[n] if gscale[n] > 0 else (0.0 if dm == 0 else np.inf)
        with np.errstate(divide="ignore", invalid="ignore"):
            pw = np.where(np.abs(ref[i]) > 0, d / np.abs(ref[i]), np.where(d == 0, 0.0, np.inf))
        rows.append((n, dm, fm, rel_f, rel_g, float(pw.max())))
    return rows


def pointwise_floor(got, ref, names, floor_frac):
    """Per field, max |got - ref| / max(|ref|, floor_frac * group scale): the
    pointwise relative error, with values below the floor measured against it
    (near a zero crossing a pure pointwise ratio says nothing about parity)."""
    _, gscale = group_scales(ref, names)
    out = []
    for i, n in enumerate(names):
        den = np.maximum(np.abs(ref[i]), max(floor_frac * gscale[n], 1e-300))
        out.append(float((np.abs(got[i] - ref[i]) / den).max()))
    return out


def report(title, rows):
    lines = [title, f"{'field':>6} {'max|diff|':>11} {'max|ref|':>11} {'rel(field)':>11} {'rel(group)':>11} "
                    f"{'pointwise':>11}"]
    for n, dm, fm, rf, rg, pw in rows:
        lines.append(f"{n:>6} {dm:11.3e} {fm:11.3e} {rf:11.3e} {rg:11.3e} {pw:11.3e}")
    return "\n".join(lines)


# A 16 x 16 KH state whose stored primitives at three interior cells are
# replaced by a guess of |v| = 1, (vx, vy) = (1, 0): the starting root guess
# Z = rho h / (1 - v^2) (srmhd.cc:438-442) is infinite, the Newton solve fails,
# and those cells take the neighbour-mean retry of SRMHD::getPrimitiveVars
# (srmhd.cc:465-507). Indices are into the padded
# (Ng = 4) arrays; no two bad cells are face neighbours and none is the last
# interior cell.
C2P_RETRY_GRID = dict(nx=16, ny=16, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
C2P_RETRY_CELLS = ((6, 7), (9, 12), (13, 5))


def c2p_retry_case(oracle):
    """(cons, true prims, true aux, guess prims, guess aux) for the retry test,
    built with the oracle's Simulation::set."""
    from method_amd import initdata as idt

    p0 = idt.kh_prims(idt.kh2d_grid(16), seed=5, shear=0.7, bz=0.5)
    cons, prims, aux = oracle.set(p0)
    gp, ga = prims.copy(), aux.copy()
    for i, j in C2P_RETRY_CELLS:
        gp[1, i, j, 0] = 1.0
        gp[2, i, j, 0] = 0.0
    return cons, prims, aux, gp, ga


# The 3-D variant: KH 12^3 with four bad guesses, two of them face neighbours
# along z (so each is the other's failed neighbour, which the retry skips)
# and one next to the z edge of the interior.
C2P_RETRY_GRID_3D = dict(nx=12, ny=12, nz=12, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0)
C2P_RETRY_CELLS_3D = ((6, 7, 7), (6, 7, 8), (11, 9, 5), (9, 13, 15))


def c2p_retry_case_3d(oracle):
    from method_amd import initdata as idt

    p0 = idt.kh_prims(idt.kh3d_grid(12), seed=6, shear=0.7, bz=0.5)
    cons, prims, aux = oracle.set(p0)
    gp, ga = prims.copy(), aux.copy()
    for i, j, k in C2P_RETRY_CELLS_3D:
        gp[1, i, j, k] = 0.0
        gp[3, i, j, k] = 1.0
        gp[2, i, j, k] = 0.0
    return cons, prims, aux, gp, ga


def c2p_step_retry_case(oracle):
    """The 16 x 16 KH state of c2p_retry_case, consistent, except that at the
    three cells the stored density is the subnormal 1e-310 and the stored h is
    infinite: the flux (which reads v and p) is untouched, but the starting
    root Z = rho h / (1 - v^2) (srmhd.cc:438-442, h from p / rho or from the
    aux) is infinite, so the cons2prim of both RK2 stages fails there (the
    stage-two guess is the step's input prims, RK2.cc:81-91) and retries."""
    from method_amd import initdata as idt

    p0 = idt.kh_prims(idt.kh2d_grid(16), seed=5, shear=0.7, bz=0.5)
    cons, prims, aux = oracle.set(p0)
    gp, ga = prims.copy(), aux.copy()
    for i, j in C2P_RETRY_CELLS:
        gp[0, i, j, 0] = 1e-310
        ga[0, i, j, 0] = float("inf")
    return cons, prims, aux, gp, ga
