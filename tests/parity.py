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


def field_errors(got, ref, names):
    """Rows of (name, max|diff|, max|ref|, diff/max|ref|, diff/group scale, max pointwise rel)."""
    scales = {n: float(np.abs(ref[i]).max()) for i, n in enumerate(names)}
    gscale = dict(scales)
    for grp in _GROUPS:
        m = max([scales[n] for n in grp if n in scales] or [0.0])
        for n in grp:
            if n in gscale:
                gscale[n] = m
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
