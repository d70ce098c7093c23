"""The allocation-free cell-list cons2prim, mgpu_c2p_cells_noalloc, through
the C ABI: on device arrays it must give exactly what mgpu_c2p_cells gives on
the same host arrays (the same kernels; only the counter scratch differs), on
two calls in a row with the second smaller, for both models. Needs an MI355X."""
import ctypes
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from method_amd import initdata as idt

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cells(a):
    """[var][i][j][k] interior -> [cell][var]."""
    return np.ascontiguousarray(a.reshape(a.shape[0], -1).T)


def _srmhd_cells(golden, n, seed):
    """n cells of the reference cons2prim vectors, each guess moved by a
    seeded few per cent so the Newton solve takes steps."""
    z = golden("srmhd_c2p_vectors")
    cons, gp, ga = (_cells(z[k][:, 4:-4, 0, 0]) for k in ("cons", "guess_prims", "guess_aux"))
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, cons.shape[0], size=n)
    c, p, a = cons[idx].copy(), gp[idx].copy(), ga[idx].copy()
    p[:, 1:4] *= 1.0 + 0.03 * rng.uniform(-1.0, 1.0, size=(n, 1))
    p[:, 4] *= 1.0 + 0.05 * rng.uniform(-1.0, 1.0, size=n)
    return c, p, a


def _srrmhd_cells(n, seed):
    """n cells of the resistive Orszag-Tang state with the Z guess moved by up to 2 %."""
    grid = dict(nx=64, ny=64, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    with eng.Engine(eng.Config(model=eng.SRRMHD, integrator=eng.SSP2, sigma=1e3, cp=0.1, cfl=0.5,
                               bc=(eng.PERIODIC,) * 3, **grid)) as e:
        e.set_prims(idt.ot_prims_resistive(idt.ot2d_grid(64)))
        c, p, a = e.download()
    cons, prims, aux = (_cells(x[:, 4:-4, 4:-4, :]) for x in (c, p, a))
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, cons.shape[0], size=n)
    cc, pp, aa = cons[idx].copy(), prims[idx].copy(), aux[idx].copy()
    aa[:, 10] *= 1.0 + 0.02 * rng.uniform(-1.0, 1.0, size=n)
    return cc, pp, aa


class _Hip:
    """Device buffers for the test through the HIP runtime libmethod_gpu.so
    itself links (the entry point under test allocates nothing; its caller
    does, beforehand)."""

    H2D, D2H = 1, 2

    def __init__(self):
        eng.load_library()
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def to_device(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        self._ok(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.ptrs.append(p)
        self._ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy H2D")
        return p.value

    def to_host(self, ptr, like):
        out = np.empty_like(like, dtype=np.float64)
        self._ok(self.L.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.ptrs:
            self._ok(self.L.hipFree(p), "hipFree")
        self.ptrs = []


def _run_noalloc(e, hip, c, p, a, stats=True):
    dc, dp, da = (hip.to_device(x) for x in (c, p, a))
    try:
        rc, st = e.c2p_cells_noalloc(c.shape[0], dc, dp, da, stats=stats)
        if not stats:
            e.sync()
        return rc, hip.to_host(dp, p), hip.to_host(da, a), st
    finally:
        hip.free()


def _check_pair(e, hip, cells, label):
    c, p, a = cells
    rc, pr, ar, st = e.c2p_cells(c, p, a)
    rc2, pn, an, st2 = _run_noalloc(e, hip, c, p, a)
    say(f"{label}: n={c.shape[0]} rc={rc}/{rc2} steps {st['iterations']}/{st2['iterations']} "
        f"failures {st['failures']}/{st2['failures']}")
    assert rc == rc2
    for k in ("cells", "iterations", "failures", "max_iterations"):
        assert st[k] == st2[k], k
    assert st2["cells"] == c.shape[0] and st2["iterations"] > 0
    # the same kernels on the same inputs: identical bits
    assert np.array_equal(pn, pr) and np.array_equal(an, ar)


@pytest.mark.parametrize("model", ["srmhd", "srrmhd"])
def test_noalloc_c2p_cells_equals_c2p_cells(golden, model):
    if model == "srmhd":
        cfg = eng.Config(nx=8, ny=8, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
        big, small = _srmhd_cells(golden, 300_001, 11), _srmhd_cells(golden, 4_097, 12)
    else:
        cfg = eng.Config(model=eng.SRRMHD, integrator=eng.SSP2, sigma=1e3, cp=0.1, cfl=0.5, nx=8, ny=8, nz=0,
                         xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, bc=(eng.PERIODIC,) * 3)
        big, small = _srrmhd_cells(200_003, 21), _srrmhd_cells(3_001, 22)
    with eng.Engine(cfg) as e:
        hip = _Hip()
        _check_pair(e, hip, big, f"{model} first call")
        # the second, smaller call reuses the scratch the first left behind
        _check_pair(e, hip, small, f"{model} second call")


def test_noalloc_c2p_cells_without_stats_and_limits(golden):
    """Without stats the call only launches; the results after mgpu_sync are
    the same. Sizes beyond the scratch's maximum and negative sizes are
    refused before anything runs; an empty call is a no-op."""
    c, p, a = _srmhd_cells(golden, 70_000, 31)
    with eng.Engine(eng.Config(nx=8, ny=8, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)) as e:
        _, pr, ar, _ = e.c2p_cells(c, p, a)
        rc, pn, an, st = _run_noalloc(e, _Hip(), c, p, a, stats=False)
        assert rc == 0 and st is None
        assert np.array_equal(pn, pr) and np.array_equal(an, ar)
        with pytest.raises(eng.EngineError, match="exceed the maximum"):
            e.c2p_cells_noalloc((1 << 22) + 1, 0, 0, 0)
        with pytest.raises(eng.EngineError, match="negative"):
            e.c2p_cells_noalloc(-1, 0, 0, 0)
        rc, st = e.c2p_cells_noalloc(0, 0, 0, 0)
        assert rc == 0 and st["cells"] == 0
