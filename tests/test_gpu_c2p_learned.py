"""The learned cons2prim guess on the GPU: training gradients against the CPU
oracle, the BF16 MFMA inference against the float64 forward pass, and the
trained guess inside the Newton cons2prim against the plain guess. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import c2p_guess
from method_amd import engine as eng
from method_amd import initdata as idt
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_training_gradients_match_oracle():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(512, 8)).astype(np.float32)
    t = rng.normal(size=(512, 2)).astype(np.float32)
    m = eng.Mlp(seed=3)
    p, _ = m.params()
    g, loss = m.gradients(x, t)
    m.close()
    ref_loss, ref = omlp.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    say(f"loss gpu {loss:.6e} oracle {ref_loss:.6e}; max grad diff {np.max(np.abs(g - ref)):.3e} "
        f"of max {np.max(np.abs(ref)):.3e}")
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss
    assert np.max(np.abs(g - ref)) <= 1e-4 * np.max(np.abs(ref))


def _bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_mfma_inference_matches_forward():
    """The BF16 MFMA kernel against the float64 forward pass with the same
    BF16 rounding of weights and activations (this also pins the MFMA operand
    layout), and against the unrounded forward pass within BF16 accuracy."""
    rng = np.random.default_rng(6)
    prims = c2p_guess.synthetic_prims(3000, seed=11)
    cons = c2p_guess.cons_from_prims(prims)[:, :8]
    m = eng.Mlp(seed=4)
    p = rng.normal(scale=0.25, size=omlp.NPARAMS).astype(np.float32)
    m.set_params(p, np.array([0.5, -2.0, 1.5, 2.5], dtype=np.float32))
    y = m.predict(cons).astype(np.float64)
    m.close()
    x = omlp.features(cons)
    w = omlp.unpack(p.astype(np.float64))
    h1 = _bf16(omlp.gelu(_bf16(x) @ _bf16(w["W1"]) + w["b1"]))
    h2 = _bf16(omlp.gelu(h1 @ _bf16(w["W2"]) + w["b2"]))
    y_emul = h2 @ _bf16(w["W3"]) + w["b3"]
    y_ref, _ = omlp.forward(p.astype(np.float64), x)
    say(f"mfma vs bf16-emulated max {np.max(np.abs(y - y_emul)):.3e}; vs float64 max {np.max(np.abs(y - y_ref)):.3e}"
        f" (output scale {np.max(np.abs(y_ref)):.2f})")
    assert np.max(np.abs(y - y_emul)) <= 2e-2 * max(1.0, np.max(np.abs(y_ref)))
    assert np.median(np.abs(y - y_emul)) <= 1e-3 * max(1.0, np.max(np.abs(y_ref)))
    assert np.max(np.abs(y - y_ref)) <= 5e-2 * max(1.0, np.max(np.abs(y_ref)))


@pytest.fixture(scope="module")
def trained():
    say("gpu: training the guess network")
    p, n, loss = c2p_guess.train_guess(n_synth=1 << 18, n_kh=1 << 15, n_ot=1 << 14, steps=3000, batch=4096, lr=2e-3)
    say(f"final training loss {loss:.4e}")
    return p, n, loss


def test_training_reduces_loss(trained):
    p, n, loss = trained
    assert loss < 0.15  # normalised outputs start near unit variance


def test_learned_guess_gives_the_same_roots_with_fewer_steps_than_cold(trained):
    """On a KH state after one step: learned-guess Newton reaches the same
    primitives as the plain (previous-step) guess, to the Newton tolerance,
    and needs fewer steps than a cold start."""
    p, n, _ = trained
    g = idt.kh3d_grid(64)
    cfg = eng.Config(nx=64, ny=64, nz=64, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     weno_order=5)
    out = {}
    with eng.Engine(cfg) as e:
        e.set_prims(idt.kh_prims(g, seed=3))
        e.update_time()
        _, p_prev, _ = e.download()
        e.update_time()
        for mode in (eng.GUESS_PLAIN, eng.GUESS_LEARNED, eng.GUESS_COLD):
            e.upload(prims=p_prev)
            e.set_c2p_guess(mode, p if mode == eng.GUESS_LEARNED else None, n if mode == eng.GUESS_LEARNED else None)
            rc, st = e.c2p()
            _, prims, _ = e.download()
            out[mode] = (prims, st)
            say(f"mode {mode}: {st['iterations'] / st['cells']:.3f} Newton steps/cell, max {st['max_iterations']}, "
                f"fallbacks {st['fallbacks']}, failures {st['failures']}")
            assert rc == 0 and st["failures"] == 0
    ref = out[eng.GUESS_PLAIN][0]
    for mode in (eng.GUESS_LEARNED, eng.GUESS_COLD):
        d = np.max(np.abs(out[mode][0] - ref)) / np.abs(ref).max()
        say(f"mode {mode}: max rel diff to plain {d:.2e}")
        assert d <= 1e-10
    it = {m: out[m][1]["iterations"] / out[m][1]["cells"] for m in out}
    assert it[eng.GUESS_LEARNED] < it[eng.GUESS_COLD]
    assert out[eng.GUESS_LEARNED][1]["fallbacks"] <= 1e-3 * out[eng.GUESS_LEARNED][1]["cells"]


def _p99(hist):
    c = np.cumsum(hist)
    return int(np.searchsorted(c, 0.99 * c[-1]))


def test_config3_at_its_own_size(trained):
    """The learned-guess workload at its own size: SRMHD cons2prim of every
    cell of the 256^3 KH state bench.py times (seed 1234), the guess being the
    state one step older, with the plain (srmhd.cc:438-442), learned and cold
    starts. Every cell converges; the learned and cold roots equal the plain
    ones to 1e-10 of each field's scale; fallbacks stay within 1e-3 of the
    cells. Mean and p99 Newton steps per cell are reported from the kernel's
    step histogram."""
    p, n, _ = trained
    g = idt.kh3d_grid(256)
    cfg = eng.Config(nx=256, ny=256, nz=256, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0, zmax=1.0,
                     weno_order=5)
    out = {}
    with eng.Engine(cfg) as e:
        e.set_prims(idt.kh_prims(g, seed=1234))
        e.update_time()
        _, p_prev, _ = e.download()
        e.update_time()
        for mode, name in ((eng.GUESS_PLAIN, "plain"), (eng.GUESS_LEARNED, "learned"), (eng.GUESS_COLD, "cold")):
            e.upload(prims=p_prev)
            e.set_c2p_guess(mode, p if mode == eng.GUESS_LEARNED else None, n if mode == eng.GUESS_LEARNED else None)
            e.set_c2p_histogram(True)
            rc, st = e.c2p()
            h = e.c2p_histogram()
            e.set_c2p_histogram(False)
            _, prims, _ = e.download()
            assert h.sum() == st["cells"] == 256 ** 3
            mean = st["iterations"] / st["cells"]
            assert abs(mean - np.dot(np.arange(64), h) / h.sum()) < 1e-9
            say(f"{name}: mean {mean:.3f} Newton steps/cell, p99 {_p99(h)}, max {st['max_iterations']}, "
                f"fallbacks {st['fallbacks']}, failures {st['failures']}")
            assert rc == 0 and st["failures"] == 0 and h[63] == 0
            out[name] = (prims[:, 4:-4, 4:-4, 4:-4], st)
    ref = out["plain"][0]
    scale = np.abs(ref).reshape(8, -1).max(axis=1)
    for name in ("learned", "cold"):
        d = np.abs(out[name][0] - ref).reshape(8, -1).max(axis=1) / np.maximum(scale, 1e-300)
        say(f"{name}: max rel diff to plain by field {d.max():.2e}")
        assert d.max() <= 1e-10
    assert out["learned"][1]["fallbacks"] <= 1e-3 * out["learned"][1]["cells"]
    assert out["learned"][1]["iterations"] < out["cold"][1]["iterations"]
