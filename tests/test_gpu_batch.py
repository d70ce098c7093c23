"""Batched ensemble of 2-D ideal SRMHD simulations through the C ABI: every
member must equal the same simulation run alone on the single-instance engine
bit for bit (the batch kernel shares its arithmetic), whatever the tile shape,
schedule or workgroup count; graph replays must equal direct launches; the
autotuner must leave the state as it found it; and a member checked against
the oracle must agree to 1e-9. Needs an MI355X."""
import sys

import numpy as np
import pytest

from method_amd import engine as eng
from method_amd import initdata as idt
from oracle import oracle as orc
from tests import parity

pytestmark = pytest.mark.gpu

# (nx, ny, shear, bz, seed): uneven sizes, none a multiple of every tile shape
SMALL = [(32, 48, 0.5, 0.1, 11), (64, 40, 0.3, 0.4, 12), (24, 72, 0.7, 0.05, 13), (96, 64, 0.25, 0.3, 14),
         (40, 36, 0.6, 0.2, 15)]
BC = (eng.PERIODIC, eng.OUTFLOW)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def grid_of(nx, ny):
    return idt.Grid(nx, ny, 0, -0.5, 0.5, -1.0, 1.0)


def prims_of(nx, ny, shear, bz, seed):
    return idt.kh_prims(grid_of(nx, ny), seed=seed, shear=shear, bz=bz)


def make_batch(members, weno=5, end_times=None):
    inst = [eng.Instance(m[0], m[1], end_time=end_times[k] if end_times else 1e10) for k, m in enumerate(members)]
    b = eng.Batch(inst, weno_order=weno, bc=BC)
    for k, m in enumerate(members):
        b.set_prims(k, prims_of(*m))
    return b


def single_run(member, steps, weno, end_time=1e10):
    nx, ny = member[:2]
    cfg = eng.Config(nx=nx, ny=ny, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, weno_order=weno,
                     bc=(eng.PERIODIC, eng.OUTFLOW, eng.PERIODIC))
    with eng.Engine(cfg) as e:
        e.set_prims(prims_of(*member))
        for _ in range(steps):
            if e.time()[0] >= end_time:  # Simulation::evolve's loop condition
                break
            rc = e.update_time(end_time)
            assert rc == 0
        return e.download(), e.time()


def assert_same(got, want, what):
    for g, w, name in zip(got, want, ("cons", "prims", "aux")):
        if not np.array_equal(g, w):
            d = np.max(np.abs(g - w))
            raise AssertionError(f"{what}: {name} differs (max |diff| {d:.3e})")


@pytest.mark.parametrize("weno", [5, 3])
def test_members_equal_single_engine_bitwise(weno):
    steps = 6
    say(f"gpu: batch of {len(SMALL)} vs single engines, WENO{weno}, {steps} steps")
    with make_batch(SMALL, weno=weno) as b:
        rc, st = b.step(steps, stats=True)
        assert rc == 0 and st["failures"] == 0
        assert st["cells"] == 2 * steps * sum(m[0] * m[1] for m in SMALL)
        for k, m in enumerate(SMALL):
            want, wt = single_run(m, steps, weno)
            assert_same(b.download(k), want, f"member {k} {m[:2]}")
            assert b.time(k) == wt


def test_tile_shape_schedule_and_grid_do_not_change_results():
    say("gpu: batch tunings give identical results")
    ref = None
    for tx, ty in ((16, 16), (8, 32), (32, 8)):
        for sched in (0, 1):
            for blocks in (64, 4096):
                with make_batch(SMALL) as b:
                    b.set_tuning(tx, ty, sched, blocks)
                    assert b.step(3) == 0
                    out = [b.download(k) for k in range(len(SMALL))]
                if ref is None:
                    ref = out
                else:
                    for k in range(len(SMALL)):
                        assert_same(out[k], ref[k], f"tile {tx}x{ty} sched {sched} blocks {blocks} member {k}")


def test_graph_replays_equal_direct_launches():
    say("gpu: batch graph replays vs direct launches")
    with make_batch(SMALL) as g, make_batch(SMALL) as d:
        d.use_graph(False)
        for rep in range(3):
            assert g.step(2) == 0 and d.step(2) == 0
            # eager work and synchronisation between the replays
            for k in (0, len(SMALL) - 1):
                assert_same(g.download(k), d.download(k), f"replay {rep} member {k}")
                assert g.time(k) == d.time(k)


def test_end_times_stop_members_like_evolve():
    say("gpu: batch members with end times")
    first_dt = [single_run(m, 1, 5)[1][1] for m in SMALL[:3]]
    ends = [3.5 * first_dt[0], 1e10, 1.0 * first_dt[2]]
    with make_batch(SMALL[:3], end_times=ends) as b:
        assert b.step(6) == 0
        for k, m in enumerate(SMALL[:3]):
            want, wt = single_run(m, 6, 5, end_time=ends[k])
            assert_same(b.download(k), want, f"member {k} end {ends[k]}")
            assert b.time(k) == wt
        assert b.time(0)[0] == ends[0] and b.time(0)[2] == 4
        assert b.time(2)[2] == 1


def test_autotune_keeps_state_and_picks_fastest():
    say("gpu: batch autotuner")
    with make_batch(SMALL) as a, make_batch(SMALL) as b:
        assert a.step(1) == 0 and b.step(1) == 0
        best, cands = a.autotune(trial_steps=2)
        say(f"  chosen {best} of {len(cands)}")
        assert len(cands) >= 6
        assert best["ms_per_step"] == min(c["ms_per_step"] for c in cands)
        b.set_tuning(best["tile_x"], best["tile_y"], best["schedule"], best["blocks"])
        assert a.step(3) == 0 and b.step(3) == 0
        for k in range(len(SMALL)):
            assert_same(a.download(k), b.download(k), f"member {k} after autotune")
            assert a.time(k) == b.time(k)


def test_member_matches_oracle():
    m = SMALL[0]
    steps = 3
    say("gpu: batch member vs oracle")
    nx, ny = m[:2]
    grid = dict(nx=nx, ny=ny, nz=0, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0)
    o = orc.Oracle(orc.Case(weno_order=5, periodic=(1, 0, 1), **grid))
    oc, op, oa, _, _ = o.run(prims_of(*m), steps)
    with make_batch(SMALL) as b:
        assert b.step(steps) == 0
        c, p, a = b.download(0)
    for got, ref, names in ((c, oc, parity.SRMHD_CONS), (p, op, parity.SRMHD_PRIMS)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report("batch member vs oracle", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, f"{n}: {rg:.3e}"


def test_full_ensemble_two_steps():
    """The 64-member ensemble at its real sizes: two steps converge everywhere,
    and two members equal their one-at-a-time runs bit for bit."""
    members = idt.kh_ensemble()
    say(f"gpu: full ensemble {len(members)} members, {sum(m[0] * m[1] for m in members)} cells")
    with make_batch(members) as b:
        rc, st = b.step(2, stats=True)
        assert rc == 0 and st["failures"] == 0
        assert st["cells"] == 2 * 2 * sum(m[0] * m[1] for m in members)
        for k in (0, 1):
            want, wt = single_run(members[k], 2, 5)
            assert_same(b.download(k), want, f"member {k} {members[k][:2]}")
            assert b.time(k) == wt


def test_member_at_real_size_matches_reference(golden):
    """A member at a real ensemble size (KH 2-D 768 x 384, WENO5, RK2), run
    inside a batch with two other members of other sizes, 3 steps against the
    reference build: dt bit for bit, the middle row along each axis of every
    cons field to 1e-9 of its group scale, and every cons and prim field's
    min, max, mean and L2 norm to 1e-9."""
    z = golden("batch_member_768x384_weno5_rk2")
    m = tuple(int(v) if k in (0, 1, 4) else float(v) for k, v in enumerate(z["member"]))
    members = [m, SMALL[1], SMALL[2]]
    say(f"gpu: batch member {m[:2]} vs reference, 3 steps")
    with make_batch(members) as b:
        for s in range(3):
            assert b.step(1) == 0
            assert b.time(0)[1] == z["dts"][s]
        c, p, _ = b.download(0)
    nx, ny = m[:2]
    c, p = parity.interior(c, 4, 2), parity.interior(p, 4, 2)
    for got, ref, what in ((c[:, nx // 2, :, 0], z["row_x"], "row x"), (c[:, :, ny // 2, 0], z["row_y"], "row y")):
        rows = parity.field_errors(got, ref, parity.SRMHD_CONS)
        say(parity.report(f"batch member vs reference, {what}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-9, f"{what} {n}: {rg:.3e}"
    for got_a, ref, fields in ((c, z["cons_stats"], parity.SRMHD_CONS), (p, z["prims_stats"], parity.SRMHD_PRIMS)):
        scale = {n: max(abs(ref[v, 0]), abs(ref[v, 1])) for v, n in enumerate(fields)}
        norm = {n: ref[v, 3] for v, n in enumerate(fields)}
        for grp in parity._GROUPS:
            mem = [n for n in grp if n in scale]
            for n in mem:
                scale[n] = max(scale[k] for k in mem)
                norm[n] = max(norm[k] for k in mem)
        for v, name in enumerate(fields):
            x = got_a[v]
            got = np.array([x.min(), x.max(), x.mean(dtype=np.float64), np.sqrt(np.sum(x * x, dtype=np.float64))])
            diff = np.abs(got - ref[v])
            assert np.all(diff[:3] <= 1e-9 * max(scale[name], 1e-300)), name
            assert diff[3] <= 1e-9 * max(norm[name], 1e-300), name


def test_batch_retry_keeps_root_for_the_aux():
    """The batch's retry (run by the last workgroup of each stage launch), on a
    member whose step fails in both stages at three cells: the stage-two retry
    keeps the root, and the aux formed from it on download, with the prims and
    cons, must match the oracle's step; the other member is untouched by it.
    Without the last-cell quirk, for the reason test_gpu_srmhd's step test
    gives."""
    quirks = 0
    say(f"gpu: batch step with failing cells, quirks={quirks}")
    g = parity.C2P_RETRY_GRID
    o = orc.Oracle(orc.Case(weno_order=5, periodic=(1, 0, 1), quirks=quirks, **g))
    cons, prims, aux, gp, ga = parity.c2p_step_retry_case(o)
    dt = o.timestep(0.0, 0)
    oc, op, oa = cons.copy(), gp.copy(), ga.copy()
    _, ost = o.step(oc, op, oa, dt, "rk2")
    other = SMALL[0]
    inst = [eng.Instance(g["nx"], g["ny"]), eng.Instance(other[0], other[1])]
    with eng.Batch(inst, weno_order=5, bc=BC, quirks=quirks) as b:
        b.upload(0, cons, gp, ga)
        b.set_prims(1, prims_of(*other))
        rc, st = b.step(1, stats=True)
        assert b.time(0)[1] == dt
        c, p, a = b.download(0)
        c1, p1, a1 = b.download(1)
    say(f"  batch failures {st['failures']} retry failures {st['fallbacks']}, oracle failures {ost.failures}")
    assert st["failures"] == ost.failures == 2 * len(parity.C2P_RETRY_CELLS)
    assert st["fallbacks"] == 0
    for got, ref, names in ((c, oc, parity.SRMHD_CONS), (p, op, parity.SRMHD_PRIMS), (a, oa, parity.SRMHD_AUX)):
        rows = parity.field_errors(parity.interior(got, 4, 2), parity.interior(ref, 4, 2), names)
        say(parity.report(f"batch step with retries, quirks={quirks}", rows))
        for n, dm, fm, rf, rg, pw in rows:
            assert rg <= 1e-10, f"{n}: {rg:.3e}"
    (sc, sp, sa), _ = single_run(other, 1, 5)
    for x, y in ((c1, sc), (p1, sp), (a1, sa)):
        assert np.array_equal(x, y)
