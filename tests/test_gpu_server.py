"""Serving loop: continuous batching of independent placements on one GPU.

Each job's result must be bit-identical to the same design placed alone with
erp_run_until(max_iterations, stop_overflow, check_every): the batched kernels
only touch a job's own buffers and reduce in a fixed order, and the server
checks convergence at the same iteration counts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHECK = 10


def _designs():
    from erplacer.netlist import generate
    specs = [(1500, 64, 11), (3000, 64, 12), (2200, 64, 13), (6000, 128, 14), (9000, 128, 15),
             (4000, 64, 16), (12000, 128, 17)]
    out = []
    for cells, grid, seed in specs:
        d = generate(cells, int(cells * 1.05), 32, 6, 0.1, 0.7, seed, name=f"job{seed}")
        out.append((d, grid))
    return out


def _alone(d, grid, max_it, stop):
    from erplacer.placement import Placement
    with Placement(d, grid, target_density=1.0) as P:
        info = P.run_until(max_it, stop, CHECK)
        x, y = P.positions()
        return info, x, y, P.final_hpwl_dbu()


@pytest.mark.parametrize("max_active,eager", [(3, False), (16, False), (16, True)])
def test_alone_equals_mix(gpu_lib, max_active, eager):
    """eager: the batched kernels launched directly each round instead of the
    captured graph (the mode for kernel tracers); the same results."""
    from erplacer.server import Server
    jobs = _designs()
    max_it, stop = 237, 0.25        # not a multiple of CHECK: the last chunk is shorter
    alone = [_alone(d, g, max_it, stop) for d, g in jobs]
    with Server(max_active=max_active, check_every=CHECK, eager=eager) as srv:
        ids = [srv.submit(d, g, max_it, stop, target_density=1.0) for d, g in jobs]
        res = {i: srv.wait(i) for i in ids}
        st = srv.stats()
        for (d, g), jid, (info, x, y, dbu) in zip(jobs, ids, alone):
            r = res[jid]
            assert r["status"] == 0
            assert r["iterations"] == info["iteration"], (jid, r, info)
            assert r["hpwl"] == info["hpwl"] and r["overflow"] == info["overflow"]
            assert r["hpwl_dbu"] == dbu
            sx, sy = srv.positions(jid)
            assert np.array_equal(sx, x) and np.array_equal(sy, y), jid
            srv.release(jid)
    assert st["max_batch"] >= 2          # same-grid jobs really ran together
    assert st["max_batch"] <= max_active
    print("server stats", st)


def test_budget_admission(gpu_lib):
    """A budget that fits two jobs keeps at most two resident."""
    from erplacer.netlist import generate
    from erplacer.server import Server, estimate_device_bytes
    jobs = [(generate(3000, 3150, 32, 6, 0.1, 0.7, 40 + k), 64) for k in range(5)]
    est = [estimate_device_bytes(d, g, target_density=1.0) for d, g in jobs]
    need = max(est)
    assert 3 * min(est) > 2 * need + 1      # no third job fits
    with Server(hbm_budget_bytes=2 * need + 1, max_active=64, check_every=CHECK) as srv:
        ids = [srv.submit(d, g, 120, 0.3, target_density=1.0) for d, g in jobs]
        srv.drain()
        got = srv.poll()
        assert sorted(r["job_id"] for r in got) == sorted(ids)
        assert all(r["status"] == 0 for r in got)
        assert srv.stats()["max_batch"] <= 2
        # intervals [t_admit, t_done] never overlap more than twice
        ev = sorted([(r["t_admit"], 1) for r in got] + [(r["t_done"], -1) for r in got],
                    key=lambda e: (e[0], e[1]))
        live = peak = 0
        for _, k in ev:
            live += k
            peak = max(peak, live)
        assert peak <= 2
    with Server(hbm_budget_bytes=need // 2, check_every=CHECK) as srv:
        from erplacer import lib as L
        with pytest.raises(L.ErpError):
            srv.submit(jobs[0][0], jobs[0][1], 100, 0.1, target_density=1.0)


def test_estimate_covers_allocation(gpu_lib):
    from erplacer.placement import Placement
    from erplacer.server import estimate_device_bytes
    for d, g in _designs()[:4]:
        est = estimate_device_bytes(d, g, target_density=1.0)
        with Placement(d, g, target_density=1.0) as P:
            real = P.device_bytes()
        assert real <= est < 2 * real + (4 << 20), (real, est)


def test_arrivals_while_running(gpu_lib):
    """Continuous batching: jobs submitted while earlier ones are on the GPU are
    admitted beside them (their [t_admit, t_done] intervals overlap), a design
    that fails validation at admission fails alone, and every other result is
    bit-identical to the job placed alone."""
    import dataclasses
    from erplacer.server import Server
    jobs = _designs()
    # first wave: one short job and two long ones (stop 0: they run to max_it)
    first = [(0, 60, 0.25), (6, 1200, 0.0), (4, 1200, 0.0)]
    second = [(1, 237, 0.25), (2, 237, 0.25), (3, 237, 0.25), (5, 237, 0.25)]
    alone = {k: _alone(jobs[k][0], jobs[k][1], it, stop) for k, it, stop in first + second}
    d_bad = jobs[1][0]
    x_bad = d_bad.x.copy()
    x_bad[d_bad.num_movable] = -5.0          # a fixed macro at a negative coordinate
    d_bad = dataclasses.replace(d_bad, x=x_bad)
    with Server(max_active=16, check_every=CHECK) as srv:
        ids = {k: srv.submit(jobs[k][0], jobs[k][1], it, stop, target_density=1.0)
               for k, it, stop in first}
        assert srv.wait(ids[0])["status"] == 0          # the long jobs are still running
        bad = srv.submit(d_bad, jobs[1][1], 100, 0.25, target_density=1.0)
        for k, it, stop in second:
            ids[k] = srv.submit(jobs[k][0], jobs[k][1], it, stop, target_density=1.0)
        res = {k: srv.wait(i) for k, i in ids.items()}
        rbad = srv.wait(bad)
        for k, (info, x, y, dbu) in alone.items():
            r = res[k]
            assert r["status"] == 0, (k, r)
            assert r["iterations"] == info["iteration"], (k, r, info)
            assert r["hpwl"] == info["hpwl"] and r["overflow"] == info["overflow"]
            assert r["hpwl_dbu"] == dbu
            sx, sy = srv.positions(ids[k])
            assert np.array_equal(sx, x) and np.array_equal(sy, y), k
    assert rbad["status"] != 0
    # a second-wave job was admitted before a long first-wave job finished
    long_done = min(res[6]["t_done"], res[4]["t_done"])
    assert any(res[k]["t_admit"] < long_done for k, _, _ in second)


def test_alone_equals_mix_large_jobs(gpu_lib):
    """The serving stream's largest jobs: 150k cells on the grid the stream's
    rule gives (256^2) and on 512^2 grids, beside smaller jobs; two jobs share
    each of the 512^2 and 256^2 batches.  Every result equals its alone run."""
    from erplacer import workloads as W
    from erplacer.netlist import generate
    from erplacer.server import Server
    specs = [(150_000, 512, 61), (150_000, W.serving_grid(150_000), 62), (90_000, 512, 63),
             (60_000, 256, 64), (20_000, 128, 65)]
    assert W.serving_grid(150_000) == 256
    jobs = [(generate(c, int(c * 1.05), 128, 20, 0.1, 0.75, s, name=f"big{s}"), g)
            for c, g, s in specs]
    max_it, stop = 400, 0.1
    alone = [_alone(d, g, max_it, stop) for d, g in jobs]
    with Server(max_active=8, check_every=CHECK) as srv:
        ids = [srv.submit(d, g, max_it, stop, target_density=1.0) for d, g in jobs]
        res = {i: srv.wait(i) for i in ids}
        st = srv.stats()
        for jid, (info, x, y, dbu) in zip(ids, alone):
            r = res[jid]
            assert r["status"] == 0
            assert r["iterations"] == info["iteration"], (jid, r, info)
            assert r["hpwl"] == info["hpwl"] and r["overflow"] == info["overflow"]
            assert r["hpwl_dbu"] == dbu
            sx, sy = srv.positions(jid)
            assert np.array_equal(sx, x) and np.array_equal(sy, y), jid
    assert st["max_batch"] >= 2
    print("server stats", st)
