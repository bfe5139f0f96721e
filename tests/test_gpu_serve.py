"""GPU tests of the request server (hsolve_serve_*): Poisson Q1 MG-IR solves
of mixed sizes with continuous batching.

What they pin:
  * a served solve equals the single-solve MG-IR handle bit for bit, and the
    oracle's Multigrid<MIXED> restatement within the solver tolerances
    (iteration counts +-1, solution 1e-6 relative);
  * the device-generated right-hand side equals problems.rhs_uniform (a
    served solve with it equals one with the host-generated b uploaded);
  * the same request gives the same bits alone and inside a mix of other
    sizes submitted while it runs, with and without graph replay;
  * admission never holds more request state than the memory budget, and a
    request larger than the budget is refused with HSOLVE_ERR_BUDGET."""
import numpy as np
import pytest

from honei_amd import problems as P

pytestmark = pytest.mark.gpu

SEED = P.SEED


def _levels(maxl, minl):
    from honei_fixtures import dinv_reference
    a32, d32 = {}, {}
    for L in range(minl, maxl + 1):
        n = P.side_of_level(L) ** 2
        a32[L] = [b.astype(np.float32) for b in P.poisson_q1_bands(n)]
        d32[L] = dinv_reference(a32[L][4])
    return a32, d32


def _server(hip, **kw):
    args = dict(min_level=2, min_served_level=6, max_served_level=9, tolerance=1e-8,
                memory_budget=4 << 30, max_slots=16, use_graphs=True)
    args.update(kw)
    return hip.Server(hip.serve_params(**args))


def _solve_alone(hip, srv, level, seed):
    n = P.side_of_level(level) ** 2
    x = hip.DeviceVector.zeros(n)
    srv.submit(level, seed=seed, x=x)
    (res,) = srv.drain(1).values()
    return x.get(), res


@pytest.mark.parametrize("level", [6, 8, 10, 11])
def test_served_solve_matches_single_solve_and_oracle(hip, oracle, level):
    n = P.side_of_level(level) ** 2
    b = P.rhs_uniform(n, seed=SEED + level)
    srv = _server(hip, max_served_level=max(level, 9), memory_budget=8 << 30)
    x_srv, res = _solve_alone(hip, srv, level, SEED + level)
    assert res["status"] == 0
    # the single-solve handle on the same b from x0 = 0
    mg = hip.MGSolver(hip.mg_params(2, level, tolerance=1e-8))
    bd, xd = hip.DeviceVector.from_numpy(b), hip.DeviceVector.zeros(n)
    info = mg.solve(bd, xd)
    assert np.array_equal(x_srv, xd.get())
    assert (res["outer"], res["inner"]) == (info["outer"], info["inner"])
    # host-generated b uploaded gives the same bits as the device-generated one
    x2 = hip.DeviceVector.zeros(n)
    srv.submit(level, b=bd, x=x2)
    (res2,) = srv.drain(1).values()
    assert np.array_equal(x2.get(), x_srv) and res2["outer"] == res["outer"]
    # the oracle's restatement of Multigrid<MIXED>
    a32, d32 = _levels(level, 2)
    ref = oracle.mg_mixed_q1(P.poisson_q1_bands(n), b, np.zeros(n), a32, d32, min_level=2, max_level=level,
                             tolerance=1e-8)
    assert abs(res["outer"] - ref["outer"]) <= 1 and abs(res["inner"] - ref["inner"]) <= 2, (res, ref)
    rel = np.linalg.norm(x_srv - ref["x"]) / np.linalg.norm(ref["x"])
    assert rel <= 1e-6
    assert res["final_defect"] < 1e-8 * res["initial_defect"]


@pytest.mark.parametrize("graphs", [True, False])
def test_same_result_alone_and_in_mix(hip, graphs):
    srv = _server(hip, use_graphs=graphs)
    probe = (7, SEED + 7)
    x_alone, r_alone = _solve_alone(hip, srv, *probe)
    # a mix of sizes; the probe enters while others are running, and more
    # arrive after it
    rng = np.random.default_rng(5)
    first = [(int(L), int(s)) for L, s in zip(rng.integers(6, 10, 6), rng.integers(1, 1 << 30, 6))]
    later = [(int(L), int(s)) for L, s in zip(rng.integers(6, 10, 6), rng.integers(1, 1 << 30, 6))]
    outs = {}
    for L, s in first:
        x = hip.DeviceVector.zeros(P.side_of_level(L) ** 2)
        outs[srv.submit(L, seed=s, x=x)] = (L, s, x)
    for _ in range(3):
        srv.step()
    xp = hip.DeviceVector.zeros(P.side_of_level(7) ** 2)
    pid = srv.submit(7, seed=probe[1], x=xp)
    srv.step()
    for L, s in later:
        x = hip.DeviceVector.zeros(P.side_of_level(L) ** 2)
        outs[srv.submit(L, seed=s, x=x)] = (L, s, x)
    got = srv.drain(len(first) + len(later) + 1)
    st = srv.stats()
    assert st["peak_running"] >= 2, st          # solves did overlap
    assert np.array_equal(xp.get(), x_alone)
    assert (got[pid]["outer"], got[pid]["inner"]) == (r_alone["outer"], r_alone["inner"])
    for rid, (L, s, x) in outs.items():
        r = got[rid]
        assert r["status"] == 0 and r["level"] == L
        assert r["final_defect"] < 1e-8 * r["initial_defect"]
        assert r["t_done"] >= r["t_admit"] >= r["t_submit"]


def test_graph_and_direct_launch_agree(hip):
    a = _server(hip, use_graphs=True)
    b = _server(hip, use_graphs=False)
    for L, s in [(6, 11), (9, 12)]:
        xa, ra = _solve_alone(hip, a, L, s)
        xb, rb = _solve_alone(hip, b, L, s)
        assert np.array_equal(xa, xb) and ra["outer"] == rb["outer"] and ra["inner"] == rb["inner"]


def test_memory_budget_admission(hip):
    probe = _server(hip)
    per = probe.request_bytes(8)
    del probe
    srv = _server(hip, memory_budget=2 * per + per // 2)
    outs = []
    for s in range(5):
        x = hip.DeviceVector.zeros(P.side_of_level(8) ** 2)
        outs.append((srv.submit(8, seed=100 + s, x=x), x))
    got = {}
    while len(got) < 5:
        srv.step()
        st = srv.stats()
        assert st["running"] <= 2 and st["bytes_in_use"] <= st["budget"], st
        assert st["bytes_held"] <= st["budget"], st
        for r in srv.poll():
            got[r["id"]] = r
    assert all(r["status"] == 0 for r in got.values())
    # two fit at once; the third waited for the first completion
    adm = sorted(r["t_admit"] for r in got.values())
    dones = sorted(r["t_done"] for r in got.values())
    assert adm[2] >= dones[0]
    # a level-9 request cannot fit this budget at all
    with pytest.raises(Exception) as e:
        srv.submit(9, seed=1)
    assert "budget" in str(e.value)


def test_run_trace_latencies(hip):
    srv = _server(hip)
    rng = np.random.default_rng(3)
    k = 24
    arrivals = np.cumsum(rng.exponential(0.002, k))
    levels = rng.integers(6, 10, k)
    seeds = rng.integers(1, 1 << 40, k)
    res = srv.run(arrivals, levels, seeds)
    assert len(res) == k
    for i, r in enumerate(res):
        assert r["status"] == 0 and r["level"] == levels[i]
        assert abs(r["t_submit"] - arrivals[i]) < 1e-9
        assert r["t_done"] > r["t_admit"] >= r["t_submit"] - 1e-9
        assert r["final_defect"] < 1e-8 * r["initial_defect"]


def test_slot_count_not_a_power_of_two(hip):
    """A tick covers the table entries up to the high-water entry rounded up
    to a power of two: with 5 slots per level that is 8 entries, past the
    configured count.  Seven concurrent requests of one level (5 running, 2
    waiting) each give the bits of the same request solved alone."""
    srv = _server(hip, max_slots=5)
    seeds = [300 + s for s in range(7)]
    alone = {s: _solve_alone(hip, srv, 6, s) for s in seeds}
    outs = {}
    for s in seeds:
        x = hip.DeviceVector.zeros(P.side_of_level(6) ** 2)
        outs[srv.submit(6, seed=s, x=x)] = (s, x)
    got = srv.drain(len(seeds))
    assert srv.stats()["peak_running"] == 5
    for rid, (s, x) in outs.items():
        assert np.array_equal(x.get(), alone[s][0]), s
        assert (got[rid]["outer"], got[rid]["inner"]) == (alone[s][1]["outer"], alone[s][1]["inner"])
