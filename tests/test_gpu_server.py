"""Proof server (continuous batching) on the GPU.

Every served proof must equal the oracle's trapdoor closed form for its
(circuit, witness, r|s) byte for byte: the same bytes whether the request ran
alone, in a batch with other circuits, in continuous or sequential mode, or
after a failed batch. Admission control: a request whose footprint exceeds the
budget fails with HZK_ERR_NOMEM while the others complete; a budget that fits
one request keeps one in flight. An injected scratch-allocation failure fails
the requests of that MSM stage only and the loop keeps serving."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 21
LOGS = (6, 8, 10)


@pytest.fixture(scope="module")
def oracles(orc):
    return {lg: orc.Groth16(seed=SEED + lg, log_m=lg, num_public=8) for lg in LOGS}


def _server(mode="continuous", budget=0, max_batch=64):
    import liblightzk_amd as z
    s = z.ProofServer(0, mem_budget=budget, max_batch=max_batch, mode=mode)
    cids = {lg: s.add_circuit(SEED + lg, lg, 8) for lg in LOGS}
    return s, cids


def _requests(s, cids, count, seed0):
    import liblightzk_amd as z
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(count):
        lg = LOGS[int(rng.integers(0, len(LOGS)))]
        w = seed0 * 1000 + i
        out.append((lg, s.witness(cids[lg], w), z.Groth16.rs(w)))
    return out


@pytest.mark.parametrize("mode", ["continuous", "sequential"])
def test_served_proofs_equal_closed_form(oracles, mode):
    s, cids = _server(mode)
    reqs = _requests(s, cids, 24, 3)
    with s:
        s.start()
        # submitted from several client threads at once
        ids = [None] * len(reqs)

        def client(k):
            for i in range(k, len(reqs), 4):
                lg, zz, rs = reqs[i]
                ids[i] = s.submit(cids[lg], zz, rs)

        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        proofs = []
        for i, (lg, zz, rs) in enumerate(reqs):
            st, proof, lat = s.wait(ids[i], 300)
            assert st == 1, (i, st, s.last_error())
            assert lat > 0
            assert np.array_equal(proof, oracles[lg].prove_trapdoor(zz, rs)), (mode, i, lg)
            proofs.append(proof)
        stats = s.stats()
        assert stats["completed"] == len(reqs) and stats["failed"] == 0
        if mode == "sequential":
            assert stats["max_inflight"] == 1
    for (lg, zz, rs), p in list(zip(reqs, proofs))[:2]:
        assert oracles[lg].verify(zz[: 4 * 9], p)


def test_alone_equals_in_mix(oracles):
    s, cids = _server()
    reqs = _requests(s, cids, 12, 4)
    with s:
        s.start()
        lg, zz, rs = reqs[0]
        st, alone, _ = s.wait(s.submit(cids[lg], zz, rs), 300)
        assert st == 1
        ids = [s.submit(cids[l], a, b) for l, a, b in reqs]
        got = [s.wait(i, 300) for i in ids]
        assert all(g[0] == 1 for g in got)
        assert np.array_equal(got[0][1], alone)
        for (l, a, b), g in zip(reqs, got):
            assert np.array_equal(g[1], oracles[l].prove_trapdoor(a, b))


def test_budget_rejects_and_serialises(oracles):
    import liblightzk_amd as z
    probe, cids = _server()
    one = probe.footprint(cids[6])
    assert probe.footprint(cids[10]) > one + 1
    probe.close()
    # room for one small request only: the 2^10 circuit can never fit
    s, cids = _server(budget=one + 1)
    with s:
        s.start()
        big_z = s.witness(cids[10], 1)
        rid_big = s.submit(cids[10], big_z, z.Groth16.rs(1))
        small = [(s.witness(cids[6], 100 + i), z.Groth16.rs(100 + i)) for i in range(5)]
        ids = [s.submit(cids[6], a, b) for a, b in small]
        st, proof, _ = s.wait(rid_big, 300)
        assert st == -2, st   # HZK_ERR_NOMEM
        for (a, b), i in zip(small, ids):
            st, proof, _ = s.wait(i, 300)
            assert st == 1, s.last_error()
            assert np.array_equal(proof, oracles[6].prove_trapdoor(a, b))
        stats = s.stats()
        assert stats["max_inflight"] == 1
        assert stats["failed"] == 1 and stats["completed"] == 5


def test_injected_failure_is_isolated(oracles):
    import liblightzk_amd as z
    s, cids = _server(mode="sequential")
    with s:
        s.start()
        zz, rs = s.witness(cids[8], 7), z.Groth16.rs(7)
        st, good, _ = s.wait(s.submit(cids[8], zz, rs), 300)
        assert st == 1
        s.inject_failure(0)
        st, _, _ = s.wait(s.submit(cids[8], zz, rs), 300)
        assert st == -2, st
        assert "injected" in s.last_error()
        # the loop keeps serving, keys intact, same bytes as before the failure
        for w in (7, 8):
            zz, rs = s.witness(cids[8], w), z.Groth16.rs(w)
            st, proof, _ = s.wait(s.submit(cids[8], zz, rs), 300)
            assert st == 1
            assert np.array_equal(proof, oracles[8].prove_trapdoor(zz, rs))
        assert s.stats()["failed"] == 1


def test_server_restart_from_key_files_with_level_tables(ctx, orc, oracles, tmp_path):
    """A server restarted from saved proving keys (hzk_server_add_circuit_file),
    one of them proving over level tables (hzk_server_set_levels) and one the
    reference's MatMul statement, serves a mixed stream whose proofs equal the
    closed form."""
    import liblightzk_amd as z
    paths = {}
    for lg in LOGS:
        g = z.Groth16(ctx, seed=SEED + lg, log_m=lg, num_public=8)
        paths[lg] = str(tmp_path / f"c{lg}.hzk")
        g.save(paths[lg])
        g.close()
    mm = z.Groth16.matmul(ctx, 6, 4)          # the reference's MatMul statement, p = 4
    mm_path = str(tmp_path / "matmul.hzk")
    mm.save(mm_path)
    mm.close()
    s = z.ProofServer(0, max_batch=16, mode="continuous")
    cids = {lg: s.add_circuit_file(paths[lg]) for lg in LOGS}
    s.set_levels(cids[LOGS[-1]], 4)
    mm_cid = s.add_circuit_file(mm_path)
    reqs = _requests(s, cids, 12, 9)
    mm_reqs = [(s.witness(mm_cid, 70 + i), z.Groth16.rs(70 + i)) for i in range(3)]
    with s:
        s.start()
        ids = [s.submit(cids[lg], zz, rs) for lg, zz, rs in reqs]
        mm_ids = [s.submit(mm_cid, zz, rs) for zz, rs in mm_reqs]
        for (lg, zz, rs), rid in zip(reqs, ids):
            st, proof, _ = s.wait(rid, 120)
            assert st == 1, s.last_error()
            assert np.array_equal(proof, oracles[lg].prove_trapdoor(zz, rs)), lg
        o = orc.Groth16(6, matmul_p=4)
        for i, ((zz, rs), rid) in enumerate(zip(mm_reqs, mm_ids)):
            st, proof, _ = s.wait(rid, 120)
            assert st == 1, s.last_error()
            assert np.array_equal(zz, o.witness(70 + i))
            assert np.array_equal(proof, o.prove_trapdoor(zz, rs))
