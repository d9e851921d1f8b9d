"""GPU tests of the continuous-batching server: every request's scores equal the same request scored alone (and the
oracle), whichever round it joins, with host or HBM-resident tapes, under a budget that forces many rounds."""

import threading

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, dna_tables, protein_tables, random_bytes, random_utf8

pytestmark = pytest.mark.gpu

szs = pytest.importorskip("stringzillas_amd")


def _kinds(rng):
    b2c, costs = protein_tables(rng)
    dna_b2c, dna_costs = dna_tables()
    return [
        (szs.LevenshteinDistances(), O.Params("levenshtein"), DNA, False),
        (szs.LevenshteinDistances(0, 4, 3, 2), O.Params("levenshtein", 0, 4, 3, 2), DNA, False),
        (szs.LevenshteinDistancesUTF8(), O.Params("levenshtein_utf8"), None, True),
        (szs.NeedlemanWunschScores(b2c, costs, -11, -1),
         O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs), PROTEIN, False),
        (szs.SmithWatermanScores(dna_b2c, dna_costs, -2, -2),
         O.Params("smith_waterman", open=-2, extend=-2, byte_to_class=dna_b2c, class_costs=dna_costs), DNA, False),
    ]


def _requests(seed, count):
    rng = np.random.default_rng(seed)
    kinds = _kinds(rng)
    requests = []
    for r in range(count):
        engine, params, alphabet, utf8 = kinds[r % len(kinds)]
        pairs = [0, 1, 3, 17, 40][r % 5]
        high = [8, 90, 700, 1500][r % 4]

        def make():
            if utf8:
                return random_utf8(rng, 0, high)
            return random_bytes(rng, alphabet, 0, high)

        firsts = szs.Strs([make() for _ in range(pairs)])
        seconds = szs.Strs([make() for _ in range(pairs)])
        requests.append((engine, params, firsts, seconds))
    return requests


def _expected(params, firsts, seconds):
    return O.pairs(params, (firsts.data, firsts.offsets), (seconds.data, seconds.offsets), threads=8)


@pytest.mark.parametrize("budget_mib,inflight", [(1024, 2), (4, 3)])
def test_server_results_equal_results_alone(budget_mib, inflight):
    requests = _requests(7 + budget_mib, 40)
    server = szs.Server(hbm_budget_bytes=budget_mib << 20, max_inflight_rounds=inflight)
    try:
        tickets = []
        for i, (engine, _, firsts, seconds) in enumerate(requests):
            if i % 3 == 2:  # HBM-resident tapes are scored in place
                firsts, seconds = szs.to_device(firsts), szs.to_device(seconds)
            tickets.append((server.submit(engine, firsts, seconds), firsts, seconds))
        for (ticket, _, _), (engine, params, firsts, seconds) in zip(tickets, requests):
            out = server.wait(ticket)
            alone = engine.pairs(firsts, seconds)
            assert out.dtype == alone.dtype
            np.testing.assert_array_equal(out, alone)
            np.testing.assert_array_equal(out, _expected(params, firsts, seconds))
        assert server.bytes_in_use == 0
    finally:
        server.close()


def test_server_concurrent_submitters_and_polling():
    requests = _requests(11, 30)
    server = szs.Server(hbm_budget_bytes=64 << 20, max_inflight_rounds=2)
    results = {}
    errors = []

    def submitter(part):
        try:
            for index in range(part, len(requests), 3):
                engine, _, firsts, seconds = requests[index]
                results[index] = server.submit(engine, firsts, seconds)
        except Exception as error:  # surfaced below
            errors.append(error)

    threads = [threading.Thread(target=submitter, args=(p,)) for p in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    try:
        assert not errors
        for index, ticket in results.items():
            engine, params, firsts, seconds = requests[index]
            while not server.poll(ticket):
                pass
            np.testing.assert_array_equal(server.wait(ticket), _expected(params, firsts, seconds))
    finally:
        server.close()


def test_server_rejects_requests_over_budget_and_bad_shapes():
    rng = np.random.default_rng(3)
    server = szs.Server(hbm_budget_bytes=1 << 20, max_inflight_rounds=1)
    try:
        big = szs.Strs([random_bytes(rng, DNA, 4000, 4000) for _ in range(300)])
        with pytest.raises(MemoryError):
            server.submit(szs.LevenshteinDistances(), big, big)
        with pytest.raises(ValueError):
            server.submit(szs.LevenshteinDistances(), szs.Strs(["a", "b"]), szs.Strs(["a"]))
        ticket = server.submit(szs.LevenshteinDistances(), szs.Strs(["kitten"]), szs.Strs(["sitting"]))
        assert server.wait(ticket).tolist() == [3]
    finally:
        server.close()


def test_server_stats_completion_times_and_scores_beyond_32_bits():
    import time

    rng = np.random.default_rng(5)
    server = szs.Server(hbm_budget_bytes=512 << 20, max_inflight_rounds=2)
    try:
        heavy = szs.LevenshteinDistances(0, 127, 127, 127)
        long_side = szs.Strs([b"A" * (17 << 20)])
        short_side = szs.Strs([b"C"])
        normal = szs.Strs([random_bytes(rng, DNA, 10, 300) for _ in range(20)])
        others = szs.Strs([random_bytes(rng, DNA, 10, 300) for _ in range(20)])
        before = time.perf_counter()
        beyond = server.submit(heavy, long_side, short_side)
        fine = server.submit(heavy, normal, others)
        assert server.wait(beyond).tolist() == [127 * (17 << 20)]  # a distance past 2^31, as u32 cells hold it
        out, completed = server.wait_timed(fine)
        assert before <= completed <= time.perf_counter()
        expected = O.pairs(O.Params("levenshtein", 0, 127, 127, 127), (normal.data, normal.offsets),
                           (others.data, others.offsets), threads=4)
        np.testing.assert_array_equal(out, expected)
        stats = server.stats()
        assert stats["requests"] == 2 and stats["rounds"] >= 1 and stats["cells"] > 0
        assert stats["pending"] == 0 and stats["inflight"] == 0
    finally:
        server.close()


class _Understated:
    """An HBM tape whose stated byte count is below what its offsets span."""

    def __init__(self, device_strs, stated):
        self._inner, self.nbytes = device_strs, stated

    def __len__(self):
        return len(self._inner)

    def tape(self):
        return self._inner.tape()


def test_understated_tape_bytes_fail_only_their_request():
    rng = np.random.default_rng(9)
    strings = [random_utf8(rng, 200, 400) for _ in range(30)]
    honest = szs.to_device(szs.Strs(strings))
    understated = _Understated(honest, honest.nbytes // 3)
    engine = szs.LevenshteinDistancesUTF8()
    with pytest.raises(ValueError):
        szs.Batch([(engine, understated, understated)])
    server = szs.Server(hbm_budget_bytes=256 << 20, max_inflight_rounds=1)
    try:
        bad = server.submit(engine, understated, understated)
        good = server.submit(engine, honest, honest)
        with pytest.raises(ValueError):
            server.wait(bad)
        assert (server.wait(good) == 0).all()  # every string against itself
    finally:
        server.close()


def test_server_rounds_with_chained_pairs():
    """Pairs taller than the chained threshold split over workgroups inside server rounds (chain buffers come from
    the server's pool); their scores equal the oracle's and the same request scored alone."""
    rng = np.random.default_rng(17)
    b2c, costs = protein_tables(rng)
    nw = szs.NeedlemanWunschScores(b2c, costs, -11, -1)
    params = O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs)
    server = szs.Server(hbm_budget_bytes=256 << 20, max_inflight_rounds=2)
    try:
        requests = []
        for length in (25_000, 26_000, 300):
            text = random_bytes(rng, PROTEIN, length, length)
            firsts = szs.Strs([text, random_bytes(rng, PROTEIN, length, length)])
            seconds = szs.Strs([random_bytes(rng, PROTEIN, length, length), text])
            requests.append((server.submit(nw, firsts, seconds), firsts, seconds))
        for ticket, firsts, seconds in requests:
            out = server.wait(ticket)
            np.testing.assert_array_equal(out, _expected(params, firsts, seconds))
            np.testing.assert_array_equal(out, nw.pairs(firsts, seconds))
    finally:
        server.close()


def test_serving_workload_shape_matches_each_request_alone_and_the_oracle():
    """The serving workload's own shape: Poisson arrivals at 1,000 req/s for 1 s, each request 1..4,096 pairs
    (log-uniform) of one kind drawn from the mixed batch's mix (50% byte Levenshtein over ACGT, 25% UTF-8
    Levenshtein, 25% affine NW) with both lengths log-uniform in [16, 16384], tapes resident in HBM. Every
    completed request must equal the same request scored alone, and a sample holding each kind's longest served
    pair must equal the oracle."""
    from stringzillas_amd import serving as S
    from stringzillas_amd import workloads as W

    templates = S.make_templates(count=250, seed=0x535A0044)
    engines = S.engines_by_kind()
    summary, results = S.run_server(templates, engines, qps=1000.0, duration=1.0, hbm_budget=16 << 30,
                                    drain_timeout=60.0)
    assert summary["dropped_after_drain_timeout"] == 0 and summary["completed"] == len(results) > 600
    assert summary["largest_round"] > 1  # requests arriving during a round joined the next one together
    check = S.verify(templates, engines, results)
    assert check["equal_to_alone"] == check["verified"] == len(results)
    b2c, costs = W.protein_matrix(np.random.default_rng(W.SEED_BASE + 1))
    params = {"levenshtein": O.Params("levenshtein"), "levenshtein_utf8": O.Params("levenshtein_utf8"),
              "needleman_wunsch": O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c,
                                           class_costs=costs)}
    rng = np.random.default_rng(4)
    served = {}
    for pick, out in results:
        served.setdefault(pick, out)
    for kind in params:
        picks = [p for p in served if templates[p].kind == kind]
        assert picks, f"no {kind} request was served"
        sample = []  # (pick, pair index): the kind's longest pair, then random pairs
        for p in picks:
            host = templates[p].firsts.host, templates[p].seconds.host
            sizes = np.diff(host[0].offsets.astype(np.int64)) * np.diff(host[1].offsets.astype(np.int64))
            sample.append((int(sizes.max()), p, int(sizes.argmax())))
        sample = [max(sample)[1:]] + [(p, int(rng.integers(0, len(templates[p].firsts))))
                                      for p in rng.choice(picks, size=min(8, len(picks)), replace=False)]
        for p, i in sample:
            firsts, seconds = templates[p].firsts.host, templates[p].seconds.host
            want = O.score(params[kind], firsts[i], seconds[i])
            assert int(served[p][i]) == want, f"{kind} request {p} pair {i}"
