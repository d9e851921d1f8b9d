"""CPU tests that pin the oracle: the reference's own known answers and scorer outputs (committed fixtures), a
live differential against the reference scorers when oracle/_ref is built, brute-force alignment enumeration on
tiny strings, and closed forms. A deliberate mistake in the oracle (a dropped border term, a swapped operand, a
wrong discard value) fails at least one of these.
"""

import base64
import itertools
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import DNA, PROTEIN, protein_tables, random_bytes, random_utf8

FIXTURES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")


@pytest.fixture(scope="module")
def fixtures():
    with open(FIXTURES) as f:
        return json.load(f)


def _params_from_vector(vector, tables):
    params = O.Params(vector["kind"], match=vector["match"], mismatch=vector["mismatch"], open=vector["open"],
                      extend=vector["extend"])
    if vector.get("table") is not None:
        table = tables[str(vector["table"])]
        params.byte_to_class = np.frombuffer(base64.b64decode(table["byte_to_class"]), dtype=np.uint8).copy()
        params.class_costs = np.frombuffer(base64.b64decode(table["class_costs"]), dtype=np.int8).reshape(32, 32).copy()
    return params


def test_known_answers_of_the_reference_tests(fixtures):
    for case in fixtures["known_answers"]:
        params = O.Params(case["kind"], **case["params"])
        got = O.score(params, case["first"].encode(), case["second"].encode())
        assert got == case["expected"], case


def test_reference_scorer_vectors(fixtures):
    for vector in fixtures["reference_vectors"]:
        params = _params_from_vector(vector, fixtures["tables"])
        firsts = [base64.b64decode(s) for s in vector["firsts"]]
        seconds = [base64.b64decode(s) for s in vector["seconds"]]
        got = O.pairs(params, O.tape(firsts), O.tape(seconds), threads=4)
        np.testing.assert_array_equal(got, np.array(vector["expected"], dtype=np.int64),
                                      err_msg=f"{vector['kind']} o={vector['open']} e={vector['extend']}")


@pytest.mark.skipif(not O.reference_available("serial"), reason="oracle/_ref is built only where the reference is mounted")
@pytest.mark.parametrize("variant", ["serial", "icelake"])
def test_live_differential_against_reference(variant):
    if not O.reference_available(variant):
        pytest.skip(f"reference {variant} scorer unavailable on this host")
    rng = np.random.default_rng(17)
    b2c, costs = protein_tables(rng, asymmetric=True)
    cases = [
        (O.Params("levenshtein"), DNA),
        (O.Params("levenshtein", 0, 4, 3, 2), DNA),
        (O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs), PROTEIN),
        (O.Params("smith_waterman", open=-5, extend=-1, byte_to_class=b2c, class_costs=costs), PROTEIN),
        (O.Params("smith_waterman", open=2, extend=1, byte_to_class=b2c, class_costs=costs), PROTEIN),
        (O.Params("levenshtein_utf8"), None),
    ]
    for params, alphabet in cases:
        for low, high in [(0, 10), (50, 400), (600, 1400)]:
            make = (lambda: random_utf8(rng, low, high)) if alphabet is None else (lambda: random_bytes(rng, alphabet, low, high))
            a, b = [make() for _ in range(16)], [make() for _ in range(16)]
            mine = O.pairs(params, O.tape(a), O.tape(b), threads=8)
            theirs = O.reference_pairs(params, O.tape(a), O.tape(b), threads=8, variant=variant)
            np.testing.assert_array_equal(mine, theirs, err_msg=f"{params.kind} {low}-{high}")


#  Brute force: enumerate every alignment of tiny strings --------------------------------------------------------


def _alignments(n, m):
    """All sequences of M (substitute), V (skip a row symbol), H (skip a column symbol) consuming n and m."""
    if n == 0 and m == 0:
        yield ()
        return
    if n and m:
        for rest in _alignments(n - 1, m - 1):
            yield ("M",) + rest
    if n:
        for rest in _alignments(n - 1, m):
            yield ("V",) + rest
    if m:
        for rest in _alignments(n, m - 1):
            yield ("H",) + rest


def _alignment_score(ops, a, b, sub, open_, extend):
    """Textbook cost: a maximal gap run of length L costs open + (L - 1) * extend."""
    i = j = 0
    total = 0
    previous = None
    for op in ops:
        if op == "M":
            total += sub(a[i], b[j])
            i += 1
            j += 1
        else:
            total += extend if previous == op else open_
            if op == "V":
                i += 1
            else:
                j += 1
        previous = op
    return total


def _brute(params, a, b):
    if params.kind in ("levenshtein", "levenshtein_utf8"):
        sub = lambda x, y: params.match if x == y else params.mismatch  # noqa: E731
        pick = min
    else:
        sub = lambda x, y: int(params.class_costs[params.byte_to_class[x], params.byte_to_class[y]])  # noqa: E731
        pick = max
    if params.kind != "smith_waterman":
        return pick(_alignment_score(ops, a, b, sub, params.open, params.extend) for ops in _alignments(len(a), len(b)))
    best = 0
    for i1, i2 in itertools.combinations(range(len(a) + 1), 2):
        for j1, j2 in itertools.combinations(range(len(b) + 1), 2):
            sa, sb = a[i1:i2], b[j1:j2]
            best = max(best, max(_alignment_score(ops, sa, sb, sub, params.open, params.extend)
                                 for ops in _alignments(len(sa), len(sb))))
    return best


@pytest.mark.parametrize("case", range(6))
def test_brute_force_enumeration(case):
    """Normal signs only (open and extend no better than ... see module doc): there the textbook run cost and the
    reference's Gotoh recurrence agree; odd signs are pinned by the reference vectors above."""
    rng = np.random.default_rng(100 + case)
    b2c = np.zeros(256, dtype=np.uint8)
    b2c[ord("a")], b2c[ord("b")], b2c[ord("c")] = 1, 2, 3
    costs = rng.integers(-3, 2, size=(32, 32)).astype(np.int8)
    np.fill_diagonal(costs, 3)
    all_params = [
        O.Params("levenshtein"),
        O.Params("levenshtein", 0, 4, 3, 2),
        O.Params("levenshtein", 1, 3, 2, 2),
        O.Params("needleman_wunsch", open=-2, extend=-2, byte_to_class=b2c, class_costs=costs),
        O.Params("needleman_wunsch", open=-4, extend=-1, byte_to_class=b2c, class_costs=costs),
        O.Params("smith_waterman", open=-2, extend=-2, byte_to_class=b2c, class_costs=costs),
        O.Params("smith_waterman", open=-3, extend=-1, byte_to_class=b2c, class_costs=costs),
    ]
    for params in all_params:
        longest = 4 if params.kind == "smith_waterman" else 5  # local alignment enumerates every substring pair
        symbols = np.frombuffer(b"abc", dtype=np.uint8)
        for _ in range(6):
            a = bytes(rng.choice(symbols, size=rng.integers(0, longest)))
            b = bytes(rng.choice(symbols, size=rng.integers(0, longest)))
            assert O.score(params, a, b) == _brute(params, a, b), (params.kind, params.open, params.extend, a, b)


#  Closed forms ----------------------------------------------------------------------------------------------------


def test_closed_forms_and_invariants():
    rng = np.random.default_rng(3)
    unit = O.Params("levenshtein")
    affine = O.Params("levenshtein", 0, 4, 3, 2)
    for _ in range(40):
        a = random_bytes(rng, DNA, 0, 60)
        b = random_bytes(rng, DNA, 0, 60)
        assert O.score(unit, a, a) == 0
        d = O.score(unit, a, b)
        assert abs(len(a) - len(b)) <= d <= max(len(a), len(b))
        assert d == O.score(unit, b, a)
        assert O.score(unit, a, b"") == len(a)
        assert O.score(affine, a, b"") == (3 + 2 * (len(a) - 1) if a else 0)
        # ASCII through the UTF-8 engine equals the byte distance.
        assert O.score(O.Params("levenshtein_utf8"), a, b) == d


def test_needleman_wunsch_with_unit_class_costs_is_negated_levenshtein():
    """test/similarities.py:290-319: classes = byte mod 32, costs 0 on the diagonal and -1 elsewhere, gaps -1."""
    rng = np.random.default_rng(4)
    b2c = (np.arange(256) % 32).astype(np.uint8)
    costs = np.full((32, 32), -1, dtype=np.int8)
    np.fill_diagonal(costs, 0)
    nw = O.Params("needleman_wunsch", open=-1, extend=-1, byte_to_class=b2c, class_costs=costs)
    for _ in range(30):
        a = random_bytes(rng, b"abcdefghijklmnopqrstuvwxyz", 0, 80)
        b = random_bytes(rng, b"abcdefghijklmnopqrstuvwxyz", 0, 80)
        assert O.score(nw, a, b) == -O.score(O.Params("levenshtein"), a, b)


def test_self_alignment_scores_sum_the_diagonal():
    """With positive diagonal costs dominating their rows and columns and non-positive gaps, NW(a, a) = SW(a, a)
    = sum of C[a_i][a_i]."""
    rng = np.random.default_rng(6)
    b2c, costs = protein_tables(rng)
    nw = O.Params("needleman_wunsch", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs)
    sw = O.Params("smith_waterman", open=-11, extend=-1, byte_to_class=b2c, class_costs=costs)
    for _ in range(20):
        a = random_bytes(rng, PROTEIN, 1, 200)
        expected = sum(int(costs[b2c[x], b2c[x]]) for x in a)
        assert O.score(nw, a, a) == expected
        assert O.score(sw, a, a) == expected
        assert O.score(sw, a, random_bytes(rng, PROTEIN, 0, 50)) >= 0


def test_utf8_decoding_table():
    """`sz_rune_decode_unchecked` (include/stringzilla/utf8_runes/serial.h:101-114) on valid and malformed input."""
    assert list(O.utf8_decode("aé中😀".encode())) == [0x61, 0xE9, 0x4E2D, 0x1F600]
    assert list(O.utf8_decode(b"\x80a")) == [0x80, 0x61]            # stray continuation byte: one rune
    assert list(O.utf8_decode(b"\xe4\xb8")) == [0x4E00]              # truncated tail reads missing bytes as 0
    assert list(O.utf8_decode(b"\xe0AB")) == [(0xE0 & 0x0F) << 12 | (0x41 & 0x3F) << 6 | (0x42 & 0x3F)]
    assert list(O.utf8_decode(b"")) == []
