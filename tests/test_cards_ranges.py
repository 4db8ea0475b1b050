"""Card encoding and range grammar against the reference's known answers
(crates/range-solver/src/card.rs:414-491, range.rs:770-970)."""
import numpy as np
import pytest

import poker_solver_rust_amd as P


def test_card_encoding():
    assert P.card_from_str("2c") == 0
    assert P.card_from_str("As") == 51
    assert P.card_from_str("3d") == 5
    assert P.card_from_str("4h") == 10
    with pytest.raises(P.PsrError):
        P.card_from_str("1c")
    with pytest.raises(P.PsrError):
        P.card_from_str("AS")   # suits are lower case only


def test_card_pair_index_exhaustive():
    k = 0
    for i in range(52):
        for j in range(i + 1, 52):
            assert P.card_pair_to_index(i, j) == k
            assert P.card_pair_to_index(j, i) == k
            assert P.index_to_card_pair(k) == (i, j)
            k += 1
    assert k == 1326
    # doc examples card.rs:48,59-60
    assert P.index_to_card_pair(0) == (0, 1)      # 2d2c
    assert P.index_to_card_pair(1) == (0, 2)      # 2h2c
    assert P.index_to_card_pair(1325) == (50, 51) # AsAh


def test_flop_from_str():
    assert P.flop_from_str("2c3d4h") == [0, 5, 10]
    assert P.flop_from_str("As Ah Ks") == [47, 50, 51]
    with pytest.raises(P.PsrError):
        P.flop_from_str("2c3d4h5s")
    with pytest.raises(P.PsrError):
        P.flop_from_str("2c2c3d")


def R(s):
    return P.Range.parse(s).data


@pytest.mark.parametrize("a,b", [
    ("88+", "AA,KK,QQ,JJ,TT,99,88"),
    ("8s8h+", "AhAs,KhKs,QhQs,JhJs,ThTs,9h9s,8h8s"),
    ("98s+", "AKs,KQs,QJs,JTs,T9s,98s"),
    ("A8o+", "AKo,AQo,AJo,ATo,A9o,A8o"),
    ("88-55", "88,77,66,55"),
    ("98s-65s", "98s,87s,76s,65s"),
    ("AQo-86o", "AQo,KJo,QTo,J9o,T8o,97o,86o"),
    ("K5-K2", "K5,K4,K3,K2"),
    ("AhAs-QhQs,JJ", "JJ,AhAs,KhKs,QhQs"),
    ("  AK  ", "AK"),
    ("K9s: .67", "K9s:.67"),
    ("88+, AQ+", "88+,AQ+"),
    ("98s - 65s: 0.25", "98s-65s:0.25"),
])
def test_range_equivalences(a, b):
    assert np.array_equal(R(a), R(b))


@pytest.mark.parametrize("s", ["", "AK,", "ak", "K9s:.67", "88+:1.", "98s-65s:0.25", "AcKh", "8h8s+:.67",
                               "9d8d-6d5d:0.25", "85s:0.5"])
def test_range_accepts(s):
    P.Range.parse(s)


@pytest.mark.parametrize("s", ["AK,,", "89", "AAo", "AQo:1.1", "AQo-AQo", "AQo-86s", "AQo-KQo", "K2-K5",
                               "AhAs-QsQh", "AKQ", "AK+-AJ", "K9s.67", "88+:2.0", "98s-21s"])
def test_range_rejects(s):
    with pytest.raises(P.PsrError):
        P.Range.parse(s)


def test_range_weights():
    r = P.Range.parse("85s:0.5")
    # suited 8-5 (ranks 6, 3): all 4 suits at 0.5, offsuit at 0
    for s in range(4):
        assert r.get_weight_by_cards(4 * 6 + s, 4 * 3 + s) == 0.5
        for t in range(4):
            if t != s:
                assert r.get_weight_by_cards(4 * 6 + s, 4 * 3 + t) == 0.0
    r = P.Range.parse("AA,AKs,AKo")
    assert r.data.sum() == 6 + 16
    r = P.Range.parse("TT+")
    assert r.get_weight_by_cards(4 * 8, 4 * 8 + 1) == 1.0 and r.get_weight_by_cards(4 * 7, 4 * 7 + 1) == 0.0
    # later groups are overridden by earlier ones (parsed right to left)
    r = P.Range.parse("AA:0.25,AA:0.75")
    assert r.get_weight_by_cards(48, 49) == 0.25
