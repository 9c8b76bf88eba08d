"""Philox4x32-10 known answers (CPU): the oracle's generator, which the
kernel restates, against the Random123 KAT vectors of Salmon et al. and
against rocRAND's own Philox rounds."""
import os
import random

import pytest

from oracle import oracle as O

# (counter, key, expected) from the Random123 distribution's kat_vectors
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_philox_random123_kat(ctr, key, expected):
    assert O.philox(ctr, key) == expected


@pytest.mark.skipif(not os.path.exists(O.ROCRAND_XCHECK), reason="rocRAND harness not built")
def test_philox_matches_rocrand_rounds():
    rng = random.Random(7)
    for _ in range(200):
        c = tuple(rng.getrandbits(32) for _ in range(4))
        k = tuple(rng.getrandbits(32) for _ in range(2))
        assert O.philox(c, k) == O.rocrand_philox(c, k)
