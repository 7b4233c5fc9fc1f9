"""Host-only drop-ins of the HIP library (no GPU needed): getHash and
qryGetReverseComplement against the reference's known answers and the
oracle."""
import json
import os
import random

import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_vectors.json")))


@pytest.fixture(scope="module")
def S():
    from swift_amd import _lib, build
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    from swift_amd import search
    return search


def test_get_hash_closed_form(S, orc):
    assert S.get_hash(b"ACG") == 36  # worked example, src/preprocess.cu:79-86
    rnd = random.Random(1)
    for _ in range(500):
        k = rnd.randint(1, 15)
        s = bytes(rnd.choice(b"ACGTacgtNn*xR") for _ in range(k))
        want = sum(4 ** i * {"C": 1, "G": 2, "T": 3}.get(chr(c).upper(), 0)
                   for i, c in enumerate(s))
        assert S.get_hash(s) == want == orc.get_hash(s, k)


@pytest.mark.parametrize("case", GOLD["revcomp"], ids=lambda c: c["seq"][:20])
def test_revcomp_known(S, case):
    assert S.revcomp(case["seq"].encode()) == case["rc"].encode()


def test_revcomp_vs_oracle(S, orc):
    rnd = random.Random(2)
    for _ in range(200):
        s = bytes(rnd.choice(b"ACGTacgtNn-*") for _ in range(rnd.randint(0, 300)))
        assert S.revcomp(s) == orc.revcomp(s)
