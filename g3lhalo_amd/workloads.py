"""Seeded request streams of the serving and fit-sweep workloads, shared by
the benchmark and the tests (host-side description only; the requests run
through the C ABI).

A serving request is what one call of the reference's drivers or fit loop
asks for (src/calculateNNMap.cu:111-154, src/calculateNNMap_twoPop.cu:137-232,
src/NNMap_Model.cu:382-415): one HOD set, a list of aperture triplets and the
terms wanted.  One request in five is a two-population request (red/blue lens
samples: two prior HODs), which, as calculateNNMap_twoPop.x does, asks for
every NNMap term at the three type pairs (1,2), (1,1) and (2,2).
"""
from __future__ import annotations

import numpy as np

from . import tables as T

TYPE_PAIRS_TWO_POP = ((1, 2), (1, 1), (2, 2))


def serve_stream(n: int, seed: int = 4):
    """n seeded requests: each carries an HOD vector from the uniform prior,
    1-200 triplets of the 550-triplet grid and a random non-empty mask over
    NNMap/NMapMap 1h/2h/3h; one request in five is a two-population NNMap
    request (two prior HODs, all three type pairs)."""
    rng = np.random.default_rng(seed)
    trips = T.triplets_550()
    hods = T.hod_prior_draws(2 * n, seed=seed + 1)
    out = []
    for i in range(n):
        two = rng.random() < 0.2
        k = int(rng.integers(1, 201))
        sel = trips[rng.choice(len(trips), size=k, replace=False)]
        while True:
            mask = int(rng.integers(1, 64))
            terms = [(0, t) for t in (1, 2, 3) if mask >> (t - 1) & 1]
            if not two:
                terms += [(1, t) for t in (1, 2, 3) if mask >> (t + 2) & 1]
            if terms:
                break
        out.append({"hod1": hods[2 * i], "hod2": hods[2 * i + 1] if two else hods[2 * i],
                    "two": two, "triplets": sel, "terms": terms})
    return out


def request_integrals(q: dict):
    """The integrals of one stream request, in submission order:
    (stat, term, theta [arcmin], type1, type2)."""
    pairs = TYPE_PAIRS_TWO_POP if q["two"] else ((1, 2),)
    out = []
    for th in q["triplets"]:
        for stat, term in q["terms"]:
            for t1, t2 in (pairs if stat == 0 else ((1, 1),)):
                out.append((stat, term, th, t1, t2))
    return out
