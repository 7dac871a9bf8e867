"""Loaders of the committed golden fixtures (tests/golden/*.json)."""
import json
import os

import numpy as np

from falm_amd import _abi as abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def dot_case():
    """(domain, a, b, expected) of the reference's test/dot.cpp set-up."""
    d = _load("dot_case.json")
    dom = abi.Domain.make(d["shape"], gc=d["gc"])
    n = d["shape"][0]
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    a = (300 - (i - 7) ** 2 - (j - 2) ** 2 - (k - 5) ** 2).astype(np.float64).ravel()
    b = (150 - i ** 2 - (j - 1) ** 2 - (k - 1) ** 2).astype(np.float64).ravel()
    return dom, a, b, d


def heat1d():
    return _load("heat1d.json")
