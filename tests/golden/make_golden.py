"""Writes the golden fixtures of tests/golden/ from the known-answer set-ups
held by the reference's own tests (restated here, not imported):

  dot_case.json   test/dot.cpp:9-41 + test/dot.py:1-29 — two integer
                  quadratics on a 12^3 box with one guide layer; sum(a*b),
                  sum(a*a) over the 10^3 interior and max|a|, max|b| there,
                  computed in exact integer arithmetic.
  heat1d.json     test/heat1d.cpp:14-79 — 1-D Poisson on N = 100 cells with
                  ghost-Dirichlet rows T_W = 0, T_E = 100; its exact discrete
                  solution is T_i = 100*(i + 1/2)/N.

    python tests/golden/make_golden.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def dot_case():
    sab = saa = 0
    maxa = maxb = 0
    for k in range(12):
        for j in range(12):
            for i in range(12):
                a = 300 - (i - 7) ** 2 - (j - 2) ** 2 - (k - 5) ** 2
                b = 150 - i ** 2 - (j - 1) ** 2 - (k - 1) ** 2
                if 0 < i < 11 and 0 < j < 11 and 0 < k < 11:
                    sab += a * b
                    saa += a * a
                    maxa = max(maxa, abs(a))
                    maxb = max(maxb, abs(b))
    return {"shape": [12, 12, 12], "gc": 1, "sumab": sab, "sumaa": saa, "maxa": maxa, "maxb": maxb,
            "source": "test/dot.cpp:9-41, test/dot.py:1-29"}


def heat1d():
    n, tw, te = 100, 0.0, 100.0
    return {"n": n, "TW": tw, "TE": te, "L": 1.0,
            "T": [te * (i + 0.5) / n for i in range(n)],
            "source": "test/heat1d.cpp:14-79"}


if __name__ == "__main__":
    for name, fn in (("dot_case.json", dot_case), ("heat1d.json", heat1d)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, indent=1)
    print("ok")
