"""Writes tests/golden/rng_golden.json from oracle/_ref/ref_rng.

ref_rng is compiled from the reference's own ModelG/xoroshiro128plus.h and
drives it through libstdc++'s seed_seq / generate_canonical /
uniform_real_distribution / normal_distribution / shuffle exactly as
ModelG/NoiseGenerator.h does.  The JSON keeps the inputs and the outputs so
that the oracle can be checked on machines without the reference tree.

    make -C oracle && python tests/golden/make_rng_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from oracle import oracle  # noqa: E402

ENGINE_SEEDS = [0, 1, 42, 2934951935, 2**64 - 1, 0x123456789ABCDEF]
SEEDSEQ_BASES = [0, 1, 10, 42, 2021, 122335456, -1, 2**31 - 1]
SEEDSEQ_N = [1, 2, 3, 5, 8, 40]
NOISE = [(42, "uuuvvvnnnnuvnvnnnvu"), (2021, "uuuu"), (10, "nnnnnnnnvvvvuuuu"),
         (122335456, "vnvnvnvnuu")]
SHUFFLE_STATES = [(1, 2), (12345, 67890), (2**63, 3), (12982847866522366295, 7775382490793364087)]
CANON = [2**64 - 1, 2**64 - 512, 2**64 - 1024, 2**64 - 1025, 2**64 - 2048, 0, 1,
         2**53, 2**53 + 1, 2**63 + 1, 0xDEADBEEFCAFEBABE]
STATE = [(2021, 2), (42, 0), (42, 1000)]


def main():
    g = {"engine": [], "seedseq": [], "noise": [], "shuffle": [], "canon": [], "state": []}
    for s in ENGINE_SEEDS:
        g["engine"].append({"seed": s, "out": [int(x) for x in oracle.ref_rng("engine", s, 12)]})
    for b in SEEDSEQ_BASES:
        for n in SEEDSEQ_N:
            g["seedseq"].append({"base": b, "n": n,
                                 "out": [int(x) for x in oracle.ref_rng("seedseq", b, n)]})
    for b, pat in NOISE:
        g["noise"].append({"base": b, "pattern": pat, "out": oracle.ref_rng("noise", b, pat)})
    for s0, s1 in SHUFFLE_STATES:
        rows = oracle.ref_rng("shuffle", s0, s1, 40)
        g["shuffle"].append({"state": [s0, s1], "out": [[int(v) for v in r.split()] for r in rows]})
    g["canon"] = {"in": CANON, "out": oracle.ref_rng("canon", *CANON)}
    for b, n in STATE:
        g["state"].append({"base": b, "ndraws": n, "out": oracle.ref_rng("state", b, n)[0]})
    with open(os.path.join(HERE, "rng_golden.json"), "w") as f:
        json.dump(g, f, indent=1)


if __name__ == "__main__":
    main()
