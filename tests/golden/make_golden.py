"""Writes the committed fixtures of tests/golden/ from the CPU oracle.

generator_4x1024x4096.json : digests of the 4-piece circuit (4 pieces x 1024 nodes x 4096
                    wires, 95% in-piece, seed 12345, reference profile), its
                    drand48 draw count and its first node records.
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    c = O.generate(4, 1024, 4096, 95, 12345, O.PROFILE_REFERENCE)
    out = {
        "config": {"num_pieces": 4, "nodes_per_piece": 1024, "wires_per_piece": 4096,
                   "pct_wire_in_piece": 95, "random_seed": 12345, "profile": "reference"},
        "sha256": {n: hashlib.sha256(getattr(c, n).tobytes()).hexdigest()
                   for n in ("nodes", "wires", "locator")},
        "draws": int(c.draws),
        "first_nodes": {str(i): [float(x) for x in c.nodes[i].tolist()] for i in range(4)},
    }
    with open(os.path.join(HERE, "generator_4x1024x4096.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
