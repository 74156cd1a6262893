"""Writes tests/golden/rng_vectors.json: the first draws and states of the
renderer's PCG stream for a few seeds, as produced by the CPU oracle
(regression fixture; the generator's definition is pinned separately by the
splitmix64 known answer and statistical checks in tests/test_oracle_pins.py)."""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
import oracle_bindings as ob  # noqa: E402

SEEDS = [0, 1, 12345, 0xDEADBEEFCAFEF00D, 2**64 - 1]


def main():
    out = []
    for s in SEEDS:
        draws, states = ob.rng_draws(s, 16)
        out.append({"seed": s, "state0": int(states[0]), "states": [int(v) for v in states[1:]],
                    "draws_bits": [int(v) for v in draws.view("uint32")]})
    path = Path(__file__).with_name("rng_vectors.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
