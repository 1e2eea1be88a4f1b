"""Copies two of the reference's own forming decks (data written by its precomp
package, python/precomp/fea/deck.py:206-231) into tests/golden/decks/:

  kin_h0833.json   benchmarks/physics_audit/challenge/cases/kin_h0833/deck/deck.json
                   (48 x 48 x 1 incompatible-mode Hex8, 2 x 2 x 5 rule, logarithmic
                   strain, Hill48 + two Chaboche backstresses + Voce: the fine-mesh
                   decks of the reference's element cost figures)
  homogeneous_uniaxial.json
                   benchmarks/physics_audit/calculix/homogeneous/uniaxial/sparlab/deck.json
                   (one standard Hex8, J2 + Voce)

Run from the repository root where /root/reference is mounted:
python tests/golden/copy_decks.py
"""
import json
import os

SRC = "/root/reference/benchmarks/physics_audit"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "decks")
DECKS = {"kin_h0833.json": "challenge/cases/kin_h0833/deck/deck.json",
         "homogeneous_uniaxial.json": "calculix/homogeneous/uniaxial/sparlab/deck.json"}


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, rel in DECKS.items():
        with open(os.path.join(SRC, rel)) as f:
            deck = json.load(f)
        with open(os.path.join(OUT, name), "w") as f:
            json.dump(deck, f, indent=1, sort_keys=True)
        print(f"wrote {os.path.join(OUT, name)}")


if __name__ == "__main__":
    main()
