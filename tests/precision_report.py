"""Writes the per-iteration precision table (markdown) of tests/precision_common.py to stdout.

    python tests/precision_report.py > profiles/precision_panel32_tv.md
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import precision_common as pc  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    rows = pc.study(oracle)
    print("# Storage precision vs the FP64 oracle: 32^3 volume, 32x32-crystal panels, 20 MLEM iterations + TV")
    print()
    print("Command: `python tests/precision_report.py` (one MI355X).  " + pc.__doc__.strip().splitlines()[2])
    print()
    labels = list(rows)
    print("| iteration | " + " | ".join(f"{l}: image rel L2 / LL rel" for l in labels) + " |")
    print("|---|" + "---|" * len(labels))
    for k in range(pc.ITERS):
        cells = [f"{rows[l][0][k]:.2e} / {rows[l][1][k]:.1e}" for l in labels]
        print(f"| {k + 1} | " + " | ".join(cells) + " |")


if __name__ == "__main__":
    main()
