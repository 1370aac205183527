"""Ansatz-name resolution shared by the theory parameter modules.

Same accepted names and error as the reference
(``theories/nuclear_skyrme/params.py:144-150``).
"""

from __future__ import annotations


def resolve_ansatz(name: str | None):
    ans = (name or "hedgehog").lower().strip()
    hedgehog = ans == "hedgehog"
    rational_map = ans in ("rational_map", "rational map", "rational-map")
    uniform = ans == "uniform"
    if not (hedgehog or rational_map or uniform):
        raise ValueError("Unknown ansatz '{0}'. Valid: hedgehog, rational_map, uniform".format(name))
    return hedgehog, rational_map, uniform
