"""Write ``initial_config_constants.json``: the numeric constants of the reference's initial data.

Reads ``theories/nuclear_skyrme/initial_config.py`` of the reference as text and
parses it with ``ast`` (nothing of the reference is imported or executed); takes
the literal constants of

* ``profilefun`` (:342-442): the fit coefficients (a, b, c, d, e) per baryon
  number of the m0 < 1 branch (first the massless fit, then the massive one) and
  of the m0 >= 1 branch, and the two radius cut-offs;
* ``rational_map_value`` (:445-549): the named coefficients of each map and the
  regulariser added to denominators;
* ``cdiv`` (:229-250): the regulariser added to |den|^2.

Usage: python tests/golden/make_initial_config_constants.py [path/to/initial_config.py]
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

DEFAULT = Path("/root/reference/skyrmion_solver/theories/nuclear_skyrme/initial_config.py")
OUT = Path(__file__).resolve().parent / "initial_config_constants.json"


def _func(tree, name):
    for node in tree.body:
        if isinstance(node, ast.FunctionDef) and node.name == name:
            return node
    raise KeyError(name)


def _bint_chain(node):
    """(bint, body) pairs of an `if bint == k: ... elif bint == k2: ...` chain."""
    out = []
    while isinstance(node, ast.If):
        t = node.test
        assert isinstance(t, ast.Compare) and isinstance(t.left, ast.Name) and t.left.id == "bint"
        out.append((t.comparators[0].value, node.body))
        node = node.orelse[0] if len(node.orelse) == 1 else None
    return out


def _tuple_assigns(body):
    return [[ast.literal_eval(e) for e in st.value.elts] for st in body
            if isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Tuple)]


def main(path: Path):
    tree = ast.parse(path.read_text())
    prof = _func(tree, "profilefun")
    ifs = [st for st in prof.body if isinstance(st, ast.If)]
    cutoffs = {}
    for st in ifs:
        t = st.test
        if isinstance(t, ast.Compare) and isinstance(t.left, ast.Name) and t.left.id == "r":
            cutoffs["r_small" if isinstance(t.ops[0], ast.Lt) else "r_large"] = t.comparators[0].value
    m0_branch = [st for st in ifs if isinstance(st.test, ast.Compare) and isinstance(st.test.left, ast.Name)
                 and st.test.left.id == "m0"][0]
    massless, massive_m0lt1 = {}, {}
    for b, body in _bint_chain([s for s in m0_branch.body if isinstance(s, ast.If)][0]):
        first, second = _tuple_assigns(body)
        massless[b], massive_m0lt1[b] = first, second
    massive = {b: _tuple_assigns(body)[0]
               for b, body in _bint_chain([st for st in ifs if isinstance(st.test.left, ast.Name)
                                           and st.test.left.id == "bint"][0])}
    rmap = _func(tree, "rational_map_value")
    eps = [ast.literal_eval(st.value) for st in rmap.body if isinstance(st, ast.Assign)
           and isinstance(st.targets[0], ast.Name) and st.targets[0].id == "eps0"][0]
    maps = {}
    for st in rmap.body:
        if isinstance(st, ast.If):
            b = st.test.comparators[0].value
            named = {s.targets[0].id: ast.literal_eval(s.value) for s in st.body if isinstance(s, ast.Assign)
                     and isinstance(s.targets[0], ast.Name) and isinstance(s.value, ast.Constant)}
            if named:
                maps[b] = named
    cdiv = _func(tree, "cdiv")
    cdiv_eps = [n.value for n in ast.walk(cdiv) if isinstance(n, ast.Constant) and isinstance(n.value, float)]
    data = dict(source="theories/nuclear_skyrme/initial_config.py",
                profile_fit_massless={str(k): v for k, v in sorted(massless.items())},
                profile_fit_massive={str(k): v for k, v in sorted(massive.items())},
                profile_fit_massive_m0_below_1={str(k): v for k, v in sorted(massive_m0lt1.items())},
                profile_cutoffs=cutoffs, map_constants={str(k): v for k, v in sorted(maps.items())},
                map_eps=eps, cdiv_eps=cdiv_eps[0])
    OUT.write_text(json.dumps(data, indent=1) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else DEFAULT)
