"""Transcribes the prepare_bonds test data of the reference into
tests/golden/topology_goldens.json: the literal arrays of the `dppc_single`
and `alanine_octapeptide` fixtures (test/conftest.py:34-200: indices,
molecules, bonds, names and the bond/angle/dihedral types of CONF) are read
from the file text (nothing is imported or run), and the expected term lists
of test/test_force.py:18-45, 103-127, 189-212 are copied by hand below.  Run
in a container that has the reference at /root/reference; the JSON it
writes is the committed fixture."""
import ast
import json
import os

REF = "/root/reference/test/conftest.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "topology_goldens.json")


def _literal(node):
    """Literal value of an array constructor / namedtuple call / range()."""
    if isinstance(node, ast.Call):
        fn = node.func.attr if isinstance(node.func, ast.Attribute) else getattr(node.func, "id", "")
        if fn in ("array",):
            return _literal(node.args[0])
        if fn == "range":
            return list(range(*[_literal(a) for a in node.args]))
        return [_literal(a) for a in node.args]          # Bond(...), Angle(...), Dihedral(...)
    if isinstance(node, ast.ListComp):                    # [0 for _ in range(16)] / [1 for _ in range(5)]
        elt = _literal(node.elt)
        rng = _literal(node.generators[0].iter)
        return [elt for _ in rng]
    if isinstance(node, (ast.List, ast.Tuple)):
        return [_literal(e) for e in node.elts]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_literal(node.operand)
    if isinstance(node, ast.Constant):
        v = node.value
        return v.decode() if isinstance(v, bytes) else v
    raise ValueError(ast.dump(node)[:120])


def fixture(tree, name):
    fn = next(n for n in ast.walk(tree) if isinstance(n, ast.FunctionDef) and n.name == name)
    out = {}
    for node in fn.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            t = node.targets[0]
            if isinstance(t, ast.Name) and t.id in ("indices", "molecules", "bonds", "names"):
                out[t.id] = _literal(node.value)
            if isinstance(t, ast.Subscript) and isinstance(t.value, ast.Name) and t.value.id == "CONF":
                key = t.slice.value
                if key in ("bond_2", "bond_3", "bond_4"):
                    out[key] = _literal(node.value)
    return out


tree = ast.parse(open(REF).read())
dppc = fixture(tree, "dppc_single")
ala = fixture(tree, "alanine_octapeptide")
data = {
    "_source": ("test/conftest.py:34-200 (dppc_single, alanine_octapeptide) and the expected term lists of "
                "test/test_force.py:18-45, 103-127, 189-212"),
    "dppc_single": dict(dppc, expected_bonds=[
        [0, 1, 0.47, 1250.0], [1, 2, 0.47, 1250.0], [2, 3, 0.37, 1250.0], [2, 4, 0.47, 1250.0],
        [3, 8, 0.47, 1250.0], [4, 5, 0.47, 1250.0], [5, 6, 0.47, 1250.0], [6, 7, 0.47, 1250.0],
        [8, 9, 0.47, 1250.0], [9, 10, 0.47, 1250.0], [10, 11, 0.47, 1250.0]],
        expected_angles=[
        [1, 2, 3, 120.0, 25.0], [1, 2, 4, 180.0, 25.0], [3, 8, 9, 180.0, 25.0], [2, 4, 5, 180.0, 25.0],
        [4, 5, 6, 180.0, 25.0], [5, 6, 7, 180.0, 25.0], [8, 9, 10, 180.0, 25.0], [9, 10, 11, 180.0, 25.0]]),
    "alanine_octapeptide": dict(ala, expected_dihedrals=[
        [0, 2, 4, 6, 0], [2, 4, 6, 8, 0], [4, 6, 8, 10, 0], [6, 8, 10, 12, 0], [8, 10, 12, 14, 0]],
        expected_dihedral_coeffs=[[1, 1, 1, 1, 1], [0, 0, 0, 0, 0]]),
}
with open(OUT, "w") as fh:
    json.dump(data, fh, indent=1)
print("wrote", OUT)
