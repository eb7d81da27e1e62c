"""Write the reference's tabulated quadrature / nodal-point data as golden fixtures.

The reference's tables are read as text and parsed with ``ast`` (nothing of the
reference is imported or executed).  Output: ``reference_tables.npz`` beside this
script, holding

* ``tet_quad_<C>``  : QuadraturePointsWeightsTet(C, Opt=3) rows [x, y, z, w]
  (Florence/QuadratureRules/QuadraturePointsWeightsTet.py:32-265)
* ``fekete_tet_<C>``: FeketePointsTet(C) rows [x, y, z]
  (Florence/QuadratureRules/FeketePointsTet.py:3-80+)
* ``gll_<n>``       : GaussLobattoQuadrature(n) rows [z, w]
  (Florence/QuadratureRules/NumericIntegrator.py:93-135)
* ``hex_face_numbering_<C>``: NodeArrangementHex(C)[0] face tables
  (Florence/MeshGeneration/NodeArrangement.py:216-260)

Run from the repository root:  python tests/golden/extract_reference_tables.py
"""
import ast
import os
import sys

import numpy as np

REF = "/root/reference/Florence"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.npz")


def _array_literals(node):
    """Yield every ``np.array(<literal>)`` argument found below ``node``."""
    for sub in ast.walk(node):
        if isinstance(sub, ast.Call) and getattr(sub.func, "attr", None) == "array" and sub.args:
            try:
                yield np.array(ast.literal_eval(sub.args[0]), dtype=np.float64)
            except ValueError:
                continue


def _branches(func, var):
    """Map ``var==k`` (or ``var==k or var==m``) branch values -> body statements."""
    out = {}

    def visit_if(node):
        test = node.test
        keys = []
        tests = test.values if isinstance(test, ast.BoolOp) else [test]
        for t in tests:
            if (isinstance(t, ast.Compare) and isinstance(t.left, ast.Name) and t.left.id == var
                    and isinstance(t.ops[0], ast.Eq) and isinstance(t.comparators[0], ast.Constant)):
                keys.append(t.comparators[0].value)
        for k in keys:
            out.setdefault(k, node.body)
        for o in node.orelse:
            if isinstance(o, ast.If):
                visit_if(o)

    for sub in ast.walk(func):
        if isinstance(sub, ast.If):
            visit_if(sub)
    return out


def _function(path, name):
    tree = ast.parse(open(path).read())
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == name:
            return node
    raise KeyError(name)


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present; fixtures are already committed")
    data = {}
    f = _function(f"{REF}/QuadratureRules/QuadraturePointsWeightsTet.py", "QuadraturePointsWeightsTet")
    for C, body in _branches(f, "C").items():
        arrs = [a for stmt in body for a in _array_literals(stmt) if a.ndim == 2 and a.shape[1] == 4]
        if arrs:
            data[f"tet_quad_{C}"] = arrs[-1]
    f = _function(f"{REF}/QuadratureRules/FeketePointsTet.py", "FeketePointsTet")
    for C, body in _branches(f, "C").items():
        if C > 4:
            continue
        arrs = [a for stmt in body for a in _array_literals(stmt) if a.ndim == 2 and a.shape[1] == 3]
        if arrs:
            data[f"fekete_tet_{C}"] = arrs[-1]
    f = _function(f"{REF}/QuadratureRules/NumericIntegrator.py", "GaussLobattoQuadrature")
    for C, body in _branches(f, "C").items():
        if C > 10:
            continue
        arrs = [a for stmt in body for a in _array_literals(stmt) if a.ndim == 2 and a.shape[1] == 2]
        if arrs:
            data[f"gll_{C}"] = arrs[-1]
    f = _function(f"{REF}/MeshGeneration/NodeArrangement.py", "NodeArrangementHex")
    for C, body in _branches(f, "C").items():
        arrs = [a for stmt in body for a in _array_literals(stmt) if a.ndim == 2]
        if arrs:
            data[f"hex_face_numbering_{C}"] = arrs[0].astype(np.int64)
    np.savez_compressed(OUT, **data)
    for k in sorted(data):
        print(k, data[k].shape)


if __name__ == "__main__":
    main()
