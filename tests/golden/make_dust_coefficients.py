"""Write tests/golden/dust_coefficients.json: the numeric tables of the
reference's Gordon et al. (2023) curve, read from the source text of
rubix/spectra/dust/extinction_models.py with Python's ``ast`` module (the
file is parsed, never imported or run).

    python tests/golden/make_dust_coefficients.py [path/to/extinction_models.py]

Extracted from ``Gordon23.evaluate``: the list literals ``ir_a``, ``opt_a``,
``opt_b``, ``optir_waves``, ``uvopt_waves``; the keyword arguments of the
``PowerLaw1d`` calls (infrared b term); the numeric positional arguments of
the ``FM90`` calls (ultraviolet a and b terms, in source order); and the
numeric bounds of the ``ir_mask`` / ``opt_mask`` / ``uv_mask`` comparisons.
"""
from __future__ import annotations

import ast
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = "/root/reference/rubix/spectra/dust/extinction_models.py"


def _num(node):
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return float(node.value)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        v = _num(node.operand)
        return None if v is None else -v
    return None


def extract(path):
    tree = ast.parse(open(path).read())
    cls = next(n for n in tree.body if isinstance(n, ast.ClassDef) and n.name == "Gordon23")
    ev = next(n for n in cls.body if isinstance(n, ast.FunctionDef) and n.name == "evaluate")
    out = {"source": "rubix/spectra/dust/extinction_models.py (Gordon23.evaluate)"}
    fm90, powerlaw = [], []
    for node in ast.walk(ev):
        if isinstance(node, ast.Assign) and len(node.targets) == 1 \
                and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name in ("ir_a", "opt_a", "opt_b", "optir_waves", "uvopt_waves") \
                    and isinstance(node.value, ast.List):
                out[name] = [_num(e) for e in node.value.elts]
            elif name in ("ir_mask", "opt_mask", "uv_mask"):
                out[name] = [_num(c) for c in ast.walk(node.value)
                             if isinstance(c, (ast.Constant, ast.UnaryOp)) and _num(c) is not None]
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
            if node.func.id == "FM90":
                args = [_num(a) for a in node.args[1:]]
                if args not in fm90:
                    fm90.append(args)
            elif node.func.id == "PowerLaw1d":
                kw = {k.arg: _num(k.value) for k in node.keywords if k.arg != "x"}
                if kw not in powerlaw:
                    powerlaw.append(kw)
    out["fm90_calls"] = fm90
    out["powerlaw_calls"] = powerlaw
    return out


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    data = extract(path)
    dst = os.path.join(HERE, "dust_coefficients.json")
    with open(dst, "w") as f:
        json.dump(data, f, indent=1)
    print(f"wrote {dst}")


if __name__ == "__main__":
    main()
