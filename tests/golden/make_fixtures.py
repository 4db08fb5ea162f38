"""Extract the reference's own known-answer values into JSON fixtures.

Run in a checkout that has the reference tree mounted (default
/root/reference); the GPU host never runs this.  It reads the reference test
sources as text and keeps only numbers (inputs and expected outputs):

* LGL points and weights, N = 2..10 and 20
  (tests/Unit/NumericalAlgorithms/Spectral/Test_LegendreGaussLobatto.cpp:34-143)
* LGL differentiation matrices, N = 2..6 and 10 (same file, :159-260)
* SpEC values of the GH right-hand side for the seeded random inputs
  (tests/Unit/Evolution/Systems/GeneralizedHarmonic/Test_DuDt.cpp:361-465)

Usage: python tests/golden/make_fixtures.py [reference_root]
"""
import ast
import json
import operator
import pathlib
import re
import sys

HERE = pathlib.Path(__file__).resolve().parent

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.USub: operator.neg, ast.UAdd: operator.pos}


def _eval(node):
    """Evaluate a numeric literal expression (numbers and + - * /)."""
    if isinstance(node, ast.Expression):
        return _eval(node.body)
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return float(node.value)
    if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
        return _OPS[type(node.op)](_eval(node.left), _eval(node.right))
    if isinstance(node, ast.UnaryOp) and type(node.op) in _OPS:
        return _OPS[type(node.op)](_eval(node.operand))
    raise ValueError(f"not a numeric expression: {ast.dump(node)}")


def num(expr):
    return _eval(ast.parse(expr.strip(), mode="eval"))


def _split_top(s):
    """Split on commas that are not nested in braces/parens."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def _datavector(expr):
    """DataVector{a, b, ...} optionally followed by / c."""
    m = re.match(r"\s*DataVector\s*\{(.*?)\}\s*(/\s*([0-9.eE+-]+))?\s*$", expr, re.S)
    if not m:
        raise ValueError(expr)
    vals = [num(v) for v in _split_top(m.group(1))]
    if m.group(3):
        vals = [v / float(m.group(3)) for v in vals]
    return vals


def lgl_tables(ref):
    src = (ref / "tests/Unit/NumericalAlgorithms/Spectral/"
           "Test_LegendreGaussLobatto.cpp").read_text()
    points = {}
    for m in re.finditer(r"test_points_and_weights\(\s*(\d+)\s*,(.*?)\);\s*\n", src, re.S):
        n = int(m.group(1))
        args = _split_top(m.group(2))
        if len(args) != 2:
            continue
        points[n] = {"x": _datavector(args[0]), "w": _datavector(args[1])}
    diff = {}
    for m in re.finditer(r"test_diff_matrix\(\s*(\d+)\s*,\s*Matrix\((.*?)\)\);", src, re.S):
        n = int(m.group(1))
        body = m.group(2).strip()
        rows = re.findall(r"\{([^{}]*)\}", body)
        mat = [[num(v) for v in _split_top(r)] for r in rows]
        diff[n] = mat
    return points, diff


def spec_values(ref):
    src = (ref / "tests/Unit/Evolution/Systems/GeneralizedHarmonic/Test_DuDt.cpp").read_text()
    start = src.index("void test_reference_impl_against_spec()")
    end = src.index("template <size_t Dim, typename Generator>", start)
    block = src[start:end]
    vals = []
    for m in re.finditer(
            r"CHECK\(dt_(psi|pi|phi)\.get\(([\d,\s]+)\)\[(\d)\]\s*==\s*(\w+)\(([-0-9.eE+]+)\)\)",
            block):
        var, idx, pt, approx_name, value = m.groups()
        vals.append({"var": var, "index": [int(i) for i in idx.split(",")],
                     "point": int(pt), "value": float(value),
                     "epsilon": 1e-13 if approx_name == "coarser_approx" else 1e-14})
    return vals


def main():
    ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    points, diff = lgl_tables(ref)
    spec = spec_values(ref)
    assert sorted(points) == [2, 3, 4, 5, 6, 7, 8, 9, 10, 20], sorted(points)
    assert sorted(diff) == [2, 3, 4, 5, 6, 10], sorted(diff)
    assert len(spec) == 100, len(spec)
    (HERE / "lgl_points_weights.json").write_text(json.dumps(
        {"source": "tests/Unit/NumericalAlgorithms/Spectral/Test_LegendreGaussLobatto.cpp:34-143",
         "tolerance": {"x": "approx (1e-14 rel, scale 1)", "w": 1e-8},
         "tables": {str(k): v for k, v in sorted(points.items())}}, indent=1))
    (HERE / "lgl_diff_matrix.json").write_text(json.dumps(
        {"source": "tests/Unit/NumericalAlgorithms/Spectral/Test_LegendreGaussLobatto.cpp:159-260",
         "tolerance": "approx with margin 1e-13",
         "tables": {str(k): v for k, v in sorted(diff.items())}}, indent=1))
    (HERE / "gh_spec_dudt.json").write_text(json.dumps(
        {"source": "tests/Unit/Evolution/Systems/GeneralizedHarmonic/Test_DuDt.cpp:361-465",
         "inputs": "std::mt19937 gen(1.), uniform_real_distribution<>(-10, 10), 2 points, "
                   "Test_DuDt.cpp:294-352",
         "values": spec}, indent=1))
    print("wrote", len(points), "LGL tables,", len(diff), "D matrices,", len(spec), "SpEC values")


if __name__ == "__main__":
    main()
