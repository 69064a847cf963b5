"""Reader of the parts of a semtex session file (FEML) that set up the
Helmholtz path: TOKENS (numeric expressions), NODES, ELEMENTS (<Q>),
SURFACES (<B> group / <P> periodic), GROUPS and BCS (src/feml.cpp,
src/mesh.cpp:88-239 reads NODES/ELEMENTS, :350-475 SURFACES, :1329-1373
BCS).  Curved sides, USER functions and BC expressions are not evaluated:
the numbering and the straight-sided geometry need none of them.

    S = read_session(path)
    topo = S.topology("u")     # -> arrays for sx_enumerate (essential flags of field u)
"""
from __future__ import annotations

import ast
import math
import operator
import re

DEFAULT_TOKENS = {"PI": math.pi, "TWOPI": 2 * math.pi, "N_P": 5.0, "N_Z": 1.0,
                  "BETA": 1.0, "N_TIME": 2.0, "KINVIS": 1.0, "D_T": 0.01}
_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "sin": math.sin,
          "cos": math.cos, "tan": math.tan, "atan": math.atan, "sinh": math.sinh,
          "cosh": math.cosh, "tanh": math.tanh, "abs": abs, "pow": math.pow}
_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.USub: operator.neg,
        ast.UAdd: operator.pos}


def evaluate(expr, names):
    """Arithmetic expression of the token language (+ - * / ^, functions,
    named tokens), evaluated without exec."""
    tree = ast.parse(expr.replace("^", "**"), mode="eval")

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name):
            if n.id in names:
                return float(names[n.id])
            raise KeyError(n.id)
        if isinstance(n, ast.BinOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and type(n.op) in _OPS:
            return _OPS[type(n.op)](ev(n.operand))
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS:
            return float(_FUNCS[n.func.id](*[ev(a) for a in n.args]))
        raise ValueError(f"unsupported expression: {expr}")
    return ev(tree)


def _sections(text):
    out = {}
    for m in re.finditer(r"<(\w+)([^>]*)>(.*?)</\1>", text, flags=re.S):
        name, attrs, body = m.group(1).upper(), m.group(2), m.group(3)
        if name in out:
            continue
        num = re.search(r"NUMBER\s*=\s*(\d+)", attrs)
        lines = [ln.split("#", 1)[0].strip() for ln in body.splitlines()]
        out[name] = (int(num.group(1)) if num else None, [ln for ln in lines if ln])
    return out


class Session:
    def __init__(self, path):
        text = open(path).read()
        self.path = path
        S = _sections(text)
        self.tokens = dict(DEFAULT_TOKENS)
        for ln in S.get("TOKENS", (None, []))[1]:
            for part in ln.split(";"):
                if "=" in part:
                    k, v = part.split("=", 1)
                    try:
                        self.tokens[k.strip()] = evaluate(v.strip(), self.tokens)
                    except (KeyError, ValueError, ZeroDivisionError):
                        pass
        self.np = int(self.tokens["N_P"])
        nn, lines = S["NODES"]
        self.nodes = {}
        for ln in lines[:nn]:
            t = ln.split()
            self.nodes[int(t[0]) - 1] = (float(t[1]), float(t[2]))
        ne, lines = S["ELEMENTS"]
        self.elements = [None] * ne
        for ln in lines[:ne]:
            t = ln.replace("<", " <").replace(">", "> ").split()
            if t[1].upper() != "<Q>":
                raise ValueError(f"{path}: only <Q> elements are read ({ln})")
            self.elements[int(t[0]) - 1] = [int(v) - 1 for v in t[2:6]]
        ns, lines = S.get("SURFACES", (0, []))
        self.surfaces = []
        for ln in lines[:ns]:
            t = ln.replace("<", " <").replace(">", "> ").split()
            e, s, tag = int(t[1]) - 1, int(t[2]) - 1, t[3].upper()
            if tag == "<B>":
                self.surfaces.append((e, s, "B", t[4], None))
            elif tag == "<P>":
                self.surfaces.append((e, s, "P", int(t[4]) - 1, int(t[5]) - 1))
            else:
                raise ValueError(f"{path}: surface tag {tag}")
        self.bcs = {}
        nb, lines = S.get("BCS", (0, []))
        i = 0
        for _ in range(nb or 0):
            t = lines[i].split()
            grp, nf = t[1], int(t[2])
            kinds = {}
            for ln in lines[i + 1:i + 1 + nf]:
                m = re.match(r"<(\w)>\s*(\w)", ln)
                if m:
                    kinds[m.group(2)] = m.group(1).upper()
            self.bcs[grp] = kinds
            i += 1 + nf

    def match_bc(self, group, field, kind):
        """Mesh::matchBC: the BC of `field` on group `group` has tag `kind`."""
        return self.bcs.get(group, {}).get(field.lower()) == kind

    def essential(self, group, field):
        """Mesh::buildMask's test for a boundary side: <D>, or <A> for the
        axis-essential names U v w P C, or <O> for the pressure."""
        return (self.match_bc(group, field, "D") or
                (field in "UvwPC" and self.match_bc(group, field, "A")) or
                (field in "pP" and self.match_bc(group, field, "O")))

    def topology(self, field):
        """(elem_nodes [nel][4], nnode, surfaces [(elmt, side, kind, a, b)])
        for sx_enumerate / the oracle: kind 0 boundary (a = essential for
        `field`), kind 1 periodic (a, b = mate element and side)."""
        surf = []
        for (e, s, k, a, b) in self.surfaces:
            if k == "B":
                surf.append((e, s, 0, int(self.essential(a, field)), 0))
            else:
                surf.append((e, s, 1, a, b))
        return [list(v) for v in self.elements], len(self.nodes), surf


def read_session(path):
    return Session(path)
