"""Writes tests/golden/reference_cases.json: every known answer of the blazert primitive test suite.

Sources (read as text, never compiled or run): test/test_primitives/test_trimesh.cpp, test_sphere.cpp,
test_plane.cpp and test_cylinder.cpp of the blazert repository, with the mesh helpers of test/test_helpers.h.
Every assert_* call of those files becomes one case:

* assert_intersect_primitive_hit  -> "intersect_primitive": primitive 0 tested alone (assert_helper.h:123-146)
* assert_traverse_bvh_hit         -> "traverse": BVH built and traversed (assert_helper.h:212-228)
* assert_traverse_bvh_hit_trimesh_distance -> "traverse_distance": hit flag and distance only (:178-190)
* assert_bounding_box / assert_bounding_box_collection / assert_primitive_center / assert_distance_to_surface
  -> "bounding_box", "bounding_box_collection", "primitive_center", "distance_to_surface"
* assert_traverse_bvh_hit_trimesh_precision is a WARN_MESSAGE in the reference; those loops are counted, not kept.

doctest semantics: a TEST_CASE body runs once per leaf SUBCASE path, so the state an assert sees is what the
statements on its path (outside sibling SUBCASEs) built. The walker below copies the state on entering a SUBCASE
and drops it on leaving, and evaluates the statements with C++ rules (integer division for integer operands,
static_cast, the test helpers' arbitraryRotationMatrix and pi<T>). Values are computed for T = double; the
FP32 tests round them to float. Unknown statements stop the script, so nothing is silently skipped.
"""
from __future__ import annotations

import copy
import json
import math
import os
import re
import sys

REF = "/root/reference"
FILES = ["test/test_primitives/test_trimesh.cpp", "test/test_primitives/test_sphere.cpp",
         "test/test_primitives/test_plane.cpp", "test/test_primitives/test_cylinder.cpp"]
HELPERS = "test/test_helpers.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_cases.json")

DBL_MAX = sys.float_info.max
DBL_MIN = sys.float_info.min
DBL_EPS = sys.float_info.epsilon
IDENTITY = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]


# ------------------------------------------------------------------------------------------ source scanning
def strip_comments(text):
    out, i, n = [], 0, len(text)
    while i < n:
        if text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i)
            out.append(" " * (j + 2 - i))  # keep line numbers through newlines below
            out[-1] = "\n" * text[i:j + 2].count("\n")
            i = j + 2
        elif text[i] == '"':
            j = i + 1
            while text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text[i] == "#" and (i == 0 or text[i - 1] == "\n"):
            j = text.find("\n", i)
            i = n if j < 0 else j
        else:
            out.append(text[i])
            i += 1
    return "".join(out)


def is_block_head(head):
    """A '{' opens a block after a parenthesised head (SUBCASE, for, a function signature) or else/do; after a
    declarator name it starts a brace initializer."""
    return head in ("", "do", "else") or re.sub(r"\s*(const|noexcept)\s*$", "", head).endswith(")")


def parse_blocks(text):
    """Splits C++ source into a tree: ("stmt", text, line) and ("block", head, line, children)."""
    root = []
    stack = [root]
    buf, buf_line, line = [], None, 1
    depth_paren = 0
    i, n = 0, len(text)
    init_depth = 0  # brace depth inside an initializer (not a block)
    while i < n:
        c = text[i]
        if c == "\n":
            line += 1
        if c.isspace() and not buf:
            i += 1
            continue
        if buf_line is None and not c.isspace():
            buf_line = line
        if c == '"':  # string literal: copied whole (a SUBCASE name may hold an unbalanced parenthesis)
            j = text.index('"', i + 1)
            buf.append(text[i:j + 1])
            i = j + 1
            continue
        if c == "(":
            depth_paren += 1
        elif c == ")":
            depth_paren -= 1
        if c == "{" and depth_paren == 0 and init_depth == 0:
            head = "".join(buf).strip()
            if is_block_head(head):
                node = ("block", head, buf_line, [])
                stack[-1].append(node)
                stack.append(node[3])
                buf, buf_line = [], None
                i += 1
                continue
            init_depth = 1
            buf.append(c)
            i += 1
            continue
        if init_depth:
            if c == "{":
                init_depth += 1
            elif c == "}":
                init_depth -= 1
            buf.append(c)
            i += 1
            continue
        if c == "}" and depth_paren == 0:
            if "".join(buf).strip():
                raise SyntaxError(f"dangling text before '}}' at line {line}: {''.join(buf)!r}")
            stack.pop()
            buf, buf_line = [], None
            i += 1
            continue
        if c == ";" and depth_paren == 0:
            stmt = " ".join("".join(buf).split())
            stack[-1].append(("stmt", stmt, buf_line))
            buf, buf_line = [], None
            i += 1
            continue
        buf.append(c)
        i += 1
    return root


# ------------------------------------------------------------------------------------------ expressions
TOKEN = re.compile(r"\s*(?:(?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?[fFuUlL]*)|"
                   r"(?P<name>[A-Za-z_]\w*(?:\s*<\s*(?:T|unsigned int|unsigned|int|double|float)\s*>)?"
                   r"(?:::[A-Za-z_]\w*(?:\s*<\s*(?:T|unsigned int|unsigned|int|double|float)\s*>)?)*)|"
                   r"(?P<op>[-+*/(){}\[\],]))")


def tokenize(s):
    toks, pos = [], 0
    s = s.strip()
    while pos < len(s):
        m = TOKEN.match(s, pos)
        if not m or m.end() == pos:
            raise SyntaxError(f"cannot tokenize {s[pos:]!r}")
        if m.group("num"):
            t = m.group("num")
            if re.search(r"[.eE]", t.rstrip("fFuUlL")) and not re.fullmatch(r"\d+[uUlL]*", t):
                toks.append(("num", float(t.rstrip("fFuUlL"))))
            else:
                toks.append(("num", int(t.rstrip("fFuUlL"))))
        elif m.group("name"):
            toks.append(("name", re.sub(r"\s+", "", m.group("name")).replace("unsignedint", "unsigned int")))
        else:
            toks.append(("op", m.group("op")))
        pos = m.end()
    return toks


class Expr:
    def __init__(self, text, env):
        self.t = tokenize(text)
        self.i = 0
        self.env = env

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.peek()
        if val is not None and tok[1] != val:
            raise SyntaxError(f"expected {val!r}, got {tok!r}")
        self.i += 1
        return tok

    def parse(self):
        v = self.expr()
        if self.i != len(self.t):
            raise SyntaxError(f"trailing tokens {self.t[self.i:]}")
        return v

    def expr(self):
        v = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            v = binop(op, v, self.term())
        return v

    def term(self):
        v = self.unary()
        while self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            v = binop(op, v, self.unary())
        return v

    def unary(self):
        if self.peek()[1] == "-":
            self.take()
            return neg(self.unary())
        if self.peek()[1] == "+":
            self.take()
            return self.unary()
        return self.postfix()

    def postfix(self):
        v = self.primary()
        while self.peek()[1] == "[":
            self.take("[")
            k = self.expr()
            self.take("]")
            v = v[k]
        return v

    def args(self, close=")"):
        out = []
        if self.peek()[1] == close:
            self.take(close)
            return out
        while True:
            out.append(self.expr())
            tok = self.take()[1]
            if tok == close:
                return out
            if tok != ",":
                raise SyntaxError(f"expected ',' or {close!r}, got {tok!r}")

    def primary(self):
        kind, val = self.take()
        if kind == "num":
            return val
        if val == "(":
            v = self.expr()
            self.take(")")
            return v
        if val == "{":  # nested initializer list (matrix row)
            return self.args("}")
        if kind != "name":
            raise SyntaxError(f"unexpected {val!r}")
        if val in ("true", "false"):
            return val == "true"
        if val in ("static_cast<T>", "T", "static_cast<double>", "double", "float", "static_cast<float>"):
            self.take("(")
            (x,) = self.args()
            return float(x)
        if val in ("static_cast<unsignedint>", "static_cast<unsigned int>", "static_cast<unsigned>"):
            self.take("(")
            (x,) = self.args()
            return int(x) & 0xFFFFFFFF
        if val in ("std::sqrt", "sqrt", "std::sin", "std::cos", "std::abs", "std::fabs"):
            self.take("(")
            (x,) = self.args()
            f = {"sqrt": math.sqrt, "sin": math.sin, "cos": math.cos, "abs": abs, "fabs": abs}[val.split("::")[-1]]
            return f(float(x)) if val.split("::")[-1] != "abs" else abs(x)
        if val.startswith("std::numeric_limits<"):
            self.take("(")
            self.take(")")
            return {"max": DBL_MAX, "min": DBL_MIN, "epsilon": DBL_EPS, "infinity": math.inf}[val.split("::")[-1]]
        if val == "pi<T>":
            return math.pi
        if val in ("Vec3r<T>", "Vec3ui"):
            self.take("{")
            xs = self.args("}")
            if len(xs) != 3:
                raise SyntaxError(f"Vec3 with {len(xs)} components")
            return [float(x) for x in xs] if val == "Vec3r<T>" else [int(x) for x in xs]
        if val == "blaze::IdentityMatrix<T>":
            self.take("(")
            self.args()
            return copy.deepcopy(IDENTITY)
        if val == "arbitraryRotationMatrix":
            self.take("(")
            axis, angle = self.args()
            return rotation_matrix(axis, angle)
        if val in self.env:
            return self.env[val]
        raise NameError(f"unknown name {val!r}")


def c_div(a, b):
    if isinstance(a, int) and isinstance(b, int) and not isinstance(a, bool):
        q = abs(a) // abs(b)
        return q if (a >= 0) == (b >= 0) else -q
    return float(a) / float(b)


def binop(op, a, b):
    if isinstance(a, list) or isinstance(b, list):
        if isinstance(a, list) and isinstance(b, list):
            return [binop(op, x, y) for x, y in zip(a, b)]
        if isinstance(a, list):
            return [binop(op, x, b) for x in a]
        return [binop(op, a, y) for y in b]
    if op == "+":
        return a + b
    if op == "-":
        return a - b
    if op == "*":
        return a * b
    return c_div(a, b)


def neg(a):
    return [neg(x) for x in a] if isinstance(a, list) else -a


def rotation_matrix(axis, angle):
    """test/test_helpers.h:439-451 (axis-angle, right-handed)."""
    n1, n2, n3 = (float(x) for x in axis)
    s, c = math.sin(angle), math.cos(angle)
    return [[n1 * n1 * (1 - c) + c, n1 * n2 * (1 - c) - n3 * s, n1 * n3 * (1 - c) + n2 * s],
            [n1 * n2 * (1 - c) + n3 * s, n2 * n2 * (1 - c) + c, n2 * n3 * (1 - c) - n1 * s],
            [n1 * n3 * (1 - c) - n2 * s, n2 * n3 * (1 - c) + n1 * s, n3 * n3 * (1 - c) + c]]


def ev(text, env):
    return Expr(text, env).parse()


# ------------------------------------------------------------------------------------------ interpreter
COLLECTIONS = {"TriangleMesh": "mesh", "SphereCollection": "spheres", "PlaneCollection": "planes",
               "CylinderCollection": "cylinders"}
COLL_FIELDS = {"mesh": ("vertices", "faces"), "spheres": ("centers", "radii"),
               "planes": ("centers", "dxs", "dys", "rotations"),
               "cylinders": ("centers", "a", "b", "heights", "rotations")}
DECL = re.compile(r"^(?:const\s+|static\s+|constexpr\s+)*(Vec3r<T>|Mat3r<T>|Ray<T>|RayHit<T>|T|float|double|bool|"
                  r"unsigned int|unsigned|int|auto)\s+(\w+)\s*(=|\{|\(|$)(.*)$")


class Walker:
    def __init__(self, fname, helpers):
        self.fname = fname
        self.helpers = helpers
        self.cases = []
        self.warn_loops = 0
        self.test_case = None
        self.path = []

    def run(self, tree):
        for node in tree:
            if node[0] == "block" and node[1].startswith("TEST_CASE"):
                self.test_case = re.search(r'"([^"]*)"', node[1]).group(1)
                self.block(node[3], {})

    def block(self, children, env):
        for node in children:
            if node[0] == "stmt":
                self.stmt(node[1], node[2], env)
            else:
                head = node[1]
                if head.startswith("SUBCASE"):
                    name = re.search(r'"([^"]*)"', head).group(1)
                    self.path.append(name)
                    self.block(node[3], copy.deepcopy(env))
                    self.path.pop()
                elif head.startswith("for"):
                    body = json.dumps(node[3])
                    if "assert_traverse_bvh_hit_trimesh_precision" not in body:
                        raise SyntaxError(f"{self.fname}:{node[2]}: loop with non-WARN asserts")
                    self.warn_loops += 1
                else:
                    raise SyntaxError(f"{self.fname}:{node[2]}: unsupported block {head!r}")

    def coll_snapshot(self, kind, args, env):
        names = [a.strip().lstrip("*") for a in args]
        vals = [copy.deepcopy(env[n]) for n in names]
        return {"geometry": kind, **dict(zip(COLL_FIELDS[kind], vals))}

    def emit(self, assert_kind, line, **kw):
        self.cases.append({"assert": assert_kind, "ref": f"{self.fname}:{line}", "test_case": self.test_case,
                           "path": " / ".join(self.path), **kw})

    def stmt(self, s, line, env):
        if s.startswith("using namespace") or s.startswith("INFO(") or s.startswith("WARN"):
            return
        m = re.match(r"^auto\s+(\w+)\s*=\s*std::make_unique<.*>\(\)$", s)
        if m:
            env[m.group(1)] = []
            return
        m = re.match(r"^(\w+)->(emplace_back|push_back)\((.*)\)$", s)
        if m:
            env[m.group(1)].append(ev(m.group(3), env))
            return
        m = re.match(r"^(?:const\s+)?(\w+)(?:<T>)?\s+(\w+)\s*[{(](.*)[})]$", s)
        if m and m.group(1) in COLLECTIONS:
            kind = COLLECTIONS[m.group(1)]
            env[m.group(2)] = self.coll_snapshot(kind, m.group(3).split(","), env)
            return
        m = re.match(r"^(\w+)\((.*)\)$", s)
        if m and m.group(1) in self.helpers:
            self.call_helper(m.group(1), m.group(2), env)
            return
        if m and m.group(1).startswith("assert_"):
            self.assertion(m.group(1), split_args(m.group(2)), line, env)
            return
        m = DECL.match(s)
        if m:
            typ, name, sep, rest = m.groups()
            if typ == "RayHit<T>":
                return
            if typ == "Ray<T>":
                args = split_args(rest.rstrip("})"))
                vals = [ev(a, env) for a in args]
                env[name] = {"org": vals[0], "dir": vals[1]}
                if len(vals) > 2:
                    env[name]["tmin"], env[name]["tmax"] = vals[2], vals[3]
                return
            if sep == "=":
                val = ev(rest, env)
            elif sep == "{":
                inner = rest[:-1]
                if typ == "Mat3r<T>":
                    val = ev("{" + inner + "}", env)
                elif typ == "Vec3r<T>":
                    val = ev("Vec3r<T>{" + inner + "}", env)
                else:
                    val = ev(inner, env)
            else:
                raise SyntaxError(f"{self.fname}:{line}: declaration without value: {s}")
            if typ in ("T", "float", "double"):
                val = float(val)
            elif typ in ("unsigned int", "unsigned"):
                val = int(val) & 0xFFFFFFFF
            elif typ == "Vec3r<T>":
                val = [float(x) for x in val]
            env[name] = val
            return
        raise SyntaxError(f"{self.fname}:{line}: unsupported statement: {s}")

    def call_helper(self, name, argtext, env):
        params, body = self.helpers[name]
        args = [a.strip() for a in split_args(argtext)]
        local = {}
        for (ptype, pname), a in zip(params, args):
            local[pname] = env[a.lstrip("*")] if a.startswith("*") else ev(a, env)
        for s in body:
            m = re.match(r"^(\w+)\.emplace_back\((.*)\)$", s)
            if not m:
                raise SyntaxError(f"helper {name}: unsupported statement {s}")
            local[m.group(1)].append(ev(m.group(2), local))

    def assertion(self, fn, args, line, env):
        if fn == "assert_traverse_bvh_hit_trimesh_precision":
            self.warn_loops += 1
            return
        coll = env[args[0].strip()]
        if fn in ("assert_intersect_primitive_hit", "assert_traverse_bvh_hit"):
            ray = env[args[1].strip()]
            hit, prim, dist, normal = (ev(a, env) for a in args[2:6])
            self.emit("intersect_primitive" if "intersect_primitive" in fn else "traverse", line,
                      scene=coll, org=ray["org"], dir=ray["dir"], hit=bool(hit), prim_id=int(prim),
                      t=float(dist), normal=[float(x) for x in normal])
        elif fn == "assert_traverse_bvh_hit_trimesh_distance":
            ray = env[args[1].strip()]
            hit, dist = (ev(a, env) for a in args[2:4])
            self.emit("traverse_distance", line, scene=coll, org=ray["org"], dir=ray["dir"], hit=bool(hit),
                      t=float(dist))
        elif fn == "assert_traverse_bvh_hit_collection":
            ray = env[args[1].strip()]
            hit, dist, normal = (ev(a, env) for a in args[2:5])
            self.emit("traverse_collection", line, scene=coll, org=ray["org"], dir=ray["dir"], hit=bool(hit),
                      t=float(dist), normal=[float(x) for x in normal])
        elif fn == "assert_bounding_box":
            prim, bmin, bmax = (ev(a, env) for a in args[1:4])
            self.emit("bounding_box", line, scene=coll, prim=int(prim), bmin=bmin, bmax=bmax)
        elif fn == "assert_bounding_box_collection":
            bmin, bmax = (ev(a, env) for a in args[1:3])
            self.emit("bounding_box_collection", line, scene=coll, bmin=bmin, bmax=bmax)
        elif fn == "assert_primitive_center":
            prim, center = (ev(a, env) for a in args[1:3])
            self.emit("primitive_center", line, scene=coll, prim=int(prim), center=center)
        elif fn == "assert_distance_to_surface":
            prim, point, dist = (ev(a, env) for a in args[1:4])
            self.emit("distance_to_surface", line, scene=coll, prim=int(prim), point=point, distance=float(dist))
        else:
            raise SyntaxError(f"{self.fname}:{line}: unknown assert {fn}")


def split_args(text):
    out, depth, cur = [], 0, []
    for c in text:
        if c in "({[":
            depth += 1
        elif c in ")}]":
            depth -= 1
        if c == "," and depth == 0:
            out.append("".join(cur).strip())
            cur = []
        else:
            cur.append(c)
    if "".join(cur).strip():
        out.append("".join(cur).strip())
    return out


def load_helpers():
    tree = parse_blocks(strip_comments(open(os.path.join(REF, HELPERS)).read()))
    helpers = {}
    for node in tree:
        if node[0] != "block":
            continue
        m = re.search(r"void\s+(\w+)\s*\((.*)\)\s*$", node[1])
        if not m:
            continue
        params = []
        for p in split_args(m.group(2)):
            pm = re.match(r"^(?:const\s+)?([\w<>:]+)\s*&?\s*(\w+)$", p.strip())
            params.append((pm.group(1), pm.group(2)))
        body = [n[1] for n in node[3] if n[0] == "stmt"]
        if all(re.match(r"^\w+\.emplace_back\(", s) for s in body):
            helpers[m.group(1)] = (params, body)
    return helpers


def main():
    helpers = load_helpers()
    all_cases, warn = [], {}
    for f in FILES:
        w = Walker(f, helpers)
        w.run(parse_blocks(strip_comments(open(os.path.join(REF, f)).read())))
        all_cases += w.cases
        warn[f] = w.warn_loops
    counts = {}
    for c in all_cases:
        k = f"{c['scene']['geometry']}:{c['assert']}"
        counts[k] = counts.get(k, 0) + 1
    doc = {"source": "every assert_* of blazert's test/test_primitives/test_{trimesh,sphere,plane,cylinder}.cpp, "
                     "extracted by tests/golden/make_reference_cases.py (values for T = double). Tolerance in the "
                     "reference: doctest::Approx default (epsilon 100*FLT_EPSILON, scale 1).",
           "approx_epsilon": 100 * 1.1920928955078125e-07, "warn_only_asserts_skipped": warn, "counts": counts,
           "cases": all_cases}
    json.dump(doc, open(OUT, "w"), indent=None, separators=(",", ":"))
    print(f"{len(all_cases)} cases -> {OUT}")
    for k in sorted(counts):
        print(f"  {k}: {counts[k]}")
    print("  WARN-only asserts skipped:", warn)


if __name__ == "__main__":
    main()
