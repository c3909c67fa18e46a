"""Writes tests/golden/cylinder_distance_cases.json from the reference's cylinder distance known answers.

Source: the SUBCASE("distance to surface") block of test/test_primitives/test_cylinder.cpp in the blazert
repository, whose CHECKs compare distance_to_surface(cylinder 0, point) with a true distance under doctest's
default Approx. The doctest SUBCASEs are nested: each leaf path re-runs the test case from the top, so the
cylinder is what the enclosing blocks pushed into the collection vectors (element 0 of each) when the collection
in scope was made. This script walks the block structure and records, for every assert_distance_to_surface, the
cylinder (center, semi axes, height, row-major rotation), the point, the expected distance and its source line.
Only this script reads the reference tree (as text); the tests use the committed JSON.
"""
import copy
import json
import math
import os
import re

SRC = "/root/reference/test/test_primitives/test_cylinder.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cylinder_distance_cases.json")
IDENTITY = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]


def cexpr(text, env):
    """Evaluates a C++ scalar expression of the test (literals, std::sqrt, static_cast<T>, named constants)."""
    t = text.replace("std::sqrt", "math.sqrt").replace("static_cast<T>", "")
    t = re.sub(r"(\d)\.(?!\d)", r"\1.0", t)
    return float(eval(t, {"math": math}, dict(env)))


def vec3(text, env):
    inner = re.search(r"Vec3r<T>\{(.*)\}", text).group(1)
    return [cexpr(x, env) for x in inner.split(",")]


def main():
    lines = open(SRC).read().split("\n")
    start = next(i for i, s in enumerate(lines) if 'SUBCASE("distance to surface")' in s)
    state = {"vec": {"centers": [], "a": [], "b": [], "h": [], "rot": []}, "env": {}, "coll": None}
    stack = [state]
    depth = 0
    cases = []
    i = start
    while i < len(lines):
        s = lines[i].strip()
        # join an assert spread over several lines
        j = i
        if s.startswith("assert_distance_to_surface(") and not s.endswith(";"):
            while not lines[j].strip().endswith(";"):
                j += 1
                s += " " + lines[j].strip()
        cur = stack[-1]
        v = cur["vec"]
        if s.startswith("SUBCASE("):
            stack.append(copy.deepcopy(cur))
        elif s.startswith("centers->emplace_back("):
            v["centers"].append(vec3(s, cur["env"]))
        elif s.startswith("semi_axes_a->emplace_back("):
            v["a"].append(cexpr(s[s.index("(") + 1:s.rindex(")")], cur["env"]))
        elif s.startswith("semi_axes_b->emplace_back("):
            v["b"].append(cexpr(s[s.index("(") + 1:s.rindex(")")], cur["env"]))
        elif s.startswith("heights->emplace_back("):
            v["h"].append(cexpr(s[s.index("(") + 1:s.rindex(")")], cur["env"]))
        elif re.match(r"(const )?Mat3r<T> rot = blaze::IdentityMatrix", s):
            cur["env"]["__rot"] = IDENTITY
        elif s.startswith("Mat3r<T> rot{"):
            rows = re.findall(r"\{([^{}]*)\}", s[s.index("{") + 1:])
            cur["env"]["__rot"] = [cexpr(x, cur["env"]) for r in rows for x in r.split(",")]
        elif s.startswith("rotations->push_back(rot)"):
            v["rot"].append(cur["env"]["__rot"])
        elif s.startswith("rotations->emplace_back(blaze::IdentityMatrix"):
            v["rot"].append(IDENTITY)
        elif s.startswith("CylinderCollection"):
            cur["coll"] = copy.deepcopy(v)
        elif s.startswith("const T sqrt_1_div_2"):
            cur["env"]["sqrt_1_div_2"] = math.sqrt(1.0 / 2.0)
        elif s.startswith("assert_distance_to_surface("):
            m = re.match(r"assert_distance_to_surface\(\s*cylinders, prim_id, (Vec3r<T>\{[^}]*\}), (.*)\);$", s)
            env = {k: x for k, x in cur["env"].items() if not k.startswith("__")}
            c = cur["coll"]
            cases.append({"line": i + 1, "center": c["centers"][0], "a": c["a"][0], "b": c["b"][0],
                          "height": c["h"][0], "rotation": c["rot"][0], "point": vec3(m.group(1), env),
                          "distance": cexpr(m.group(2), env)})
        depth += s.count("{") - s.count("}")
        # a line that closes blocks pops them (SUBCASE bodies are the only blocks here)
        closes = s.count("}") - s.count("{")
        if closes > 0 and not s.startswith("SUBCASE("):
            for _ in range(closes):
                stack.pop()
        if depth <= 0:  # the closing brace of the distance block
            break
        i = j + 1
    doc = {"source": "test/test_primitives/test_cylinder.cpp (SUBCASE \"distance to surface\"), extracted by "
                     "tests/golden/make_cylinder_distance_cases.py; tolerance in the reference: doctest::Approx "
                     "default (epsilon 100*FLT_EPSILON, scale 1)",
           "approx_epsilon": 100 * 1.1920928955078125e-07, "cases": cases}
    json.dump(doc, open(OUT, "w"), indent=1)
    print(f"{len(cases)} cases -> {OUT}")


if __name__ == "__main__":
    main()
