"""Writes reference_kats.json, reference_vector_kats.json and
reference_normal_kats.json: the scalar, directed and normal known-answer
tests of the reference crate's own unit tests, as data.

Reads the Rust test modules of src/implicit/{sphere,torus,cylinder,
rectangular_prism,csg}.rs (only in the build container, where the reference
is mounted) and records, for every `assert_eq!(x.sample_scalar(p).0, v)`,
the source that `x` names (as an iso_node post-order program), the point and
the expected f32 value; and for every `assert_eq!(x.sample_vector(p).0, e)`
the expected Vec3, with the f32 expressions the tests write (such as
`Vec2::from_scalar(1.0 - (4.0f32 - 1.0f32).sqrt()).extend(-3.0)`) evaluated
in f32 and f32::MAX kept as "MAX"; and for every
`assert_eq!(x.sample_normal(p).normalised().unwrap(), n)` the expected unit
normal.  The JSON is committed; tests never read the reference itself.

usage: python tests/golden/make_kats.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

import numpy as np

OPS = {"Sphere": 1, "Torus": 2, "RectangularPrism": 3, "Cylinder": 4, "Union": 16,
       "Intersection": 17}
FILES = ["sphere.rs", "torus.rs", "cylinder.rs", "rectangular_prism.rs", "csg.rs"]

NUM = r"-?\d+(?:\.\d+)?"


def vec3(text):
    text = text.strip()
    if text == "Vec3::zero()":
        return [0.0, 0.0, 0.0]
    m = re.fullmatch(rf"Vec3::new\(({NUM}),\s*({NUM}),\s*({NUM})\)", text)
    assert m, text
    return [float(m.group(i)) for i in range(1, 4)]


class F32:
    """An f32 value with the arithmetic the expected-value expressions use."""

    def __init__(self, v):
        self.v = np.float32(v)

    def __add__(self, o):
        return F32(self.v + F32.of(o))

    def __radd__(self, o):
        return F32(F32.of(o) + self.v)

    def __sub__(self, o):
        return F32(self.v - F32.of(o))

    def __rsub__(self, o):
        return F32(F32.of(o) - self.v)

    def __mul__(self, o):
        return F32(self.v * F32.of(o))

    def __rmul__(self, o):
        return F32(F32.of(o) * self.v)

    def __neg__(self):
        return F32(-self.v)

    def sqrt(self):
        return F32(np.sqrt(self.v))

    @staticmethod
    def of(o):
        return o.v if isinstance(o, F32) else np.float32(o)


MAX = F32(np.finfo(np.float32).max)


def f32_expr(text):
    """Evaluates one Rust f32 expression of the tests (numbers, MAX, + - *,
    parentheses, .sqrt())."""
    py = re.sub(r"(\d+\.\d+)(?:f32)?", r"F32(\1)", text.strip())
    assert re.fullmatch(r"[F32()\d.\s+\-*MAXsqrt]*", py), text
    return eval(py, {"F32": F32, "MAX": MAX}).v  # noqa: S307 - fixed grammar above


def vec3_expr(text):
    text = " ".join(text.split())
    if text == "Vec3::zero()":
        vals = [np.float32(0)] * 3
    elif text.startswith("Vec3::from_scalar("):
        vals = [f32_expr(text[len("Vec3::from_scalar("):-1])] * 3
    elif text.startswith("Vec3::new("):
        vals = [f32_expr(t) for t in text[len("Vec3::new("):-1].split(",")]
    else:
        m = re.fullmatch(r"Vec2::from_scalar\((.*)\)\.extend\((.*)\)", text)
        assert m, text
        xy = f32_expr(m.group(1))
        vals = [xy, xy, f32_expr(m.group(2))]
    fmax = np.finfo(np.float32).max
    return ["MAX" if v == fmax else float(v) for v in vals]


def vector_kats(name, text, progs):
    out = []
    pat = re.compile(r"assert_eq!\(\s*(\w+)\.sample_vector\((Vec3::\w+\([^)]*\))\)\.0,\s*(.*?),?\s*\);",
                     re.S)
    for m in pat.finditer(text):
        var, p, e = m.groups()
        lineno = text.count("\n", 0, m.start()) + 1
        out.append({"where": f"src/implicit/{name}:{lineno}", "program": progs[var],
                    "point": vec3(p), "expected": vec3_expr(e)})
    return out


def normal_kats(name, text, progs):
    out = []
    pat = re.compile(r"assert_eq!\(\s*(\w+)\s*\.sample_normal\((Vec3::\w+\([^)]*\))\)\s*"
                     r"\.normalised\(\)\s*\.unwrap\(\),\s*(.*?),?\s*\);", re.S)
    for m in pat.finditer(text):
        var, p, e = m.groups()
        lineno = text.count("\n", 0, m.start()) + 1
        out.append({"where": f"src/implicit/{name}:{lineno}", "program": progs[var],
                    "point": vec3(p), "expected": vec3_expr(e)})
    return out


def main(ref_root: str) -> None:
    kats = []
    vkats = []
    nkats = []
    for name in FILES:
        path = Path(ref_root) / "src" / "implicit" / name
        lines = path.read_text().splitlines()
        progs = {}
        for lineno, line in enumerate(lines, 1):
            m = re.search(r"let (\w+) = (\w+)::new\((.*)\);", line)
            if m and m.group(2) in OPS:
                var, typ, args = m.groups()
                if typ in ("Union", "Intersection"):
                    a, b = [s.strip().replace(".clone()", "") for s in args.split(",")]
                    progs[var] = progs[a] + progs[b] + [[OPS[typ], []]]
                elif typ == "RectangularPrism":
                    progs[var] = [[OPS[typ], vec3(args)]]
                else:
                    progs[var] = [[OPS[typ], [float(v) for v in args.split(",")]]]
            m = re.search(rf"assert_eq!\((\w+)\.sample_scalar\((Vec3::\w+\([^)]*\))\)\.0,\s*({NUM})\);",
                          line)
            if m:
                var, p, v = m.groups()
                kats.append({"where": f"src/implicit/{name}:{lineno}", "program": progs[var],
                             "point": vec3(p), "expected": float(v)})
        vkats += vector_kats(name, "\n".join(lines), progs)
        nkats += normal_kats(name, "\n".join(lines), progs)
    out = Path(__file__).with_name("reference_kats.json")
    out.write_text(json.dumps(kats, indent=1) + "\n")
    print(f"{len(kats)} known-answer tests -> {out}")
    vout = Path(__file__).with_name("reference_vector_kats.json")
    vout.write_text(json.dumps(vkats, indent=1) + "\n")
    print(f"{len(vkats)} directed known-answer tests -> {vout}")
    nout = Path(__file__).with_name("reference_normal_kats.json")
    nout.write_text(json.dumps(nkats, indent=1) + "\n")
    print(f"{len(nkats)} normal known-answer tests -> {nout}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
