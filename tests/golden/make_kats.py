"""Writes reference_kats.json: the scalar known-answer tests of the
reference crate's own unit tests, as data.

Reads the Rust test modules of src/implicit/{sphere,torus,cylinder,
rectangular_prism,csg}.rs (only in the build container, where the reference
is mounted) and records, for every `assert_eq!(x.sample_scalar(p).0, v)`,
the source that `x` names (as an iso_node post-order program), the point and
the expected f32 value.  The JSON is committed; tests never read the
reference itself.

usage: python tests/golden/make_kats.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

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


def main(ref_root: str) -> None:
    kats = []
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
    out = Path(__file__).with_name("reference_kats.json")
    out.write_text(json.dumps(kats, indent=1) + "\n")
    print(f"{len(kats)} known-answer tests -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
