"""Sources: Python mirror of the crate's implicit shapes and CSG operators.

Each source lowers itself to the C ABI's post-order iso_node program, which
the GPU evaluates with the reference's f32 operation order:

  Sphere(radius)                       implicit/sphere.rs:22-39
  Torus(radius, tube_radius)           implicit/torus.rs:23-46
  RectangularPrism(half_extent)        implicit/rectangular_prism.rs:22-41
  Cylinder(radius, half_length)        implicit/cylinder.rs:21-48
  Union(a, b) / Intersection(a, b)     implicit/csg.rs:22-72
  Difference(a, b)  (b minus a)        implicit/csg.rs:70-100
  DemoSource(s): s sampled at p - 0.5  examples/common/sources.rs:26-56

Translate(offset, s) has no reference type; it is the user-level wrapper the
examples build (DemoSource is Translate((0.5, 0.5, 0.5), s)).  SmoothUnion
is the build's own operator for the smooth-min workload.

Any other source (an object with `sample_scalar(x, y, z)` over numpy arrays)
is sampled on the host into a dense grid by the extractors.
"""
from __future__ import annotations

from typing import Iterable, List, Sequence, Tuple

from . import _abi as A

Node = Tuple[int, Tuple[float, ...]]


class Source:
    """A source the GPU can evaluate: lowers to an iso_node program.

    Each source describes itself as (nodes before its children, children,
    nodes after them); program() walks the tree with an explicit stack, so
    chains thousands of operators deep (a smooth-min fold) lower too."""

    def program(self) -> List[Node]:
        out: List[Node] = []
        stack: List[object] = [self]
        while stack:
            item = stack.pop()
            if isinstance(item, Source):
                pre, children, post = item._parts()
                out.extend(pre)
                stack.append(("nodes", post))
                stack.extend(reversed(children))
            else:
                out.extend(item[1])
        return out

    def _parts(self) -> Tuple[List[Node], List["Source"], List[Node]]:  # pragma: no cover
        raise NotImplementedError

    def _emit(self, out: List[Node]) -> None:
        out.extend(self.program())


class Sphere(Source):
    def __init__(self, radius: float):
        self.radius = float(radius)

    def _parts(self):
        return [], [], [(A.ISO_OP_SPHERE, (self.radius,))]


class Torus(Source):
    def __init__(self, radius: float, tube_radius: float):
        self.radius = float(radius)
        self.tube_radius = float(tube_radius)

    def _parts(self):
        return [], [], [(A.ISO_OP_TORUS, (self.radius, self.tube_radius))]


class RectangularPrism(Source):
    def __init__(self, half_extent: Sequence[float]):
        self.half_extent = tuple(float(v) for v in half_extent)
        assert len(self.half_extent) == 3

    def _parts(self):
        return [], [], [(A.ISO_OP_BOX, self.half_extent)]


class Cylinder(Source):
    def __init__(self, radius: float, half_length: float):
        self.radius = float(radius)
        self.half_length = float(half_length)

    def _parts(self):
        return [], [], [(A.ISO_OP_CYLINDER, (self.radius, self.half_length))]


class _Binary(Source):
    OP = 0

    def __init__(self, a: Source, b: Source):
        self.a = a
        self.b = b

    def _parts(self):
        return [], [self.a, self.b], [(self.OP, ())]


class Union(_Binary):
    OP = A.ISO_OP_UNION


class Intersection(_Binary):
    OP = A.ISO_OP_INTERSECTION


class Difference(_Binary):
    """Solid where `b` is solid except where `a` is: max(b, -a)."""

    OP = A.ISO_OP_DIFFERENCE


class SmoothUnion(Source):
    def __init__(self, a: Source, b: Source, k: float):
        self.a, self.b, self.k = a, b, float(k)

    def _parts(self):
        return [], [self.a, self.b], [(A.ISO_OP_SMOOTH_UNION, (self.k,))]


class Translate(Source):
    """`source` sampled at p - offset."""

    def __init__(self, offset: Sequence[float], source: Source):
        self.offset = tuple(float(v) for v in offset)
        self.source = source

    def _parts(self):
        return [(A.ISO_OP_TRANSLATE, self.offset)], [self.source], [(A.ISO_OP_END_TRANSLATE, ())]


class CentralDifference:
    """CentralDifference::new / new_with_epsilon (source.rs:52-94): a
    HermiteSource whose normals are central differences of `source`
    (step epsilon, default 1e-6), as dual contouring needs."""

    def __init__(self, source: Source, epsilon: float = 0.000001):
        if not isinstance(source, Source):
            raise TypeError("CentralDifference wraps a program source")
        self.source = source
        self.epsilon = float(epsilon)

    def program(self) -> List[Node]:
        return self.source.program()


def DemoSource(source: Source) -> Source:
    """The examples' recentring wrapper: q = p - Vec3::from_scalar(0.5)."""
    return Translate((0.5, 0.5, 0.5), source)


def union_all(sources: Iterable[Source]) -> Source:
    """Left fold of Union over `sources` (Union(Union(s0, s1), s2) ...)."""
    it = iter(sources)
    acc = next(it)
    for s in it:
        acc = Union(acc, s)
    return acc


def smooth_union_all(sources: Iterable[Source], k: float) -> Source:
    it = iter(sources)
    acc = next(it)
    for s in it:
        acc = SmoothUnion(acc, s, k)
    return acc


def to_nodes(program: Sequence[Node]):
    """Packs a program into a ctypes iso_node array."""
    arr = (A.IsoNode * len(program))()
    for i, (op, params) in enumerate(program):
        arr[i].op = op
        arr[i].reserved = 0
        for j, v in enumerate(params):
            arr[i].p[j] = v
    return arr
