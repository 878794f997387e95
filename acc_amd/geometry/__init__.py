"""Shapes for samples: primitives, CSG operations, XML parsing and encoding.

The reference receives shapes from mcvine's `instrument.nixml` parser and
renders them into Numba closures with a visitor
(acc/geometry/arrow_intersect.py:18-142, acc/geometry/locate.py:8-65).  Here a
shape tree is (a) encoded as a node table for the C ABI (MCV_CSG_* in
include/mcv_acc.h; node 0 is the root) and (b) rendered as the C++ type the
fuser instantiates (mcv::CsgSphere<i>, mcv::CsgDifference<i, A, B>, ...).

Cylinders have their axis along local z and a height centred on the origin,
as in the reference (arrow_intersect.py:171-177).
"""
import ast
import operator
import xml.etree.ElementTree as ET

import numpy as np

from .._abi import (MCV_CSG_NODE_SIZE, MCV_CSG_SPHERE, MCV_CSG_CYLINDER, MCV_CSG_BLOCK, MCV_CSG_UNION,
                    MCV_CSG_DIFFERENCE, MCV_CSG_INTERSECTION)

# acc/geometry/__init__.py:2 and location.py
epsilon = 1e-10
outside = 1
onborder = 0
inside = -1


class Shape:
    pass


class Sphere(Shape):
    def __init__(self, radius):
        self.radius = float(radius)

    def identify(self, visitor):
        return visitor.onSphere(self)


class Cylinder(Shape):
    def __init__(self, radius, height):
        self.radius = float(radius)
        self.height = float(height)

    def identify(self, visitor):
        return visitor.onCylinder(self)


class Block(Shape):
    def __init__(self, width, height, thickness):
        self.width = float(width)
        self.height = float(height)
        self.thickness = float(thickness)

    def identify(self, visitor):
        return visitor.onBlock(self)


class Union(Shape):
    def __init__(self, s1, s2):
        self.shapes = (s1, s2)

    def identify(self, visitor):
        return visitor.onUnion(self)


class Difference(Shape):
    def __init__(self, op1, op2):
        self.op1, self.op2 = op1, op2

    def identify(self, visitor):
        return visitor.onDifference(self)


class Intersection(Shape):
    def __init__(self, s1, s2):
        self.shapes = (s1, s2)

    def identify(self, visitor):
        return visitor.onIntersection(self)


def _children(s):
    if isinstance(s, (Union, Intersection)):
        return list(s.shapes)
    if isinstance(s, Difference):
        return [s.op1, s.op2]
    return []


def encode(shape):
    """Node table (n_nodes, MCV_CSG_NODE_SIZE), root first, pre-order."""
    nodes = []

    def visit(s):
        idx = len(nodes)
        nodes.append(None)
        if isinstance(s, Sphere):
            nodes[idx] = [MCV_CSG_SPHERE, s.radius, 0, 0, -1, -1]
        elif isinstance(s, Cylinder):
            nodes[idx] = [MCV_CSG_CYLINDER, s.radius, s.height, 0, -1, -1]
        elif isinstance(s, Block):
            nodes[idx] = [MCV_CSG_BLOCK, s.width, s.height, s.thickness, -1, -1]
        else:
            kind = {Union: MCV_CSG_UNION, Difference: MCV_CSG_DIFFERENCE,
                    Intersection: MCV_CSG_INTERSECTION}[type(s)]
            a, b = _children(s)
            ia = visit(a)
            ib = visit(b)
            nodes[idx] = [kind, 0, 0, 0, ia, ib]
        return idx

    visit(shape)
    return np.array(nodes, dtype=np.float64).reshape(-1, MCV_CSG_NODE_SIZE)


def cpp_type(shape):
    """C++ type of the shape for the fused kernel (node indices as in encode())."""
    counter = [0]

    def visit(s):
        idx = counter[0]
        counter[0] += 1
        if isinstance(s, Sphere):
            return f"mcv::CsgSphere<{idx}>"
        if isinstance(s, Cylinder):
            return f"mcv::CsgCylinder<{idx}>"
        if isinstance(s, Block):
            return f"mcv::CsgBlock<{idx}>"
        name = {Union: "CsgUnion", Difference: "CsgDifference", Intersection: "CsgIntersection"}[type(s)]
        a, b = _children(s)
        ta = visit(a)
        tb = visit(b)
        return f"mcv::{name}<{idx}, {ta}, {tb}>"

    return visit(shape)


def max_roots(shape):
    if isinstance(shape, (Sphere, Cylinder, Block)):
        return 2
    a, b = _children(shape)
    return min(10, max_roots(a) + max_roots(b))


# ---------------------------------------------------------------- XML
_UNITS = {"m": 1.0, "meter": 1.0, "cm": 1e-2, "mm": 1e-3, "um": 1e-6, "angstrom": 1e-10, "AA": 1e-10,
          "km": 1e3, "deg": 1.0, "degree": 1.0}

_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
        ast.Pow: operator.pow, ast.USub: operator.neg, ast.UAdd: operator.pos}


def eval_quantity(text, units=None):
    """Evaluate '2.5*cm', '10./m', '4.*mm' with SI units (arithmetic only)."""
    table = dict(_UNITS if units is None else units)

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](ev(node.operand))
        if isinstance(node, ast.Name) and node.id in table:
            return table[node.id]
        raise ValueError(f"unsupported quantity expression: {text!r}")

    return ev(ast.parse(text.strip(), mode="eval"))


def shape_from_element(el):
    tag = el.tag.lower()
    if tag == "sphere":
        return Sphere(eval_quantity(el.attrib["radius"]))
    if tag == "cylinder":
        return Cylinder(eval_quantity(el.attrib["radius"]), eval_quantity(el.attrib["height"]))
    if tag == "block":
        return Block(eval_quantity(el.attrib["width"]), eval_quantity(el.attrib["height"]),
                     eval_quantity(el.attrib["thickness"]))
    kids = [c for c in el if isinstance(c.tag, str)]
    if tag in ("union", "difference", "intersection"):
        shapes = [shape_from_element(c) for c in kids]
        if len(shapes) != 2:
            raise ValueError(f"<{tag}> needs exactly two operands")
        return {"union": Union, "difference": Difference, "intersection": Intersection}[tag](*shapes)
    raise ValueError(f"unsupported shape element <{el.tag}>")


def parse_file(path):
    """Shapes under a <geometry> document (instrument.nixml.parse_file's role
    in tests/geometry/test_from_xml.py)."""
    root = ET.parse(path).getroot()
    return [shape_from_element(c) for c in root if isinstance(c.tag, str)]
