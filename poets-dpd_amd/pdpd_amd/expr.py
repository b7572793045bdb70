"""Parameter expressions of WorldState files (the language of the reference's
include/dpd/core/expression.hpp; the C++ side is include/pdpd/expression.hpp):
numbers, variables, + - * / with precedence, parentheses. Evaluation is in
double precision, binary operations left to right, as in the C++ tree, so a
table evaluated here equals the one the C++ engine host evaluates."""
from __future__ import annotations

import re

# a decimal literal as strtod reads it (a leading '-' is an operator here)
_NUMBER = re.compile(r"(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?")


def _tokens(src):
    out, i = [], 0
    while i < len(src):
        if src[i].isspace():
            i += 1
            continue
        c = src[i]
        if c in "()+-*/":
            out.append(("op", c)); i += 1
        elif c.isalpha() or c == "_":
            m = re.match(r"[A-Za-z_][A-Za-z0-9_]*", src[i:])
            out.append(("var", m.group(0))); i += len(m.group(0))
        else:
            m = _NUMBER.match(src, i)
            if not m:
                raise ValueError(f"Couldn't parse expression from {src[i:]}")
            out.append(("num", float(m.group(0)))); i = m.end()
    return out


def parse(src):
    """A tree of tuples: ('c', value) | ('v', name) | (op, lhs, rhs)."""
    toks = _tokens(src)
    if not toks:
        raise ValueError("empty expression")
    pos = 0

    def at(op):
        return pos < len(toks) and toks[pos] == ("op", op)

    def atom():
        nonlocal pos
        if pos >= len(toks):
            raise ValueError("expression ends early")
        kind, v = toks[pos]
        pos += 1
        if kind == "num":
            return ("c", v)
        if kind == "var":
            return ("v", v)
        if v == "(":
            e = total()
            if not at(")"):
                raise ValueError("Un-matched '(' bracket.")
            pos += 1
            return e
        raise ValueError(f"Unexpected token in parse_expression_base, token = {v}")

    def product():
        nonlocal pos
        e = atom()
        while pos < len(toks) and not (at(")") or at("+") or at("-")):
            if at("*") or at("/"):
                op = toks[pos][1]; pos += 1
                e = (op, e, atom())
            else:
                raise ValueError("Unexpected token in parse_expression_factors.")
        return e

    def total():
        nonlocal pos
        e = product()
        while pos < len(toks) and not at(")"):
            if at("+") or at("-"):
                op = toks[pos][1]; pos += 1
                e = (op, e, product())
            else:
                raise ValueError("Unexpected token in parse_expression_terms.")
        return e

    e = total()
    if pos != len(toks):
        raise ValueError("Unconsumed tokens.")
    return e


def evaluate(tree, bindings):
    kind = tree[0]
    if kind == "c":
        return tree[1]
    if kind == "v":
        if tree[1] not in bindings:
            raise ValueError(f"Could not find variable binding for '{tree[1]}'")
        return bindings[tree[1]]
    a, b = evaluate(tree[1], bindings), evaluate(tree[2], bindings)
    if kind == "+":
        return a + b
    if kind == "-":
        return a - b
    if kind == "*":
        return a * b
    if b == 0:
        raise ValueError("Divide by 0")
    return a / b


def variables(tree, out=None):
    out = set() if out is None else out
    if tree[0] == "v":
        out.add(tree[1])
    elif tree[0] != "c":
        variables(tree[1], out)
        variables(tree[2], out)
    return out


def text(tree):
    """The reference's text form: binary operations in parentheses, constants
    as std::to_string writes them."""
    if tree[0] == "c":
        return f"{tree[1]:f}"
    if tree[0] == "v":
        return tree[1]
    return f"({text(tree[1])}{tree[0]}{text(tree[2])})"


SPACE = ("x", "y", "z", "ux", "uy", "uz")


def cell_bindings(c, dims):
    """Space variables of cell c = (cx, cy, cz) in a box of dims cells
    (naive_dpd_engine.hpp:168-173)."""
    return {"x": float(c[0]), "y": float(c[1]), "z": float(c[2]),
            "ux": (c[0] + 0.5) / dims[0], "uy": (c[1] + 0.5) / dims[1], "uz": (c[2] + 0.5) / dims[2]}
