"""Runner of the reference's known-answer cases (tests/golden/reference_known_answers.json,
written by tests/golden/make_known_answers.py): builds the case's Fluid with `make`, applies
its set-up ops and evaluates its checks. `make` is the oracle (OracleFluid) in the CPU tests
and the device library (Fluid) in the GPU tests; both expose the reference's API."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
F = np.float32

CMP = {
    "eq": lambda a, b: a == b,
    "ne": lambda a, b: a != b,
    "lt": lambda a, b: a < b,
    "le": lambda a, b: a <= b,
    "gt": lambda a, b: a > b,
    "ge": lambda a, b: a >= b,
}


def load_cases():
    with open(os.path.join(HERE, "golden", "reference_known_answers.json")) as fh:
        return json.load(fh)["cases"]


def _edit(f, name, fn):
    a = f.get(name)
    fn(a)
    f.set(name, a)


def apply_ops(f, ops):
    for op in ops:
        kind, args = op[0], op[1:]
        if kind == "fill":
            _edit(f, args[0], lambda a: a.fill(args[1]))
        elif kind == "set":
            name, i, j, v = args
            _edit(f, name, lambda a: a.__setitem__((i, j), v))
        elif kind == "set_row":
            name, j, v = args
            _edit(f, name, lambda a: a.__setitem__((slice(None), j), v))
        elif kind == "fill_interior":
            _edit(f, args[0], lambda a: a.__setitem__((slice(1, -1), slice(1, -1)), args[1]))
        elif kind == "fill_index":
            _edit(f, args[0], lambda a: a.__setitem__(Ellipsis, np.arange(a.size).reshape(a.shape) + args[1]))
        elif kind == "open_all":
            for i in range(f.NumX):
                for j in range(f.NumY):
                    f.SetSolid(i, j, False)
        elif kind == "call":
            getattr(f, args[0])(*args[1:])
        elif kind == "knob":
            setattr(f, args[0], args[1])
        elif kind == "pass":
            f.run_pass(args[0], args[1])
        else:
            raise ValueError("unknown op %r" % (op,))


def _check(f, c, fields):
    def field(name):
        if name not in fields:
            fields[name] = f.get(name)
        return fields[name]

    kind = c[0]
    if kind == "field":
        name, i, j, op, v = c[1:]
        return CMP[op](F(field(name)[i, j]), F(v))
    if kind == "field_lt_field":
        a, i, j, b, k, l = c[1:]
        return field(a)[i, j] < field(b)[k, l]
    if kind in ("view", "view_abs", "view_near"):
        val = F(getattr(f, c[1])().Value(c[2], c[3])[0])
        if kind == "view":
            return CMP[c[4]](val, F(c[5]))
        if kind == "view_abs":
            return CMP[c[4]](abs(val), F(c[5]))
        return abs(float(val) - c[4]) <= c[5]
    if kind == "solid":
        return bool(f.IsSolid(c[1], c[2])) == c[3]
    if kind == "count_eq":
        return int((field(c[1]) == F(c[2])).sum()) == c[3]
    if kind == "count_ne":
        return int((field(c[1]) != F(c[2])).sum()) == c[3]
    if kind == "any":
        return any(_check(f, x, fields) for x in c[1])
    raise ValueError("unknown check %r" % (c,))


def run_case(make, case):
    """Returns (checks run, failures) of one case; `make(density, width, height, h)`."""
    f = make(*case["new"])
    apply_ops(f, case["ops"])
    fields = {}
    fails = []
    checks = [("ref", c) for c in case["ref"]] + [("exact", c) for c in case["exact"]]
    for tag, c in checks:
        if not _check(f, c, fields):
            got = f.get(c[1])[c[2], c[3]] if c[0] == "field" else None
            fails.append("%s %s: %r (got %r)" % (case["name"], tag, c, got))
    return len(checks), fails
