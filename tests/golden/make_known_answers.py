"""Writes reference_known_answers.json: the known answers of the reference's own tests for the
simulation path, as data (inputs, then expected outputs), run against the oracle on the CPU and
against the device library on the GPU by tests/known_answers.py.

Every case names the Go test it restates (pkg/fluid/*_test.go, file:line). `ref` checks are
the reference test's own assertions; `exact` checks are the closed forms the mathematics
fixes for the same state (float32 arithmetic in the reference's operation order, worked out
here from the formulas in pkg/fluid/fluid.go and walls.go, not from any implementation).
The reference is Go and no Go toolchain exists here, so nothing in this file is produced by
running the reference; it is transcribed.

    python3 tests/golden/make_known_answers.py   # rewrites the JSON next to this file
"""
import json
import math
import os

import numpy as np

F = np.float32


def f32(x):
    return float(F(x))


def exp_weight(d2, r2):
    # ApplyForceRadius (fluid.go:774-796): float32(math.Exp(float64(-3.0 * dist2 / r2)))
    return F(math.exp(float(F(F(-3.0) * F(d2)) / F(r2))))


def cases():
    out = []

    # TestCopyBorder (fluid_test.go:8-35): New(1.0, 3, 3, 1.0), src[c] = c + 1, copyBorder.
    nx = ny = 5
    ring = [(i, j) for i in range(nx) for j in range(ny) if i in (0, nx - 1) or j in (0, ny - 1)]
    out.append(dict(
        name="TestCopyBorder", reference="pkg/fluid/fluid_test.go:8-35",
        new=[1.0, 3, 3, 1.0],
        ops=[["fill_index", "U", 1.0], ["fill_index", "V", 1.0], ["pass", "copyBorderUV", 0.0]],
        ref=[["field", "newU", i, j, "eq", float(i * ny + j + 1)] for (i, j) in ring] +
            [["field", "newV", i, j, "eq", float(i * ny + j + 1)] for (i, j) in ring],
        exact=[["field", "newU", i, j, "eq", 0.0] for i in range(1, nx - 1) for j in range(1, ny - 1)],
    ))

    # TestApplyForce (fluid_test.go:1134-1160): all S = 1, force on (5,5), then on a solid
    # cell (6,6) and on the border cell (0,0), both no-ops.
    out.append(dict(
        name="TestApplyForce", reference="pkg/fluid/fluid_test.go:1134-1160",
        new=[1.0, 10, 10, 1.0],
        ops=[["fill", "S", 1.0], ["call", "ApplyForce", 5, 5, 3.0, -2.0], ["set", "S", 6, 6, 0.0],
             ["call", "ApplyForce", 6, 6, 10.0, 10.0], ["call", "ApplyForce", 0, 0, 1.0, 1.0]],
        ref=[["field", "U", 5, 5, "eq", 3.0], ["field", "V", 5, 5, "eq", -2.0],
             ["field", "U", 6, 6, "eq", 0.0], ["field", "V", 6, 6, "eq", 0.0]],
        exact=[["field", "U", 0, 0, "eq", 0.0], ["field", "V", 0, 0, "eq", 0.0],
               ["count_ne", "U", 0.0, 1], ["count_ne", "V", 0.0, 1]],
    ))

    # TestApplyForceRadius (fluid_test.go:1162-1190): all S = 1, ApplyForceRadius(10, 10, 5, 0, 3).
    r2 = 9
    exact = []
    n_in = 0
    for a in range(-3, 4):
        for b in range(-3, 4):
            d2 = a * a + b * b
            if d2 <= r2:
                n_in += 1
                exact.append(["field", "U", 10 + a, 10 + b, "eq", float(F(5.0) * exp_weight(d2, r2))])
    out.append(dict(
        name="TestApplyForceRadius", reference="pkg/fluid/fluid_test.go:1162-1190",
        new=[1.0, 20, 20, 1.0],
        ops=[["fill", "S", 1.0], ["call", "ApplyForceRadius", 10, 10, 5.0, 0.0, 3]],
        ref=[["field", "U", 10, 10, "ge", 4.5], ["field", "U", 13, 10, "gt", 0.0],
             ["field_lt_field", "U", 13, 10, "U", 10, 10], ["field", "U", 14, 10, "eq", 0.0]],
        exact=exact + [["count_ne", "U", 0.0, n_in], ["count_ne", "V", 0.0, 0]],
    ))

    # TestVorticityField (fluid_test.go:1381-1402): a 4-face vortex around (5,5), h = 1.
    # curl = (V[6,5]-V[4,5])*0.5/h - (U[5,6]-U[5,4])*0.5/h = 1 - (-1) = 2 exactly. Each face
    # also enters the curl of the cell on its far side: (5,3), (5,7), (3,5), (7,5) get -0.5;
    # the cells between, (5,4) and (4,5), see no face difference.
    out.append(dict(
        name="TestVorticityField", reference="pkg/fluid/fluid_test.go:1381-1402",
        new=[1.0, 10, 10, 1.0],
        ops=[["fill", "S", 1.0], ["set", "U", 5, 4, 1.0], ["set", "U", 5, 6, -1.0],
             ["set", "V", 4, 5, -1.0], ["set", "V", 6, 5, 1.0]],
        ref=[["view_abs", "Vorticity", 5, 5, "ge", 0.1]],
        exact=[["view", "Vorticity", 5, 5, "eq", 2.0]] +
              [["view", "Vorticity", i, j, "eq", -0.5] for (i, j) in ((5, 3), (5, 7), (3, 5), (7, 5))] +
              [["view", "Vorticity", i, j, "eq", 0.0] for (i, j) in ((5, 4), (4, 5), (3, 3))],
    ))

    # TestVelocityMagnitudeField (fluid_test.go:1404-1425): faces 3 and 4 around (5,5):
    # u = (3+3)*0.5 = 3, v = (4+4)*0.5 = 4, |(u,v)| = 5 exactly.
    out.append(dict(
        name="TestVelocityMagnitudeField", reference="pkg/fluid/fluid_test.go:1404-1425",
        new=[1.0, 10, 10, 1.0],
        ops=[["fill", "S", 1.0], ["set", "U", 5, 5, 3.0], ["set", "U", 6, 5, 3.0],
             ["set", "V", 5, 5, 4.0], ["set", "V", 5, 6, 4.0]],
        ref=[["view_near", "VelocityMagnitude", 5, 5, 5.0, 0.1]],
        exact=[["view", "VelocityMagnitude", 5, 5, "eq", 5.0],
               # (4,5) sees only U[5,5]: u = 1.5; (5,4) only V[5,5]: v = 2
               ["view", "VelocityMagnitude", 4, 5, "eq", 1.5], ["view", "VelocityMagnitude", 5, 4, "eq", 2.0]],
    ))

    # TestSetCircularObstacle (fluid_test.go:1427-1447): disc r = 3 at (10,10); the disc test
    # dx^2 + dy^2 <= r^2 (fluid.go:894-907) selects 29 cells.
    disc = [(10 + a, 10 + b) for a in range(-3, 4) for b in range(-3, 4) if a * a + b * b <= 9]
    out.append(dict(
        name="TestSetCircularObstacle", reference="pkg/fluid/fluid_test.go:1427-1447",
        new=[1.0, 20, 20, 1.0],
        ops=[["fill", "S", 1.0], ["fill", "U", 7.0], ["fill", "V", 7.0], ["fill", "newU", 7.0],
             ["fill", "newV", 7.0], ["call", "SetCircularObstacle", 10, 10, 3]],
        ref=[["solid", 10, 10, True], ["solid", 12, 10, True], ["solid", 14, 10, False]],
        # SetSolid(true) zeroes faces U[i,j], U[i+1,j], V[i,j], V[i,j+1] in U,V and newU,newV
        # (walls.go:19-48): the zeroed U faces are the disc and the disc shifted by +1 in i.
        exact=[["count_eq", "S", 0.0, len(disc)]] + [["solid", i, j, True] for (i, j) in disc] +
              [["count_eq", k, 0.0, len(set(disc) | {(i + 1, j) for (i, j) in disc})] for k in ("U", "newU")] +
              [["count_eq", k, 0.0, len(set(disc) | {(i, j + 1) for (i, j) in disc})] for k in ("V", "newV")],
    ))

    # TestBoundaryConditions (fluid_test.go:287-329): NX = NY = 8, all S = 1 but (3,3),
    # U = 1 and V = 0.5 on the interior, handleBorders. The reference's own check (U = 0 where
    # a border cell or its inner neighbour is solid) has no such cell in this state; the
    # closed form of handleBorders (fluid.go:236-289): rows j = 0 and NY-1 extrapolate
    # 2*1 - 1 = 1 for 0 < i < NX-1 and copy U[i,1] = 0 at i = 0, NX-1; columns likewise for V.
    ex = []
    for i in range(8):
        u = 1.0 if 0 < i < 7 else 0.0
        ex += [["field", "U", i, 0, "eq", u], ["field", "U", i, 7, "eq", u]]
    for j in range(8):
        v = 0.5 if 0 < j < 7 else 0.0
        ex += [["field", "V", 0, j, "eq", v], ["field", "V", 7, j, "eq", v]]
    out.append(dict(
        name="TestBoundaryConditions", reference="pkg/fluid/fluid_test.go:287-329",
        new=[1.0, 6, 6, 1.0],
        ops=[["fill", "S", 1.0], ["set", "S", 3, 3, 0.0], ["fill_interior", "U", 1.0],
             ["fill_interior", "V", 0.5], ["pass", "handleBorders", 0.0]],
        ref=[], exact=ex,
    ))
    # The same pass with the branches the reference test leaves out: a solid bottom row
    # (U = 0 on j = 0), and distinct inner values on the top rows (2*3 - 5 = 1).
    ex = [["field", "U", i, 0, "eq", 0.0] for i in range(8)]
    ex += [["field", "U", i, 7, "eq", 1.0 if 0 < i < 7 else 3.0] for i in range(8)]
    out.append(dict(
        name="TestBoundaryConditions (solid row, extrapolation)", reference="pkg/fluid/fluid.go:236-289",
        new=[1.0, 6, 6, 1.0],
        ops=[["fill", "S", 1.0], ["set_row", "S", 0, 0.0], ["fill", "U", 1.0], ["set_row", "U", 6, 3.0],
             ["set_row", "U", 5, 5.0], ["pass", "handleBorders", 0.0]],
        ref=[], exact=ex,
    ))

    # TestApplyVorticityConfinement (confinement_test.go:6-28): a 2x2 vortex, Confinement 5,
    # one confinement pass with dt = 1 must change the velocity of cell (2,2).
    out.append(dict(
        name="TestApplyVorticityConfinement", reference="pkg/fluid/confinement_test.go:6-28",
        new=[1.0, 4, 4, 1.0],
        ops=[["open_all"], ["call", "SetVelocity", 3, 2, 1.0, 0.0], ["call", "SetVelocity", 2, 3, -1.0, 0.0],
             ["call", "SetVelocity", 2, 2, 0.0, 1.0], ["call", "SetVelocity", 3, 3, 0.0, -1.0],
             ["knob", "Confinement", 5.0], ["pass", "applyVorticityConfinement", 1.0]],
        ref=[["any", [["field", "U", 2, 2, "ne", 0.0], ["field", "V", 2, 2, "ne", 1.0]]]],
        exact=[],
    ))
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_known_answers.json")
    with open(path, "w") as fh:
        json.dump({"source": "transcribed from the reference's Go tests; see make_known_answers.py",
                   "cases": cases()}, fh, indent=1)
        fh.write("\n")
    print("wrote", path)
