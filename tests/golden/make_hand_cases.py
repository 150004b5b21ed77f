"""Writes hand_cases.json: small computational-particle graphs for
Solver::searchneighboroctant (solver.cpp:456-570) with outputs traced by hand
from the reference's text, each with its trace.  Nothing here runs the oracle
or the library: the expected lists are literals worked out step by step
(bucketing, expansion, penalised keys, every partition of the final
quicksort, solver.cpp:600-632).

Octants (solver.cpp:464-492): the rotation is a few milliradians, so for an
offset (ex, ey, ez) whose nonzero components are at least 1 in size, the
x bit is the sign of ex (of -ey, then -ez, when ex = 0), the y bit the sign
of ey (of ex, then ez), the z bit the sign of ez (of -ey, then ex);
octant = x + 2 y + 4 z.  All graphs use lm = 9 (nq = 90), CellLength 1.

Run: python tests/golden/make_hand_cases.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def hub_case(name, center, box, bc):
    cx, cy, cz = center
    # node: (offset, label); Neigh24 of the hub in this order
    pts = {
        1: ((3, 3, 3), "A"), 2: ((1, 1, 1), "B"), 3: ((2, 2, 2), "C"),
        4: ((-2, -2, -2), "D"), 5: ((-1, -1, -1), "E"), 6: ((-3, -3, -3), "F"),
        7: ((2, -1, -1), "G"), 8: ((1, -1, -1), "H"), 9: ((-1, 1, 1), "I"),
        10: ((-2, 2, 1), "J"), 11: ((-1, 2, -1), "K"), 12: ((-1, -1, 2), "L"),
    }
    L = box[1] - box[0]

    def wrap(v):
        return (v - box[0]) % L + box[0] if bc == 1 else v

    x, y, z = [cx], [cy], [cz]
    for k in range(1, 13):
        o = pts[k][0]
        x.append(wrap(cx + o[0]))
        y.append(wrap(cy + o[1]))
        z.append(wrap(cz + o[2]))
    neigh = [[1, 4, 7, 2, 5, 11, 3, 12, 6, 9, 8, 10]] + [[0] for _ in range(12)]
    trace = [
        "row 0 at the centre; Neigh24 in order A1 D4 G7 B2 E5 K11 C3 L12 F6 I9 H8 J10;",
        "offsets: A(3,3,3) B(1,1,1) C(2,2,2) oct 7; D(-2,-2,-2) E(-1,-1,-1) F(-3,-3,-3) oct 0;",
        "G(2,-1,-1) H(1,-1,-1) oct 1; K(-1,2,-1) oct 2; L(-1,-1,2) oct 4; I(-1,1,1) J(-2,2,1) oct 6",
        "buckets (insertion order): o0 D E F | o1 G H | o2 K | o4 L | o6 I J | o7 A B C; l = 12 >= 9",
        "octant 0 alone is sorted (solver.cpp:557-561): E(3) D(12) F(27); octants 1-7 stay in",
        "insertion order; key = d^2 * 1e4^rank:",
        "pos 0 E 3 | 1 D 1.2e5 | 2 F 2.7e9 | 3 G 6 | 4 H 3e4 | 5 K 6 | 6 L 6 | 7 I 3 | 8 J 9e4 |",
        "9 A 27 | 10 B 3e4 | 11 C 1.2e9",
        "part(0,11) pivot 3: i=1, j=7, swap 1<->7; i=2, j=1: stop; swap 0<->1 -> I E F G H K L D J A B C, j=1",
        "part(2,11) pivot 2.7e9: i runs to 12, j=11; swap 2<->11 -> ... C ... F at 11, j=11",
        "part(2,10) pivot 1.2e9: i to 11, j=10; swap 2<->10 -> I E B G H K L D J A C F",
        "part(2,9) pivot 3e4 (B): i=7 (D), j=9 (A): swap -> pos7 A, pos9 D; i=8 (J), j=7: stop;",
        "swap 2<->7 -> I E A G H K L B J D C F, j=7",
        "part(2,6) pivot 27 (A): i=4 (H), j=6 (L): swap -> L at 4, H at 6; i=6, j=5: stop; swap 2<->5",
        "-> I E K G L A H B J D C F, j=5",
        "part(2,4) pivot 6 (K): i to 5, j=4: swap 2<->4 -> L G K; part(2,3) pivot 6 (L): swap 2<->3",
        "-> G L K",
        "part(8,9) pivot 9e4: i=9, j=8: stop, no move",
        "result: I E G L K A H B J D C F; the first 9 are the output",
        "a plain 9-nearest list would hold B, H, E, I (d^2 3), G K L (6), J (9) and D or C (12), not A;",
        "sorting octant 7 would drop A (rank 2); leaving octant 0 unsorted would take D for E;",
        "a stable sort would give E before I and K before L",
    ]
    return dict(
        name=name, bc=bc, bbox=[box[0], box[1]] * 3, lm=9, nc=13, nv=0, nb=0,
        x=x, y=y, z=z, h=[1.0] * 13, np=[1] * 13, neigh=neigh,
        rows={"0": dict(nbr=[9, 5, 7, 12, 11, 1, 8, 2, 10], d2=[3, 3, 6, 6, 6, 27, 3, 3, 9])},
        error=None, trace=trace)


def expansion_case():
    c = (8.0, 8.0, 8.0)
    offs = {1: (1, 1, 1), 2: (-1, -1, -1), 3: (1, 1, 2), 4: (3, 3, 3), 5: (2, -1, -1),
            6: (-1, 2, -1), 7: (-1, -1, 2), 8: (1, 1, -1), 9: (1, -1, 1), 10: (-1, 1, 1)}
    x, y, z = [c[0]], [c[1]], [c[2]]
    for k in range(1, 11):
        x.append(c[0] + offs[k][0])
        y.append(c[1] + offs[k][1])
        z.append(c[2] + offs[k][2])
    neigh = [[1, 2],                      # 0: P (oct 7) then Q (oct 0)
             [3, 4, 0, 8, 9, 10],         # 1 = P: Xa, Xb, the row itself, Y4, Y5, Y6
             [4, 0, 5, 6, 7],             # 2 = Q: Xb, the row itself, Y1, Y2, Y3
             [0], [0], [0], [0], [0], [0], [0], [0]]
    trace = [
        "row 0; P=1 (1,1,1) oct 7, Q=2 (-1,-1,-1) oct 0; Neigh24(0) = P, Q: l = 2 < 9",
        "Xa=3 (1,1,2) and Xb=4 (3,3,3) oct 7; Y1=5 (2,-1,-1) oct 1; Y2=6 (-1,2,-1) oct 2;",
        "Y3=7 (-1,-1,2) oct 4; Y4=8 (1,1,-1) oct 3; Y5=9 (1,-1,1) oct 5; Y6=10 (-1,1,1) oct 6",
        "round 1 walks the snapshot in octant order (solver.cpp:499-507): Q first, then P",
        "Q: Xb new -> o7 [P, Xb]; 0 is the row itself (skipped); Y1, Y2, Y3 new",
        "P: Xa new -> o7 [P, Xb, Xa]; Xb already listed; 0 skipped; Y4, Y5, Y6 new",
        "l = 10 >= 9: stop.  keys in octant order: Q 3 | Y1 6 | Y2 6 | Y4 3 | Y3 6 | Y5 3 | Y6 3 |",
        "P 3 | Xb 27e4 | Xa 6e8",
        "part(0,9) pivot 3: swaps 1<->7, 2<->6, 4<->5, then 0<->4 -> Y5 P Y6 Y4 Q Y3 Y2 Y1 Xb Xa, j=4",
        "part(0,3): i to 4, j=3: swap 0<->3 -> Y4 P Y6 Y5; part(0,2): swap 0<->2 -> Y6 P Y4;",
        "part(0,1): swap 0<->1 -> P Y6",
        "part(5,9) pivot 6 (Y3): i=8, j=7: stop; swap 5<->7 -> Y1 Y2 Y3; part(5,6): swap -> Y2 Y1",
        "part(8,9): no move",
        "result: P Y6 Y4 Y5 Q Y2 Y1 Y3 Xb Xa; the first 9 are the output",
        "walking the snapshot in insertion order (P before Q) would make Xa rank 1 (key 6e4) and",
        "put Xa in the list instead of Xb; counting Xb twice or the row itself would change l",
    ]
    return dict(
        name="expansion_order", bc=0, bbox=[0.0, 16.0] * 3, lm=9, nc=11, nv=0, nb=0,
        x=x, y=y, z=z, h=[1.0] * 11, np=[1] * 11, neigh=neigh,
        rows={"0": dict(nbr=[1, 10, 8, 9, 2, 6, 5, 7, 4], d2=[3, 3, 3, 3, 3, 6, 6, 6, 27])},
        error=None, trace=trace)


def few_case():
    return dict(
        name="not_enough_neighbours", bc=0, bbox=[0.0, 16.0] * 3, lm=9, nc=2, nv=0, nb=0,
        x=[1.0, 2.0], y=[1.0, 2.0], z=[1.0, 2.0], h=[1.0, 1.0], np=[1, 1], neigh=[[1], [0]],
        rows={}, error="APC_ERR_NEIGH_FEW",
        trace=["row 0: Neigh24 [1]; every round walks 1, whose only entry is row 0 itself;",
               "l stays 1 < 9; after the sixth round ite = 6 > 5: 'Not enough neighbours'",
               "(solver.cpp:544-550); row 1 likewise"])


def many_case():
    # row 0 at (1,1,1); ring nodes 1..8 at offsets (1,1,k); outer nodes
    # 9 + 24 (k-1) + s at offsets (2+s, 2+k, 4): everything in octant 7
    x, y, z = [1.0], [1.0], [1.0]
    for k in range(1, 9):
        x.append(2.0); y.append(2.0); z.append(1.0 + k)
    for k in range(1, 9):
        for s in range(24):
            x.append(3.0 + s); y.append(3.0 + k); z.append(5.0)
    neigh = [list(range(1, 9))]
    for k in range(1, 9):
        neigh.append([9 + 24 * (k - 1) + s for s in range(24)])
    for k in range(1, 9):
        for s in range(24):
            neigh.append([k])
    n = len(x)
    return dict(
        name="too_many_in_an_octant", bc=0, bbox=[0.0, 64.0] * 3, lm=9, nc=n, nv=0, nb=0,
        x=x, y=y, z=z, h=[1.0] * n, np=[1] * n, neigh=neigh, rows={},
        error="APC_ERR_NEIGH_MANY",
        trace=["row 0: Neigh24 = the 8 ring nodes, all octant 7: l = 8 < 9",
               "round 1 walks them; each names 24 distinct outer nodes, all octant 7 and new;",
               "the 83rd makes octant 7 hold 91 > nq = 90 entries: 'Too many neighbours'",
               "(solver.cpp:535-539); ring rows (24 entries) and outer rows (23 found in one",
               "round) succeed"])


def invalid_case():
    return dict(
        name="invalid_neighbour", bc=0, bbox=[0.0, 16.0] * 3, lm=9, nc=2, nv=0, nb=0,
        x=[1.0, 2.0], y=[1.0, 2.0], z=[1.0, 2.0], h=[1.0, 1.0], np=[1, 1],
        neigh=[[1], [0, 99]], rows={}, error="APC_ERR_NEIGH_INVALID",
        trace=["row 0: Neigh24 [1]; round 1 walks 1's entries 0, 99; 99 > nc+nv+nb = 2:",
               "'Invalid neighbour!' (solver.cpp:511-518)"])


def main():
    cases = [hub_case("octant_order_and_penalty", (8.0, 8.0, 8.0), (0.0, 16.0), 0),
             hub_case("octant_order_periodic", (0.5, 8.0, 15.5), (0.0, 16.0), 1),
             expansion_case(), few_case(), many_case(), invalid_case()]
    with open(os.path.join(HERE, "hand_cases.json"), "w") as f:
        json.dump({"source": "tests/golden/make_hand_cases.py", "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
