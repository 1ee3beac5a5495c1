"""Writes tests/golden/handmade.json: small hand-built instances whose optimal GST
weights follow from their construction (stated per case), used to pin the CPU oracles.

Each case: n, edges [u, v, w], groups [[members]], queries [{gids, D, expect}], where
D = -1 means no diameter bound. Run: python tests/golden/make_golden.py
"""
import json
import os

cases = []

# 1. path 0-1-2-3-4 with weights 1,2,3,4; groups at the ends and the middle.
#    Any tree covering {0} and {4} contains the whole path: 1+2+3+4 = 10.
#    Hop: the path has diameter 4, so D < 4 has no solution.
cases.append(dict(
    name="path5", n=5, edges=[[0, 1, 1], [1, 2, 2], [2, 3, 3], [3, 4, 4]],
    groups=[[0], [4], [2], [1, 3]],
    queries=[dict(gids=[0, 1], D=-1, expect=10), dict(gids=[0, 2], D=-1, expect=3),
             dict(gids=[2, 3], D=-1, expect=2),   # vertex 2 to vertex 1 (w=2) or 3 (w=3)
             dict(gids=[0, 1, 2, 3], D=-1, expect=10),
             dict(gids=[0, 1], D=4, expect=10), dict(gids=[0, 1], D=3, expect=100000),
             dict(gids=[3], D=-1, expect=0)]))

# 2. star: centre 0, leaves 1..5 with weights 1..5, plus a heavy chord 1-2 (w=100).
#    Covering leaves {1},{2},{3}: star edges 1+2+3 = 6 (chord never helps). D=2 ok.
cases.append(dict(
    name="star", n=6, edges=[[0, 1, 1], [0, 2, 2], [0, 3, 3], [0, 4, 4], [0, 5, 5], [1, 2, 100]],
    groups=[[1], [2], [3], [4], [5], [1, 5]],
    queries=[dict(gids=[0, 1, 2], D=-1, expect=6), dict(gids=[0, 1, 2, 3, 4], D=-1, expect=15),
             dict(gids=[0, 1, 2], D=2, expect=6), dict(gids=[0, 1, 2], D=1, expect=100000),
             dict(gids=[5, 1], D=-1, expect=2)]))   # group {1,5} already holds 1, +edge 0-2 +0-1 = 3? see below

# fix the last expectation: groups[5] = {1,5} and groups[1] = {2}: cheapest tree joins
# vertex 2 to vertex 1 (via 0: 2+1 = 3) or to vertex 5 (via 0: 2+5 = 7) -> 3.
cases[-1]["queries"][-1]["expect"] = 3

# 3. square with diagonal: 0-1 (1), 1-2 (1), 2-3 (1), 3-0 (1), 0-2 (5).
#    Groups {0},{2}: path 0-1-2 weight 2 (diagonal 5 is worse). D=1 forces the diagonal: 5.
#    All four corners: a 3-edge path (3) has diameter 3; at D=2 the only spanning trees
#    are stars, and only vertex 0 or 2 reaches the other three (1 + 1 + 5 = 7).
cases.append(dict(
    name="square", n=4, edges=[[0, 1, 1], [1, 2, 1], [2, 3, 1], [3, 0, 1], [0, 2, 5]],
    groups=[[0], [2], [1], [3]],
    queries=[dict(gids=[0, 1], D=-1, expect=2), dict(gids=[0, 1], D=1, expect=5),
             dict(gids=[0, 1, 2, 3], D=-1, expect=3), dict(gids=[0, 1, 2, 3], D=2, expect=7),
             dict(gids=[2, 3], D=-1, expect=2), dict(gids=[2, 3], D=1, expect=100000)]))

# 4. Steiner point: terminals 1,2,3 each joined to hub 0 by weight 10 and to each other
#    by weight 19. Star through the hub: 30; best terminal-only tree: 19+19 = 38.
cases.append(dict(
    name="steiner_hub", n=4, edges=[[0, 1, 10], [0, 2, 10], [0, 3, 10], [1, 2, 19], [2, 3, 19], [1, 3, 19]],
    groups=[[1], [2], [3]],
    queries=[dict(gids=[0, 1, 2], D=-1, expect=30), dict(gids=[0, 1], D=-1, expect=19),
             dict(gids=[0, 1, 2], D=1, expect=100000), dict(gids=[0, 1, 2], D=2, expect=30)]))

# 5. two components: groups in different components have no solution.
cases.append(dict(
    name="disconnected", n=4, edges=[[0, 1, 3], [2, 3, 4]],
    groups=[[0], [1], [2], [3]],
    queries=[dict(gids=[0, 1], D=-1, expect=3), dict(gids=[0, 2], D=-1, expect=100000),
             dict(gids=[2, 3], D=0, expect=100000), dict(gids=[2, 3], D=1, expect=4)]))

# 6. a vertex in every group: the answer is 0 (single-vertex tree), at any D.
cases.append(dict(
    name="common_vertex", n=3, edges=[[0, 1, 7], [1, 2, 7]],
    groups=[[1, 0], [1, 2], [1]],
    queries=[dict(gids=[0, 1, 2], D=-1, expect=0), dict(gids=[0, 1, 2], D=0, expect=0)]))

# 7. ladder 2x4: rungs weight 1, rails weight 2; groups at opposite corners and the
#    far rung.  Corners (0,0) and (1,3): rail 3 edges (6) + one rung (1) = 7.
#    vertices: top i -> i, bottom i -> 4 + i
lad = []
for i in range(4):
    lad.append([i, 4 + i, 1])
for i in range(3):
    lad.append([i, i + 1, 2])
    lad.append([4 + i, 4 + i + 1, 2])
cases.append(dict(
    name="ladder", n=8, edges=lad, groups=[[0], [7], [3], [4]],
    queries=[dict(gids=[0, 1], D=-1, expect=7), dict(gids=[0, 1, 2], D=-1, expect=7),
             dict(gids=[0, 1, 2, 3], D=-1, expect=8), dict(gids=[0, 1], D=4, expect=7),
             dict(gids=[0, 1], D=3, expect=100000)]))

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "handmade.json")
    with open(out, "w") as f:
        json.dump(cases, f, indent=1)
    print("wrote", out)
