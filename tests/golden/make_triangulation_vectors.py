"""Writes tests/golden/triangulation_vectors.json: polygons and the index
triples LiAR's polygon triangulation returns for them
(src/liar/tools/geometry.py:4-93, triangulate_3d), which decides how
liar.tools.wavefront_obj splits OBJ faces.  The module is pure Python with
no compiled dependencies; it is loaded from the reference tree by file path.
Run from the repository root:  python tests/golden/make_triangulation_vectors.py
"""
import importlib.util
import json
import math
import random
from pathlib import Path

SRC = Path("/root/reference/src/liar/tools/geometry.py")
OUT = Path(__file__).resolve().parent / "triangulation_vectors.json"

spec = importlib.util.spec_from_file_location("liar_geometry", SRC)
geo = importlib.util.module_from_spec(spec)
spec.loader.exec_module(geo)


def polygons():
    yield "quad_z_up", [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    yield "quad_z_down", [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)]
    yield "cube_face", [(-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1)]
    yield "pentagon_x", [(0, math.cos(a), math.sin(a)) for a in [2 * math.pi * k / 5 for k in range(5)]]
    yield "concave_L", [(0, 0, 0), (2, 0, 0), (2, 1, 0), (1, 1, 0), (1, 2, 0), (0, 2, 0)]
    yield "concave_L_down", [(0, 0, 0), (0, 2, 0), (1, 2, 0), (1, 1, 0), (2, 1, 0), (2, 0, 0)]
    yield "arrow_y", [(0, 0, 0), (2, 0, 1), (0, 0, 2), (0.5, 0, 1)]
    yield "degenerate_line", [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0)]
    yield "triangle", [(0, 0, 0), (1, 0, 0), (0, 0, 1)]
    rng = random.Random(7)
    for k in range(12):
        # star-shaped polygons with random radii in a random plane orientation
        n = rng.randint(4, 9)
        pts2 = []
        for q in range(n):
            a = 2 * math.pi * q / n
            r = rng.uniform(0.4, 1.0)
            pts2.append((r * math.cos(a), r * math.sin(a)))
        axis = rng.randint(0, 2)
        sign = rng.choice([1, -1])
        pts = []
        for x, y in pts2:
            p = [0.0, 0.0, 0.0]
            p[(axis + 1) % 3], p[(axis + 2) % 3] = x, y * sign
            p[axis] = 0.01 * x  # slightly tilted
            pts.append(tuple(p))
        yield f"star_{k}", pts


def main():
    out = []
    for name, poly in polygons():
        tris = [list(t) for t in geo.triangulate_3d(*poly)]
        out.append({"name": name, "polygon": [list(map(float, p)) for p in poly], "triangles": tris})
    OUT.write_text(json.dumps(out, indent=1))
    print(f"wrote {len(out)} cases to {OUT}")


if __name__ == "__main__":
    main()
