"""Writes the small input fixtures under tests/golden/ from the reference's own
test inputs (data only; run in the build container, where the reference tree
is mounted at /root/reference).

neohookean_cube2_elem0.json -- inputs of test/test_neohookean_accumulate.cpp:
  element 0 of resource/cube2.msh (first tetrahedron of the $Elements block,
  gmsh node tags -> 0-based vertices), lambda = 1e7, mu = 1e6, kd = 10,
  V = 5 (test :37), vertex perturbations (test :45-48), dt = 1e-3 (test :50).
"""
import json
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def read_msh41(path):
    lines = path.read_text().split("\n")
    i = lines.index("$Nodes")
    n_blocks = int(lines[i + 1].split()[0])
    k = i + 2
    coords = {}
    for _ in range(n_blocks):
        dim, tag, param, n = map(int, lines[k].split())
        tags = [int(t) for t in lines[k + 1:k + 1 + n]]
        xyz = [list(map(float, l.split())) for l in lines[k + 1 + n:k + 1 + 2 * n]]
        coords.update(zip(tags, xyz))
        k += 1 + 2 * n
    j = lines.index("$Elements")
    n_blocks = int(lines[j + 1].split()[0])
    k = j + 2
    tets = []
    for _ in range(n_blocks):
        dim, tag, etype, n = map(int, lines[k].split())
        rows = [list(map(int, l.split())) for l in lines[k + 1:k + 1 + n]]
        if etype == 4:
            tets += [r[1:5] for r in rows]
        k += 1 + n
    return coords, tets


def main():
    coords, tets = read_msh41(REF / "resource" / "cube2.msh")
    elem = tets[0]
    rest = [coords[t] for t in elem]
    fx = {
        "source": "test/test_neohookean_accumulate.cpp:19-50, resource/cube2.msh element 0",
        "element_node_tags": elem,
        "rest_positions": rest,
        "perturbation": [[0.3, 0.3, 0.3], [0.1, 0.2, 0.3], [-0.2, -0.3, -0.2], [-0.1, 0.1, 0.1]],
        "lambda": 1e7,
        "mu": 1e6,
        "kd": 10.0,
        "V": 5.0,
        "dt": 1e-3,
    }
    (OUT / "neohookean_cube2_elem0.json").write_text(json.dumps(fx, indent=1) + "\n")


if __name__ == "__main__":
    main()
