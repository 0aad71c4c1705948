"""Writes lattice_counts.json: particle and directed-bond counts of the lattice
specimens of the benchmark workloads, by shift-overlap counting (for every
offset xi with 0 < |xi| < horizon, the number of particles p with p + xi in the
specimen), independent of the cell-list family builds it checks.

    python tests/golden/make_lattice_counts.py
"""
import json
import math
import os

import numpy as np

H = 3.015


def offsets(h):
    r = int(math.ceil(h))
    out = []
    for a in range(-r, r + 1):
        for b in range(-r, r + 1):
            for c in range(-r, r + 1):
                d2 = a * a + b * b + c * c
                if 0 < d2 < h * h:
                    out.append((a, b, c))
    return out


def box_bonds(n1, n2, n3, h):
    return sum(max(n1 - abs(a), 0) * max(n2 - abs(b), 0) * max(n3 - abs(c), 0) for a, b, c in offsets(h))


def mask_bonds(mask, h):
    """mask: boolean occupancy grid (any dimension 3); bonds = sum over offsets
    of |mask AND shift(mask)|."""
    tot = 0
    for off in offsets(h):
        sl_a, sl_b = [], []
        for o, n in zip(off, mask.shape):
            sl_a.append(slice(max(o, 0), n + min(o, 0)))
            sl_b.append(slice(max(-o, 0), n + min(-o, 0)))
        tot += int(np.count_nonzero(mask[tuple(sl_a)] & mask[tuple(sl_b)]))
    return tot


def disk_mask(r, t):
    ri = int(math.floor(r))
    ax = np.arange(-ri, ri + 1)
    Y, X = np.meshgrid(ax, ax, indexing="ij")
    m2 = (X * X + Y * Y) <= r * r
    return np.repeat(m2[None, :, :], t, axis=0)


def sphere_mask(r):
    ri = int(math.floor(r))
    ax = np.arange(-ri, ri + 1)
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    return (X * X + Y * Y + Z * Z) <= r * r


def main():
    out = {"horizon": H, "interior_degree": len(offsets(H))}
    out["tensile_box_20"] = {"N": 20 ** 3, "nnz": box_bonds(20, 20, 20, H)}
    out["plate_50x50x20"] = {"N": 50 * 50 * 20, "nnz": box_bonds(50, 50, 20, H)}
    d = disk_mask(150.0, 10)
    out["disk_r150_t10"] = {"N": int(d.sum()), "nnz": mask_bonds(d, H)}
    s = sphere_mask(15.0)
    out["sphere_r15"] = {"N": int(s.sum()), "nnz": mask_bonds(s, H)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattice_counts.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
