"""Writes the small Matrix Market fixtures the tests read (deterministic,
no reference code involved):

  stencil5_8x8_sym.mtx   2D 5-point Laplacian on 8x8, symmetric header,
                         lower triangle stored column-major (the SuiteSparse
                         layout), real values 4 / -1
  er_96_general.mtx      96x96 Erdos-Renyi pattern, general header, entries
                         column-major, integer values 1..9
  er_80x96_general.mtx   80x96 rectangular real general, row-major with the
                         rows' entries shuffled (unsorted columns)

Run: python tests/golden/make_mtx_fixtures.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def stencil5(nx, ny):
    ent = {}
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            ent[(i, i)] = 4.0
            for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                xx, yy = x + dx, y + dy
                if 0 <= xx < nx and 0 <= yy < ny:
                    ent[(i, xx + nx * yy)] = -1.0
    return nx * ny, ent


def write(path, header, m, n, entries, fmt):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {header}\n")
        f.write("% fixture written by make_mtx_fixtures.py\n")
        f.write(f"{m} {n} {len(entries)}\n")
        for r, c, v in entries:
            f.write(fmt(r + 1, c + 1, v) + "\n")


def main():
    n, ent = stencil5(8, 8)
    low = sorted(((r, c, v) for (r, c), v in ent.items() if r >= c), key=lambda t: (t[1], t[0]))
    write(os.path.join(HERE, "stencil5_8x8_sym.mtx"), "real symmetric", n, n, low,
          lambda r, c, v: f"{r} {c} {v:g}")

    rng = np.random.default_rng(2026)
    m = 96
    mask = rng.random((m, m)) < 0.06
    rows, cols = np.nonzero(mask)
    order = np.lexsort((rows, cols))
    vals = rng.integers(1, 10, size=len(rows))
    write(os.path.join(HERE, "er_96_general.mtx"), "integer general", m, m,
          [(int(rows[k]), int(cols[k]), int(vals[k])) for k in order], lambda r, c, v: f"{r} {c} {v}")

    m, n = 80, 96
    mask = rng.random((m, n)) < 0.07
    ent = []
    for r in range(m):
        cs = np.nonzero(mask[r])[0]
        rng.shuffle(cs)
        ent += [(r, int(c), float(rng.uniform(-1, 1))) for c in cs]
    write(os.path.join(HERE, "er_80x96_general.mtx"), "real general", m, n, ent,
          lambda r, c, v: f"{r} {c} {v:.17g}")


if __name__ == "__main__":
    main()
