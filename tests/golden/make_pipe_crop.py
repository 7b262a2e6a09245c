"""Writes tests/golden/pipe_crop.vtk.gz: a block of the reference's own fluid
mask build/scenarios/pipe.vtk (legacy ASCII STRUCTURED_POINTS, 250 x 54 x 54
points, 408,838 fluid) around the pipe's constriction, x = 80 .. 119 with the
full 54 x 54 cross-section, in the same format: the header keeps the file's
blank lines and padded keyword columns, ORIGIN is moved to the block's first
point (origin + 80 * spacing along x), SPACING is kept, and the point data are
the mask values in the file's x-fastest order, 20 to a line.

    python tests/golden/make_pipe_crop.py [/root/reference/build/scenarios/pipe.vtk]

Parsing here is deliberately independent of the readers under test: the
header is split into lines, the data are every whitespace-separated token
after the LOOKUP_TABLE line.
"""
from __future__ import annotations

import gzip
import sys
from pathlib import Path

X0, X1 = 80, 120  # cropped x range [X0, X1) of the point grid


def main(src: str) -> None:
    lines = Path(src).read_text().split("\n")
    head = {}
    data_at = None
    for i, line in enumerate(lines):
        t = line.split()
        if not t:
            continue
        if t[0] in ("DIMENSIONS", "ORIGIN", "SPACING"):
            head[t[0]] = t[1:4]
        if t[0] == "LOOKUP_TABLE":
            data_at = i + 1
            break
    nx, ny, nz = (int(v) for v in head["DIMENSIONS"])
    ox, oy, oz = (float(v) for v in head["ORIGIN"])
    sx, sy, sz = (float(v) for v in head["SPACING"])
    vals = " ".join(lines[data_at:]).split()
    assert len(vals) == nx * ny * nz, (len(vals), nx * ny * nz)
    crop = [vals[x + nx * (y + ny * z)] for z in range(nz) for y in range(ny) for x in range(X0, X1)]
    cx = X1 - X0
    out = ["# vtk DataFile Version 2.0",
           f"block x = {X0} .. {X1 - 1} of build/scenarios/pipe.vtk",
           "ASCII",
           "",
           "DATASET STRUCTURED_POINTS",
           f"DIMENSIONS    {cx}   {ny}   {nz}",
           "",
           f"ORIGIN    {ox + X0 * sx:.6f}   {oy:.6f}   {oz:.6f}",
           f"SPACING   {sx:.6f}   {sy:.6f}   {sz:.6f}",
           "",
           f"POINT_DATA   {cx * ny * nz}",
           "SCALARS inputfluidMask unsigned_char",
           "LOOKUP_TABLE default",
           ""]
    out += [" ".join(crop[i:i + 20]) + " " for i in range(0, len(crop), 20)]
    dst = Path(__file__).resolve().parent / "pipe_crop.vtk.gz"
    with gzip.GzipFile(dst, "wb", mtime=0) as f:
        f.write(("\n".join(out) + "\n").encode())
    print(f"{dst}: {cx} x {ny} x {nz} points, {sum(v != '0' for v in crop)} fluid")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/build/scenarios/pipe.vtk")
