"""Writes the reference generator's beam case (griddata.txt, spdata.txt).

The beam is the first case block of data/generate/solid/input.txt, kept here
as beam_input.txt (a data file). This script restates what
SolidGenerator::BuildData (data/generate/solid/solid_generator.cpp:93-191)
does with it and writes the two files in the generator's text formats:
  griddata.txt  WriteGlobalMeshHeader (data/generate/output_util.cpp:16-37)
                + BCOutput u, v, w (module/bc.cpp:30-43)
  spdata.txt    solid_generator.cpp:199-207 with OutputVector (module/data_io.h:49-60)
Streams are opened with ios::right | ios::scientific (module/data_io.cpp:23-28),
so doubles print as %15.6e and integers as %10d / %15d.

It is its own small parser and writer, independent of the library's, so that
the library's reader and writer can be checked against it.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BEAM_INPUT = os.path.join(HERE, "beam_input.txt")

GAUSS = {1: [0.0], 2: [-0.25, 0.25], 3: [-1 / 3, 0.0, 1 / 3], 4: [-0.375, -0.125, 0.125, 0.375]}


def parse_input(path=BEAM_INPUT):
    """The first case block as a dict: value lines are the even lines
    after the header (LoadInput reads a header line, then values)."""
    lines = open(path).read().splitlines()
    vals = [lines[i].split() for i in range(1, len(lines), 2)]
    c = {}
    c["solswitch"], c["nr_flag"], c["rstflag"], c["nlstep"] = vals[0][0], int(vals[0][1]), int(vals[0][2]), int(vals[0][3])
    c["ista"], c["iend"], c["iout"] = map(int, vals[1][:3])
    c["dt"], c["mtol"] = float(vals[2][0]), float(vals[2][1])
    c["xymin"] = [float(v) for v in vals[3][:3]]
    c["xymax"] = [float(v) for v in vals[4][:3]]
    c["xyelem"] = [int(v) for v in vals[5][:3]]
    c["idimc"] = [int(v) for v in vals[6][:3]]
    c["npxye"] = [int(v) for v in vals[7][:3]]
    c["rho"], c["nmat"] = float(vals[8][0]), int(vals[8][1])
    c["materials"] = []
    k = 9
    for _ in range(c["nmat"]):
        iprop, nprop = int(vals[k][0]), int(vals[k][1])
        c["materials"].append((iprop, [float(v) for v in vals[k + 1][:nprop]]))
        k += 2
    c["bb"] = [float(v) for v in vals[k][:3]]
    return c


def build(c):
    """BuildData for the beam: BC node lists and the Gauss-point particles
    with zp in [3.5, 4.5] and xp <= 4.0."""
    xe, ye, ze = c["xyelem"]
    nxc, nyc, nzc = (c["xyelem"][a] + c["idimc"][a] for a in range(3))
    dx = [(c["xymax"][a] - c["xymin"][a]) / c["xyelem"][a] for a in range(3)]
    npx = c["npxye"]
    volp = dx[0] * dx[1] * dx[2] / float(npx[0] * npx[1] * npx[2])
    msp = c["rho"] * volp
    ubc, vbc, wbc = [], [], []
    for k in range(nzc):
        for j in range(nyc):
            for i in range(nxc):
                nid = i + nxc * j + nxc * nyc * k
                if i == 0 or i == nxc - 1 or k == 0 or k == nzc - 1:
                    ubc.append(nid)
                if k == 0 or j == 0 or j == nyc - 1:
                    vbc.append(nid)
                if i == 0 or i == nxc - 1 or k == 0 or k == nzc - 1:
                    wbc.append(nid)
    pts = []
    for k in range(ze):
        for j in range(ye):
            for i in range(xe):
                ec = [dx[0] * (i + 0.5) + c["xymin"][0], dx[1] * (j + 0.5) + c["xymin"][1],
                      dx[2] * (k + 0.5) + c["xymin"][2]]
                for kp in range(npx[2]):
                    zp = ec[2] + GAUSS[npx[2]][kp] * dx[2]
                    for jp in range(npx[1]):
                        yp = ec[1] + GAUSS[npx[1]][jp] * dx[1]
                        for ip in range(npx[0]):
                            xp = ec[0] + GAUSS[npx[0]][ip] * dx[0]
                            if 3.5 <= zp <= 4.5 and xp <= 4.0:
                                pts.append((xp, yp, zp))
    return {"bc": (ubc, vbc, wbc), "coord": np.array(pts), "msp": msp, "volp": volp}


def e15(v):
    return "%15.6e" % v


def write(c, b, outdir):
    nxe, nye, nze = c["xyelem"]
    with open(os.path.join(outdir, "griddata.txt"), "w") as f:
        f.write("".join("%10d" % v for v in c["idimc"]) + "\n")
        f.write("".join(e15(v) for v in c["xymin"]) + "\n")
        f.write("".join(e15(v) for v in c["xymax"]) + "\n")
        f.write("%10d" % (nxe * nye * nze) + "".join("%10d" % v for v in c["xyelem"]) + "\n")
        node = [v + 1 for v in c["xyelem"]]
        f.write("%10d" % int(np.prod(node)) + "".join("%10d" % v for v in node) + "\n")
        nodec = [c["xyelem"][a] + c["idimc"][a] for a in range(3)]
        f.write("%10d" % int(np.prod(nodec)) + "".join("%10d" % v for v in nodec) + "\n")
        for name, ids in zip(("usbc", "vsbc", "wsbc"), b["bc"]):
            f.write("%15d%15s%s.ibc ----\n" % (len(ids), "! ---- ", name))
            for nid in ids:
                f.write("%15d%s\n" % (nid, e15(0.0)))
    with open(os.path.join(outdir, "spdata.txt"), "w") as f:
        n = len(b["coord"])
        f.write("%10d\n" % n)
        for p in b["coord"]:
            f.write("".join(e15(v) for v in p) + "\n")
        for i in range(n):
            f.write("%10d%10d%s%s\n" % (i, 0, e15(b["msp"]), e15(b["volp"])))


def make(outdir):
    """Writes griddata.txt and spdata.txt of the beam into outdir; returns
    (parsed input, built case)."""
    c = parse_input()
    b = build(c)
    write(c, b, outdir)
    return c, b


if __name__ == "__main__":
    make(sys.argv[1] if len(sys.argv) > 1 else ".")
