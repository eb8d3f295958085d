"""Decode the README figure of the reference into a numeric fixture.

The reference ships ``readmefig.svg`` (matplotlib 3.6.1 output of the README
example, ``README.md:15-60``).  Its axes hold, as marker positions and
``fill_between`` polygons:

* ``axes_1`` (title "Sequential"): ``sequential_eks_solve(fhn, linspace(0,100,300), order=3)``
  means (``line2d_14``, ``line2d_15``) and ``mean +- 2 sqrt(diag(c c^T))`` bands
  (``PolyCollection_1``, ``PolyCollection_2``);
* ``axes_2`` (title "Parallel"): ``solve(fhn, linspace(0,100,100), order=3, init="constant")``
  means (``line2d_29``, ``line2d_30``) and bands (``PolyCollection_3``, ``PolyCollection_4``).

This script reads the SVG as text (data, not code), maps pixel coordinates
to data coordinates through the tick positions, and writes
``readme_fhn.npz``.  Pixel positions carry 6 decimals of a point, i.e.
about 5e-8 data units on the y axis.

Usage: ``python tests/golden/make_readme_fixture.py /root/reference/readmefig.svg``
"""
import re
import sys
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

NS = {"svg": "http://www.w3.org/2000/svg", "xlink": "http://www.w3.org/1999/xlink"}
XL = "{http://www.w3.org/1999/xlink}href"


def _find(root, gid):
    for g in root.iter("{http://www.w3.org/2000/svg}g"):
        if g.get("id") == gid:
            return g
    raise KeyError(gid)


def _ticks(axes, kind):
    out = []
    for g in axes.iter("{http://www.w3.org/2000/svg}g"):
        gid = g.get("id") or ""
        if gid.startswith(kind + "tick_"):
            use = next(g.iter("{http://www.w3.org/2000/svg}use"))
            out.append((float(use.get("x")), float(use.get("y")), gid))
    return out


def _markers(line):
    pts = []
    for use in line.iter("{http://www.w3.org/2000/svg}use"):
        pts.append((float(use.get("x")), float(use.get("y"))))
    return np.array(pts)


def _polygon(root, poly):
    path = next(poly.iter("{http://www.w3.org/2000/svg}path"))
    use = next(poly.iter("{http://www.w3.org/2000/svg}use"))
    ox, oy = float(use.get("x")), float(use.get("y"))
    nums = re.findall(r"[ML]\s+(-?[0-9.]+)\s+(-?[0-9.]+)", path.get("d"))
    return np.array([(float(a) + ox, float(b) + oy) for a, b in nums])


def _tick_values(svg_text, axes_start, kind, n):
    # tick labels are emitted as XML comments "<!-- 20 -->" right after each tick
    seg = svg_text[axes_start:]
    vals = []
    for m in re.finditer(r'id="%stick_\d+".*?<!-- (.*?) -->' % kind, seg, flags=re.S):
        vals.append(float(m.group(1).replace("−", "-")))
        if len(vals) == n:
            break
    return vals


def decode(svg_path):
    text = Path(svg_path).read_text()
    root = ET.fromstring(text)
    out = {}
    for axes_id, lines, polys, tag in (
        ("axes_1", ("line2d_14", "line2d_15"), ("PolyCollection_1", "PolyCollection_2"), "seq"),
        ("axes_2", ("line2d_29", "line2d_30"), ("PolyCollection_3", "PolyCollection_4"), "par"),
    ):
        axes = _find(root, axes_id)
        start = text.index('id="%s"' % axes_id)
        xt = _ticks(axes, "x")
        yt = _ticks(axes, "y")
        xv = _tick_values(text, start, "x", len(xt))
        yv = _tick_values(text, start, "y", len(yt))
        ax = np.polyfit([p[0] for p in xt], xv, 1)
        ay = np.polyfit([p[1] for p in yt], yv, 1)
        means = []
        for lid in lines:
            pts = _markers(_find(root, lid))
            means.append((np.polyval(ax, pts[:, 0]), np.polyval(ay, pts[:, 1])))
        t = means[0][0]
        out[tag + "_t"] = t
        out[tag + "_means"] = np.stack([m[1] for m in means], axis=1)
        bands = []
        for pid in polys:
            v = _polygon(root, _find(root, pid))
            n = len(t)
            # fill_between vertex order: (x0, y2_0), (x_i, y1_i) i=0..n-1, (x_n-1, y2_n-1), y2 reversed
            assert len(v) == 2 * n + 2, (pid, len(v), n)
            y1 = np.polyval(ay, v[1 : n + 1, 1])
            y2 = np.polyval(ay, v[n + 2 : 2 * n + 2, 1])[::-1]
            bands.append(np.stack([y1, y2], axis=-1))
        out[tag + "_bands"] = np.stack(bands, axis=1)  # [n, d, 2]
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/readmefig.svg"
    data = decode(src)
    dst = Path(__file__).with_name("readme_fhn.npz")
    np.savez(dst, **data)
    for k, v in data.items():
        print(k, v.shape)
