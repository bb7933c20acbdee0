"""Reader of the reference's ray / Frechet-row dump (oracle/raydump.inc):
test infrastructure only."""
from __future__ import annotations

import numpy as np


def read_raydump(path: str, max_rays: int | None = None):
    """Records of the first forward in the dump: a list of dicts with keys
    src, rcv (1-based), path, valid, t, nsec, sections [(npts, tf, first, last,
    points or None)], pdev (indices 1-based int array, values)."""
    rays = []
    seen = set()
    cur = None
    with open(path) as f:
        for ln in f:
            tag = ln[0]
            if tag == "R":
                p = ln[1:].split()
                key = (int(p[0]), int(p[1]))
                if key in seen:
                    break  # next forward of the inversion
                seen.add(key)
                if max_rays is not None and len(rays) >= max_rays:
                    break
                cur = {"src": int(p[0]), "rcv": int(p[1]), "path": int(p[2]), "valid": p[3] == "T",
                       "t": float(p[4]), "nsec": int(p[5]), "npdev": int(p[6]), "sections": [], "gi": [], "gv": []}
                rays.append(cur)
            elif tag == "S":
                p = ln[1:].split()
                np_ = int(p[1])
                sec = {"n": int(p[0]), "npts": np_, "tf": int(p[2]), "pts": []}
                if np_ > 0:
                    sec["first"] = (float(p[3]), float(p[4]))
                    sec["last"] = (float(p[5]), float(p[6]))
                cur["sections"].append(sec)
            elif tag == "P":
                p = ln[1:].split()
                cur["sections"][-1]["pts"].append((float(p[0]), float(p[1])))
            elif tag == "G":
                p = ln[1:].split()
                cur["gi"].append(int(p[0]))
                cur["gv"].append(float(p[1]))
    for r in rays:
        r["gi"] = np.array(r["gi"], dtype=np.int64)
        r["gv"] = np.array(r["gv"], dtype=np.float64)
    return rays
