"""Angular and centroid error figures of GPU normals against the FP64 oracle (shared by the GPU tests)."""
import json
import os

import numpy as np


def angles_deg(got, ref):
    """Sign-agnostic angle in degrees between rows of got (fp32 normals) and ref (fp64 unit normals)."""
    g = np.asarray(got, np.float64)
    r = np.asarray(ref, np.float64)
    gn = np.linalg.norm(g, axis=1)
    rn = np.linalg.norm(r, axis=1)
    c = np.abs((g * r).sum(1)) / np.maximum(gn * rn, 1e-300)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def report(name, ang, extra=None):
    """Median / p99 / max angle, printed as one JSON line (and appended to $RVP_PRECISION_LOG when set)."""
    out = {"config": name, "points": int(len(ang)), "angle_deg_median": float(np.median(ang)),
           "angle_deg_p99": float(np.percentile(ang, 99)), "angle_deg_max": float(np.max(ang))}
    if extra:
        out.update(extra)
    line = json.dumps(out)
    print("precision:", line, flush=True)
    log = os.environ.get("RVP_PRECISION_LOG")
    if log:
        with open(log, "a") as f:
            f.write(line + "\n")
    return out
