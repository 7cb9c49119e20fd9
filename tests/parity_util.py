"""Shared helpers of the GPU parity tests: physical groups and per-field
errors against the oracle, and the oracle's own forward-error floor.

Fields are compared normwise inside physical groups.  Inside a group a
component can be many orders smaller than the group (vys ~ 1e-6 vxs and
SXX ~ 1e-6 SXY in simple shear); such a component is rounding noise of the
solve and is meaningful only relative to its group, so per-field errors are
reported for every field, and a field is held to the group's tolerance on
its own norm when it carries at least a tenth of the group's norm.

The Darcy fluxes are cancellation-limited: they are differences of nearly
equal fluid pressures (|pf| / |dpf| ~ 1e9 here), so a relative error of the
pressures of a few ulps becomes ~1e-7 in the fluxes, for any FP64 solver.
The oracle's own forward error per unknown is measured by refining its
PARDISO solution with residuals in extended precision (PARDISO's factor as
the correction solver).  The fluxes are held to ten times the oracle's own
flux error, scaled by how far the GPU's fluid pressures are from the
oracle's relative to the oracle's own pressure error (>= 1): over many steps
the two pressure histories drift apart by ulps, and the fluxes must be as
accurate as those pressures allow.  The pressures themselves are held to
1e-10 (measured ~1e-16).
"""
import json
import os

import numpy as np

GROUPS = {"velocity": (["vxs", "vys"], 1e-10), "pressure": (["pt", "pf"], 1e-10),
          "stress": (["SXX", "SYY", "SXY"], 1e-10), "slip_rate": (["VSLIPB"], 1e-10),
          "state": (["OM"], 1e-10), "viscosity": (["ETA"], 1e-10),
          "fault_stress": (["SigmaY", "SII_fault"], 1e-10),
          "darcy": (["vxD", "vyD"], None)}          # None: 10 x the oracle's floor
ALL_FIELDS = [f for fs, _ in GROUPS.values() for f in fs]
# the unknowns of the Darcy group in the solution vector (k = node * 6 + v)
DARCY_VARS = (4, 5)


def rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    n = np.linalg.norm(np.asarray(b))
    return d / n if n > 0 else d


def field_errors(x, y, fields=ALL_FIELDS):
    """Per-field errors of x against y: relative to the field's own norm and
    to its group's norm."""
    out = {}
    for name, (fs, _) in GROUPS.items():
        gy = np.sqrt(sum(np.linalg.norm(y.get(f)) ** 2 for f in fs))
        for f in fs:
            if f not in fields:
                continue
            a, b = x.get(f), y.get(f)
            d = np.linalg.norm(a - b)
            nb = np.linalg.norm(b)
            out[f] = {"rel": float(d / nb) if nb > 0 else float(d),
                      "rel_group": float(d / gy) if gy > 0 else float(d),
                      "share": float(nb / gy) if gy > 0 else 0.0}
    return out


def group_errors(x, y):
    out = {}
    for name, (fs, tol) in GROUPS.items():
        ga = np.concatenate([x.get(f).ravel() for f in fs])
        gb = np.concatenate([y.get(f).ravel() for f in fs])
        out[name] = (rel(ga, gb), tol)
    return out


SOLUTION_VARS = {"pt": 0, "vxs": 1, "vys": 2, "pf": 3, "vxD": 4, "vyD": 5}


def oracle_floor(o, refinements=3):
    """Relative forward error of the oracle's last solution per unknown
    (pt, vxs, vys, pf, vxD, vyD) and per group (velocity, pressure, darcy)
    against the solution refined with extended-precision residuals; None
    without PARDISO."""
    rp, ci, v, b = o.csr()
    x = o.solution()
    vl = v.astype(np.longdouble)
    bl = b.astype(np.longdouble)
    xl = x.astype(np.longdouble)
    for _ in range(refinements):
        r = bl - np.add.reduceat(vl * xl[ci], rp[:-1])
        dx = o.resolve(r.astype(np.float64))
        if dx is None:
            return None
        xl = xl + dx.astype(np.longdouble)
    xt = xl.astype(np.float64)
    out = {}
    for name, vars_ in (("velocity", (1, 2)), ("pressure", (0, 3)), ("darcy", DARCY_VARS)):
        idx = np.concatenate([np.arange(k, len(b), 6) for k in vars_])
        out[name] = rel(x[idx], xt[idx])
    for name, k in SOLUTION_VARS.items():
        out[name] = rel(x[k::6], xt[k::6])
    return out


def check_step(ctx, o, tag, floor=None, strict=None, loose=None, report=None):
    """Assert the group tolerances of ctx against the oracle o after a step,
    and each field's own tolerance when it carries >= 0.1 of its group.
    floor: oracle_floor(o) of this step (the Darcy bound); strict: groups to
    check (default all); loose: {group: tol} overrides.  Returns the errors."""
    loose = loose or {}
    errs = group_errors(ctx, o)
    fe = field_errors(ctx, o)
    for name, (e, tol) in errs.items():
        if strict is not None and name not in strict:
            continue
        ftol = {}
        if name == "darcy":
            if floor is None:
                continue
            drift = max(1.0, rel(ctx.get("pf"), o.get("pf")) / floor["pf"])
            tol = max(1e-10, 10 * floor["darcy"] * drift)
            ftol = {f: max(1e-10, 10 * floor[f] * drift) for f in GROUPS[name][0]}
        tol = loose.get(name, tol)
        assert e <= tol, (tag, name, e, tol)
        for f in GROUPS[name][0]:
            if fe[f]["share"] >= 0.1:
                t = loose.get(name, ftol.get(f, tol))
                assert fe[f]["rel"] <= t, (tag, f, fe[f], t)
    if report is not None:
        report.append({"tag": tag, "groups": {k: float(v[0]) for k, v in errs.items()},
                       "fields": fe, "floor": floor})
    return errs, fe


def write_report(name, rows):
    """Per-field errors to $HMEC_PARITY_REPORT/<name>.json when that
    directory is set (a GPU run's evidence); otherwise nothing."""
    d = os.environ.get("HMEC_PARITY_REPORT")
    if not d or not rows:
        return
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, name + ".json"), "w") as f:
        json.dump(rows, f, indent=1)
