"""Error of an FP32-mixed result against an FP64 one, stated per result.

rho grids: per level L-inf = max|d| / max|ref| and L2 = ||d||_2 / ||ref||_2;
hab blocks, forces and virial: max|d| / max|ref|.  Bars of the FP32-mixed
mode: rho L-inf <= 2e-6 on every level, forces <= 1e-6.
"""
import numpy as np

RHO_LINF_BAR = 2e-6
FORCE_BAR = 1e-6


def linf_rel(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    d = float(np.max(np.abs(got - ref)))
    m = float(np.max(np.abs(ref)))
    return d / m if m > 0 else d


def l2_rel(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    d = float(np.linalg.norm((got - ref).ravel()))
    m = float(np.linalg.norm(ref.ravel()))
    return d / m if m > 0 else d


def rho_errors(got_levels, ref_levels):
    """[{"linf": .., "l2": ..}] per grid level."""
    return [{"linf": linf_rel(g, r), "l2": l2_rel(g, r)}
            for g, r in zip(got_levels, ref_levels)]


def report(rho=None, hab=None, forces=None, virial=None):
    """Per-result errors of (got, ref) pairs; absent results are left out."""
    out = {}
    if rho is not None:
        lv = rho_errors(*rho)
        out["rho_linf"] = max(e["linf"] for e in lv)
        out["rho_l2"] = max(e["l2"] for e in lv)
        out["rho_levels"] = lv
    if hab is not None:
        out["hab_linf"] = linf_rel(*hab)
    if forces is not None:
        out["forces_linf"] = linf_rel(*forces)
    if virial is not None:
        out["virial_linf"] = linf_rel(*virial)
    return out
