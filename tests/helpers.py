"""Shared test helpers: lock-step comparison of the HIP path with the oracle."""
import numpy as np

from anuga_cuda_amd import _lib
from oracle.oracle import OracleDomain

FIELDS = (("stage", _lib.SW_STAGE), ("xmomentum", _lib.SW_XMOM), ("ymomentum", _lib.SW_YMOM))


def device_array(domain, name, loc):
    h = domain._handle
    q = dict(FIELDS)[name] if name in dict(FIELDS) else {
        "elevation": _lib.SW_ELEVATION, "height": _lib.SW_HEIGHT,
        "xvelocity": _lib.SW_XVELOCITY, "yvelocity": _lib.SW_YVELOCITY}[name]
    locs = {"centroid_values": (_lib.SW_CENTROID, domain.number_of_elements),
            "edge_values": (_lib.SW_EDGE, 3 * domain.number_of_elements),
            "vertex_values": (_lib.SW_VERTEX, 3 * domain.number_of_elements),
            "boundary_values": (_lib.SW_BOUNDARY, domain.number_of_boundary_elements),
            "explicit_update": (_lib.SW_EXPLICIT_UPDATE, domain.number_of_elements),
            "semi_implicit_update": (_lib.SW_SEMI_IMPLICIT_UPDATE, domain.number_of_elements)}
    code, n = locs[loc]
    out = np.empty(n)
    h.download(q, code, out)
    return out


def compare(got, exp, what, tol=1e-12, stats=None):
    """Normwise relative check; also counts bitwise mismatches."""
    got = np.asarray(got, dtype=np.float64).ravel()
    exp = np.asarray(exp, dtype=np.float64).ravel()
    assert got.shape == exp.shape, what
    nbit = int(np.count_nonzero(got.view(np.int64) != exp.view(np.int64)))
    scale = max(float(np.max(np.abs(exp))) if exp.size else 0.0, 1e-300)
    err = float(np.max(np.abs(got - exp))) / scale if exp.size else 0.0
    if stats is not None:
        stats[what] = (nbit, err)
    assert np.all(np.isfinite(got)), what + ": non-finite values"
    assert err <= tol, "%s: normwise relative error %.3e > %.1e (%d bit mismatches)" % (what, err, tol, nbit)
    return nbit, err


def compare_state(domain, od, tol=1e-12, locs=("centroid_values", "vertex_values", "edge_values"),
                  stats=None, tag=""):
    for name, _ in FIELDS:
        for loc in locs:
            compare(device_array(domain, name, loc), od.q(name, loc), "%s%s.%s" % (tag, name, loc),
                    tol, stats)


def wet_mask(stage, bed, eps):
    return (stage - bed) > eps
