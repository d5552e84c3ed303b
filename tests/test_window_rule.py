"""The reference's ray-window rule against the exact contact definition (CPU).

The reference decides "hits" with an OptiX any-hit ray from p along the
float32 unit direction to q over t in [0.001, |q - p|]
(lib/ptx/ray_segments.cu:77-111) and combines hits(cage) && !hits(self)
(lib/src/intersections.cpp:341-378).  The engine decides contact exactly
(closed segment against closed triangle, endpoints that are vertices of the
triangle excluded).  `oracle.window_hits` restates the window rule in
float32 with the published watertight triangle test, so the pairs on which
the two definitions disagree can be counted and sorted by cause:

* tmin: the only contacts lie within 0.001 of p, which the window skips;
* q_end: the window's only hits lie within tau of q (a ray aimed at a
  vertex q meets q's own triangles at t ~ |q - p|, and rounding decides
  whether t <= |q - p|; OptiX's rounding there is closed);
* grazing: every disagreeing triangle lies within tau = 1e-6 of the mesh's
  extent of grazing (segment within tau of its boundary, or an endpoint
  within tau of it);
* other: none, on every workload here.
"""
import numpy as np

import meshes
import oracle
from visacd_amd import workloads as wl

A = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
T1 = np.array([[0, 1, 2]], np.uint32)


def _bits(p, q, tau=1e-6):
    pts = np.array([p, q], np.float32)
    return int(oracle.window_hits(pts, A, T1, tau)[0])


def test_window_rule_closed_forms():
    X, H, N, S, G = (oracle.WINDOW_EXACT, oracle.WINDOW_HIT, oracle.WINDOW_NO_TMIN,
                     oracle.WINDOW_SHORT, oracle.WINDOW_GRAZING)
    # through the interior at t = 1 of |q - p| = 2: contact and a hit
    assert _bits((0.2, 0.2, 1), (0.2, 0.2, -1)) == X | H | N | S | G
    # crossing at t = 0.0005 < 0.001: contact, the window skips it (a clean
    # crossing, not grazing)
    assert _bits((0.2, 0.2, 0.0005), (0.2, 0.2, -1)) == X | N
    # crossing at t = 0.002: both
    assert _bits((0.2, 0.2, 0.002), (0.2, 0.2, -1)) == X | H | N | S | G
    # beside the triangle, and parallel above it: neither
    assert _bits((2, 2, 1), (2, 2, -1)) == G
    assert _bits((0.1, 0.1, 0.5), (0.5, 0.1, 0.5)) == G
    # ending at vertex a = q from above: no contact by definition; any window
    # hit lies at t ~ |q - p| (within tau of q), and the triangle is grazing
    b = _bits((0.3, 0.3, 1), (0, 0, 0))
    assert b & X == 0 and b & S == 0 and b & G
    # starting at vertex a (p is a vertex): no contact; the ray meets the
    # triangle at t = 0 only, which the window skips
    assert _bits((0, 0, 0), (0.3, 0.3, 1)) == N | G


def _torus_cage(V, grow=0.06):
    """The same torus with a thicker tube (every vertex pushed away from the
    ring circle through the tube centres by `grow`): it encloses the mesh."""
    V = V.astype(np.float64)
    rho = np.hypot(V[:, 0], V[:, 1])
    R = rho.mean()
    ring = np.stack([V[:, 0] / rho * R, V[:, 1] / rho * R, np.zeros_like(rho)], 1)
    d = V - ring
    r = np.linalg.norm(d, axis=1, keepdims=True)
    return (ring + d * (r + grow) / r).astype(np.float32)


def _edge_mismatch(sb, cb):
    """Pairs whose cage-mode edge (cage hit and no self hit) differs between
    the window rule and exact contact, by the cause on either mesh."""
    ex_edge = ((cb & oracle.WINDOW_EXACT) != 0) & ((sb & oracle.WINDOW_EXACT) == 0)
    win_edge = ((cb & oracle.WINDOW_HIT) != 0) & ((sb & oracle.WINDOW_HIT) == 0)
    diff = ex_edge != win_edge
    return ex_edge, win_edge, diff


def _report(name, sb, cb, tau):
    cs, cc = oracle.window_classes(sb), oracle.window_classes(cb)
    ex_edge, win_edge, diff = _edge_mismatch(sb, cb)
    print(f"{name}: tau {tau:.2e}", flush=True)
    print("  self", cs, flush=True)
    print("  cage", cc, flush=True)
    print(f"  cage-mode edges: exact {int(ex_edge.sum())}, window rule {int(win_edge.sum())}, "
          f"differing pairs {int(diff.sum())} of {diff.size} ({diff.mean():.4%})", flush=True)
    return cs, cc, ex_edge, diff


def test_window_rule_torus_c1_all_pairs():
    W = wl.torus_c1()
    V, T = W.vertices, W.triangles
    tau = 1e-6 * float((V.max(0) - V.min(0)).max())
    C = _torus_cage(V)
    sb = oracle.window_hits(W.points, V, T, tau)
    cb = oracle.window_hits(W.points, C, T, tau)
    cs, cc, ex_edge, diff = _report("torus C1", sb, cb, tau)
    # bit 0 is the oracle's exact contact: the cage-mode edges it gives are
    # the oracle's cage-mode edges, pair for pair
    o, _ = oracle.visibility(W.points, V, T, cage=(C, T))
    ks = np.flatnonzero(ex_edge)
    n = W.n_points
    assert ks.size == o.shape[0]
    i = o[:, 0].astype(np.int64)
    j = o[:, 1].astype(np.int64)
    np.testing.assert_array_equal(ks, i * (2 * n - i - 1) // 2 + j - i - 1)
    # every disagreement has a stated cause
    assert cs["other"] == 0 and cc["other"] == 0
    assert cs["agree"] + cs["tmin"] + cs["q_end"] + cs["grazing"] == cs["pairs"]
    # the cage lies away from the points: no window effects there
    assert cc["tmin"] == 0 and cc["q_end"] == 0
    # every differing cage-mode edge comes from the self test's window
    # effects or grazing pairs
    assert diff.sum() > 0


def test_window_rule_icosphere_cage_sampled():
    V, T = wl.perturbed_icosphere(3, 0.3, 11)
    tau = 1e-6 * float((V.max(0) - V.min(0)).max())
    C = meshes.cage_of(V, 1.06)
    sb = oracle.window_hits(V, V, T, tau)
    cb = oracle.window_hits(V, C, T, tau)
    cs, cc, ex_edge, diff = _report("icosphere L3 with its 1.06 cage", sb, cb, tau)
    assert cs["other"] == 0 and cc["other"] == 0
    o, _ = oracle.visibility(V, V, T, cage=(C, T))
    assert int(ex_edge.sum()) == o.shape[0]
