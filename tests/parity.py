"""Parity checks between a tracer under test and the CPU oracle.

The bar (for the same BVH and the same rays):
  * bit-exact results where the kernel evaluates the reference's arithmetic in
    the reference's order (exact triangle records, plain traversal);
  * otherwise the hit triangle must match on >= 99.999% of rays, and every
    mismatch must be a tie or a grazing hit: both candidates' FP64 ray
    parameters agree within TIE_REL relative (or the hit that only one side
    reports lies within GRAZE of a triangle edge or of t_min/t_max);
  * t, beta, gamma of matching hits within TOL_REL of the FP64 intersection,
    relative to max(|x|, 1) for the barycentrics and |t| for t -- or, where
    the fp32 CPU tracer itself misses that bound, no further from FP64 than
    twice the CPU's own error.
"""
import numpy as np

FLT_MAX = np.float32(3.4028234663852886e38)
TIE_REL = 1e-6
GRAZE = 1e-5
TOL_REL = 1e-6


def uniform_sphere_rays(n, seed, lo=0.0, hi=1.0, tmin=1e-4):
    """Rays from origins U[lo,hi]^3 in uniform random directions (8, n) float32."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    z = 1.0 - 2.0 * rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    rays = np.empty((8, n), np.float32)
    rays[0:3] = o.T
    rays[3] = tmin
    rays[4:7] = d.T
    rays[7] = FLT_MAX
    return rays


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def bit_exact(a, b):
    """Equal bit patterns of (t, u, v, id) for every ray."""
    return all(np.array_equal(bits(x), bits(y)) for x, y in zip(a, b))


def classify(orc, pos, idx, rays, test, ref):
    """Compare hits `test` against `ref` (tuples t, u, v, id). Returns a report dict."""
    tt, ut, vt, it = test
    tr, ur, vr, ir = ref
    vt_ok, vr_ok = tt != FLT_MAX, tr != FLT_MAX
    n = len(tt)
    same = (vt_ok == vr_ok) & (~vt_ok | (it == ir))
    bad = np.nonzero(~same)[0]
    rep = {"n": n, "mismatch": int(len(bad)), "agree_frac": 1.0 - len(bad) / max(n, 1),
           "unexplained": 0, "bit_exact": bit_exact(test, ref)}
    # Classify mismatches with the FP64 oracle.  When the two sides report
    # different triangles, the side with the larger t missed the nearer
    # candidate (or, for a miss, the only candidate): that candidate must be
    # grazing -- within GRAZE of an edge or of t_min/t_max -- unless both t
    # agree within TIE_REL (a tie).
    unexplained = []
    for k in bad:
        r1 = rays[:, k:k + 1]
        cands = []
        for valid, tri, tf in ((vt_ok[k], it[k], tt[k]), (vr_ok[k], ir[k], tr[k])):
            if valid:
                t64, b64, g64, _ = orc.fp64_intersect(pos, idx, np.array([tri], np.uint32), r1)
                cands.append((float(tf), t64[0], b64[0], g64[0]))
        if len(cands) == 2 and abs(cands[0][1] - cands[1][1]) <= TIE_REL * max(abs(cands[0][1]), abs(cands[1][1])):
            continue
        _, t64, b64, g64 = min(cands)
        edge = min(b64, g64, 1.0 - b64 - g64)
        tmin, tmax = float(r1[3, 0]), float(r1[7, 0])
        near_t = abs(t64 - tmin) <= GRAZE * max(abs(t64), 1e-30) or abs(t64 - tmax) <= GRAZE * abs(t64)
        if abs(edge) <= GRAZE or near_t:
            continue
        unexplained.append(int(k))
    rep["unexplained"] = len(unexplained)
    rep["unexplained_idx"] = unexplained[:10]
    return rep


def fp64_error(orc, pos, idx, rays, hits, sample=None):
    """Relative errors of t, u, v of valid hits vs the FP64 intersection."""
    t, u, v, i = hits
    valid = np.nonzero(t != FLT_MAX)[0]
    if sample is not None and len(valid) > sample:
        valid = np.random.default_rng(0).choice(valid, sample, replace=False)
    t64, b64, g64, _ = orc.fp64_intersect(pos, idx, i[valid], rays[:, valid])
    et = np.abs(t[valid] - t64) / np.abs(t64)
    eu = np.abs(u[valid] - b64) / np.maximum(np.abs(b64), 1.0)
    ev = np.abs(v[valid] - g64) / np.maximum(np.abs(g64), 1.0)
    return valid, et, eu, ev


def accuracy_ok(orc, pos, idx, rays, test, ref, sample=20000):
    """t/u/v of the tested hits within TOL_REL of FP64, or within twice the
    reference tracer's own FP64 error on rays where the reference misses it."""
    tt, ut, vt, it = test
    tr, ur, vr, ir = ref
    both = np.nonzero((tt != FLT_MAX) & (tr != FLT_MAX) & (it == ir))[0]
    if sample is not None and len(both) > sample:
        both = np.random.default_rng(1).choice(both, sample, replace=False)
    sub = lambda h: tuple(x[both] for x in h)
    r = rays[:, both]
    _, et, eu, ev = fp64_error(orc, pos, idx, r, sub(test))
    _, rt_, ru, rv = fp64_error(orc, pos, idx, r, sub(ref))
    ok = ((et <= TOL_REL) | (et <= 2 * rt_ + 1e-12)) & ((eu <= TOL_REL) | (eu <= 2 * ru + 1e-12)) & \
         ((ev <= TOL_REL) | (ev <= 2 * rv + 1e-12))
    return {"checked": int(len(both)), "fail": int((~ok).sum()),
            "max_err_t": float(et.max()) if len(et) else 0.0,
            "frac_t_within_1e-6": float((et <= TOL_REL).mean()) if len(et) else 1.0,
            "ref_frac_t_within_1e-6": float((rt_ <= TOL_REL).mean()) if len(rt_) else 1.0}
