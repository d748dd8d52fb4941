"""Local-optimality certificate of a clustering, recomputed from scratch in
FP64 with numpy (test infrastructure, independent of the C++ oracles and of
the device code).

The reference's last phase runs until a loop makes no modification
(Common/vtkUniformClustering.h:773-805): for every cut edge (I1, I2) it keeps
the labels when Try1 <= Try2 and Try1 <= Try3 (:960-991), where Try2 moves
I1 to the cluster of I2 unless I1's cluster has one item or the ring check
(DiscreteRemeshing/vtkVerticesProcessing.h:168-237, while the connexity
constraint is on) forbids it, and Try3 moves I2 likewise (:922-956).  So at
the end of a run no allowed single move across a cut edge lowers the
energy.  `violations` lists the allowed moves whose energy change, computed
here from the cluster sums, is below -tol.

Energies (shifted, as the reference's; the constant sum w|x|^2 cancels in
every difference):
  isotropic  E_c = -|S|^2 / W                     (vtkIsotropicMetricForClustering.h:114-121)
  QEM        E_c = W |q|^2 - 2 q.S, q the rank-truncated quadric point of the
             cluster quadric from the centroid S/W (vtkQEMetricForClustering.h:194-208,
             Common/vtkQuadricTools.cxx:83-177: singular values kept by
             decreasing magnitude while |s| / s_max > 1e-3, at most `level`).
"""
import numpy as np


def rep_points(Q9, c, level=3, thr=1e-3):
    """Batched rank-truncated quadric points q = c + A+_trunc (b - A c)."""
    Q9 = np.asarray(Q9, np.float64).reshape(-1, 9)
    c = np.asarray(c, np.float64).reshape(-1, 3)
    A = np.empty((len(Q9), 3, 3))
    A[:, 0, 0], A[:, 0, 1], A[:, 0, 2] = Q9[:, 0], Q9[:, 1], Q9[:, 2]
    A[:, 1, 0], A[:, 1, 1], A[:, 1, 2] = Q9[:, 1], Q9[:, 4], Q9[:, 5]
    A[:, 2, 0], A[:, 2, 1], A[:, 2, 2] = Q9[:, 2], Q9[:, 5], Q9[:, 7]
    b = -Q9[:, [3, 6, 8]]
    r = b - np.einsum("nij,nj->ni", A, c)
    lam, V = np.linalg.eigh(A)
    mag = np.abs(lam)
    order = np.argsort(-mag, axis=1, kind="stable")
    smax = mag.max(axis=1)
    q = c.copy()
    alive = np.ones(len(Q9), bool)
    for k in range(min(level, 3)):
        idx = order[:, k]
        lk = np.take_along_axis(lam, idx[:, None], 1)[:, 0]
        mk = np.abs(lk)
        keep = alive & (smax > 0) & (mk / np.where(smax > 0, smax, 1.0) > thr)
        alive &= keep
        vk = np.take_along_axis(V, idx[:, None, None].repeat(3, 1), 2)[:, :, 0]
        f = np.where(keep, np.einsum("ni,ni->n", vk, r) / np.where(keep, lk, 1.0), 0.0)
        q += f[:, None] * vk
    return q


def cluster_energy(metric, W, S, Q9=None, level=3, constrained=True, apos=None):
    """apos: per row the anchor position of a fixed QEM cluster, NaN for a
    free one (its point is the anchor in every phase,
    vtkQEMetricForClustering.h:270-274)."""
    W = np.asarray(W, np.float64)
    S = np.asarray(S, np.float64).reshape(-1, 3)
    if metric == 0 or not constrained:
        E = -np.einsum("ni,ni->n", S, S) / W
    else:
        q = rep_points(Q9, S / W[:, None], level)
        E = W * np.einsum("ni,ni->n", q, q) - 2.0 * np.einsum("ni,ni->n", q, S)
    if metric == 1 and apos is not None:
        fixed = ~np.isnan(apos[:, 0])
        a = apos[fixed]
        E[fixed] = W[fixed] * np.einsum("ni,ni->n", a, a) - 2.0 * np.einsum("ni,ni->n", a, S[fixed])
    return E


def neighbour_sets(nv, a, b):
    nb = [set() for _ in range(nv)]
    for u, v in zip(a.tolist(), b.tolist()):
        nb[u].add(v)
        nb[v].add(u)
    return nb


def ring_problem(nb, lab, i, c):
    """Would removing item i from cluster c disconnect c's items among i's
    neighbours (connected through edges between those neighbours)?"""
    ring = [v for v in nb[i] if lab[v] == c]
    if not ring:
        return False
    ring_set = set(ring)
    seen = {ring[0]}
    stack = [ring[0]]
    while stack:
        u = stack.pop()
        for v in nb[u]:
            if v in ring_set and v not in seen:
                seen.add(v)
                stack.append(v)
    return len(seen) != len(ring_set)


def violations(metric, xyz, a, b, lab, K, w, q9=None, level=3, connexity=True, constrained=True, rel_tol=1e-9,
               accept_eps=0.0, fixed=None):
    """Allowed single moves across cut edges that lower the energy by more
    than the tolerance: a list of (item, from, to, dE, tol).

    rel_tol scales |E_from| + |E_to| (FP64 rounding of the shifted energies);
    accept_eps adds accept_eps * M, M the sum of the magnitudes of the terms
    of dE: w (t_a + t_b) with t_a = W_a / (W_a - w) |x - c_a|^2 and
    t_b = W_b / (W_b + w) |x - c_b|^2 (the isotropic part, Hartigan form), plus
    T = W |q - c|^2 of both clusters before and after (QEM) -- the device
    takes a move only when dE < -2^-18 M.  fixed: item of fixed cluster k
    (SetFixedClusters); that item never moves and, with the QEM metric, is
    the cluster's point."""
    x = np.asarray(xyz, np.float64).reshape(-1, 3)
    lab = np.asarray(lab)
    w = np.asarray(w, np.float64)
    nv = len(x)
    W = np.bincount(lab, weights=w, minlength=K + 1)[:K]
    S = np.stack([np.bincount(lab, weights=w * x[:, k], minlength=K + 1)[:K] for k in range(3)], 1)
    size = np.bincount(lab, minlength=K + 1)[:K]
    Q = None
    if metric == 1:
        q9 = np.asarray(q9, np.float64).reshape(-1, 9)
        Q = np.zeros((K + 1, 9))
        np.add.at(Q, lab, q9)
        Q = Q[:K]
    apos = None
    anchor_item = np.full(K, -1, np.int64)
    if fixed is not None and len(fixed):
        fixed = np.asarray(fixed, np.int64)
        anchor_item[:len(fixed)] = fixed
        apos = np.full((K, 3), np.nan)
        apos[:len(fixed)] = x[fixed]
    E = cluster_energy(metric, W, S, Q, level, constrained, apos)
    cut = (lab[a] != lab[b]) & (lab[a] < K) & (lab[b] < K)  # null items (label K) are not moved
    src = np.concatenate([a[cut], b[cut]])
    dst_item = np.concatenate([b[cut], a[cut]])
    ca = lab[src]
    cb = lab[dst_item]
    ok = (size[ca] > 1) & (anchor_item[ca] != src)
    src, ca, cb = src[ok], ca[ok], cb[ok]
    wi = w[src][:, None]
    xi = x[src]
    Wa, Sa = W[ca] - w[src], S[ca] - wi * xi
    Wb, Sb = W[cb] + w[src], S[cb] + wi * xi
    Qa = Qb = None
    if metric == 1:
        Qa, Qb = Q[ca] - q9[src], Q[cb] + q9[src]
    pa = pb = None
    if apos is not None:
        pa, pb = apos[ca], apos[cb]
    dE = (cluster_energy(metric, Wa, Sa, Qa, level, constrained, pa)
          + cluster_energy(metric, Wb, Sb, Qb, level, constrained, pb) - E[ca] - E[cb])
    tol = rel_tol * (np.abs(E[ca]) + np.abs(E[cb]))
    if accept_eps:
        ca_c, cb_c = S[ca] / W[ca][:, None], S[cb] / W[cb][:, None]
        ta = W[ca] / Wa * np.einsum("ni,ni->n", xi - ca_c, xi - ca_c)
        tb = W[cb] / Wb * np.einsum("ni,ni->n", xi - cb_c, xi - cb_c)
        M = w[src] * (ta + tb)
        if metric == 1 and (constrained or apos is not None):
            def T(Wx, Sx, Qx, px):
                return cluster_energy(1, Wx, Sx, Qx, level, constrained, px) - cluster_energy(0, Wx, Sx)
            M = (M + T(W[ca], S[ca], Q[ca], pa) + T(Wa, Sa, Qa, pa) + T(W[cb], S[cb], Q[cb], pb)
                 + T(Wb, Sb, Qb, pb))
        tol = tol + accept_eps * M
    out = []
    nb = None
    for k in np.flatnonzero(dE < -tol):
        i, f, t = int(src[k]), int(ca[k]), int(cb[k])
        slack = float(tol[k])
        if connexity:
            if nb is None:
                nb = neighbour_sets(nv, a, b)
            if ring_problem(nb, lab, i, f):
                continue
        out.append((i, f, t, float(dE[k]), slack))
    return out
