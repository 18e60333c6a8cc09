"""CPU oracle of track building (test infrastructure only).

Restates recon.py:14-56 (build_tracks) with numpy and scipy 1.15.3's
scipy.sparse.csgraph.connected_components, the library routine the reference calls:

* node of feature f of photo i = offsets[i] + f, offsets = exclusive cumsum of nfeat
  (recon.py:17-23);
* components of the undirected graph of all match edges over all nodes (recon.py:26-30);
  scipy labels them in order of first visit, scanning nodes upwards, i.e. in order of each
  component's smallest node (tests/test_oracle_tracks.py checks this on random graphs);
* matched nodes only (recon.py:31-37), their photo by a right-sided search of offsets
  (recon.py:38-39);
* a component holding two nodes of one photo is dropped, and so is one with fewer than two
  matched nodes (recon.py:40-48);
* surviving components numbered by label, observations stably sorted by that number, so nodes
  ascend inside each track (recon.py:49-56).

Pinned in tests/test_oracle_tracks.py by a pure-Python breadth-first reference on small
random graphs (components, photo clash, order), hand-built cases, and invariants at flight
size.  ``brute_force`` below is that breadth-first reference.
"""
from __future__ import annotations

from collections import deque

import numpy as np


def _edges(pairs, nfeat):
    nfeat = np.asarray(nfeat, np.int64)
    offsets = np.concatenate([[0], np.cumsum(nfeat)]).astype(np.int64)
    if not pairs:
        return offsets, np.zeros(0, np.int64), np.zeros(0, np.int64)
    u = np.concatenate([offsets[i] + np.asarray(q, np.int64) for i, _, q, _ in pairs])
    v = np.concatenate([offsets[j] + np.asarray(t, np.int64) for _, j, _, t in pairs])
    return offsets, u, v


def build_tracks(pairs, nfeat, n_img):
    """recon.py:14-56: (tid, img, feat) int64 and the number of tracks."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    z = np.zeros(0, np.int64)
    if not pairs:
        return z, z.copy(), z.copy(), 0
    offsets, u, v = _edges(pairs, nfeat)
    n = int(offsets[-1])
    g = sp.csr_matrix((np.ones(len(u), np.int8), (u, v)), shape=(n, n))
    n_comp, label = connected_components(g, directed=False)
    matched = np.zeros(n, bool)
    matched[u] = True
    matched[v] = True
    node = np.nonzero(matched)[0]
    comp = label[node].astype(np.int64)
    img = np.searchsorted(offsets, node, side="right") - 1
    # a component is dropped if some photo appears twice in it
    key = np.unique(np.stack([comp, img], 1), axis=0, return_counts=True)
    clash = np.zeros(n_comp, bool)
    clash[key[0][key[1] > 1, 0]] = True
    members = np.bincount(comp, minlength=n_comp)
    keep = ~clash[comp] & (members[comp] >= 2)
    comp, img, node = comp[keep], img[keep], node[keep]
    alive = np.zeros(n_comp, bool)
    alive[comp] = True
    number = np.cumsum(alive) - 1
    tid = number[comp]
    order = np.argsort(tid, kind="stable")
    return (tid[order].astype(np.int64), img[order].astype(np.int64),
            (node - offsets[img])[order].astype(np.int64), int(alive.sum()))


def brute_force(pairs, nfeat, n_img):
    """Breadth-first search over an adjacency list, pure Python (small inputs only): the
    same definition of a track, built without scipy or numpy sorting."""
    nfeat = [int(x) for x in nfeat]
    first = [0]
    for c in nfeat:
        first.append(first[-1] + c)
    adj = {}
    for i, j, q, t in pairs:
        for a, b in zip(q, t):
            x, y = first[i] + int(a), first[j] + int(b)
            adj.setdefault(x, set()).add(y)
            adj.setdefault(y, set()).add(x)
    seen = set()
    tracks = []
    for s in sorted(adj):               # components in order of their smallest node
        if s in seen:
            continue
        comp, dq = [], deque([s])
        seen.add(s)
        while dq:
            x = dq.popleft()
            comp.append(x)
            for y in adj[x]:
                if y not in seen:
                    seen.add(y)
                    dq.append(y)
        comp.sort()
        photos = []
        for x in comp:
            k = 0
            while first[k + 1] <= x:
                k += 1
            photos.append(k)
        if len(comp) >= 2 and len(set(photos)) == len(photos):
            tracks.append([(k, x - first[k]) for k, x in zip(photos, comp)])
    tid = [n for n, tr in enumerate(tracks) for _ in tr]
    img = [k for tr in tracks for k, _ in tr]
    feat = [f for tr in tracks for _, f in tr]
    return (np.array(tid, np.int64), np.array(img, np.int64), np.array(feat, np.int64), len(tracks))


def synthetic_flight_matches(n_img=400, nfeat=8000, grid=20, seed=0, p_obs=0.45, keep=0.9, outlier=0.02):
    """Seeded match lists shaped like a 400-photo survey.  Photos sit on a grid x grid raster;
    each ground point is seen by its centre photo and by each of the 8 raster neighbours with
    probability p_obs (points seen once are dropped), taking the next free feature index of
    each photo (at most nfeat).  Every two photos within one raster step of each other are a
    pair, matched on the points both see (each match kept with probability `keep`), plus
    `outlier` x as many random wrong matches that merge or spoil tracks.  Returns (pairs,
    features per photo) with pairs in (i, j), i < j, order."""
    rng = np.random.default_rng(seed)
    n_pts = n_img * nfeat // 3
    dr = np.repeat(np.arange(-1, 2), 3)
    dc = np.tile(np.arange(-1, 2), 3)
    centre = rng.integers(0, n_img, n_pts)
    cr, cc = np.divmod(centre, grid)
    pr = cr[:, None] + dr[None, :]
    pc = cc[:, None] + dc[None, :]
    seen = (rng.random((n_pts, 9)) < p_obs) | (np.arange(9) == 4)[None, :]
    seen &= (pr >= 0) & (pr < n_img // grid) & (pc >= 0) & (pc < grid)
    seen &= seen.sum(1, keepdims=True) >= 2
    photo = np.where(seen, pr * grid + pc, -1)
    # feature index: order of the point among the points each photo sees, capped at nfeat
    pt, slot = np.nonzero(seen)
    ph = photo[pt, slot]
    o = np.argsort(ph, kind="stable")
    first = np.searchsorted(ph[o], np.arange(n_img))
    rank = np.empty(len(o), np.int64)
    rank[o] = np.arange(len(o)) - first[ph[o]]
    feat = np.full((n_pts, 9), -1, np.int64)
    feat[pt, slot] = np.where(rank < nfeat, rank, -1)
    edges = []
    for a in range(9):
        for b in range(9):
            if abs(dr[a] - dr[b]) > 1 or abs(dc[a] - dc[b]) > 1 or a == b:
                continue
            m = (feat[:, a] >= 0) & (feat[:, b] >= 0) & (photo[:, a] < photo[:, b])
            edges.append(np.stack([photo[m, a], photo[m, b], feat[m, a], feat[m, b]], 1))
    e = np.concatenate(edges)
    e = e[rng.random(len(e)) < keep]
    e = e[np.lexsort((e[:, 2], e[:, 1], e[:, 0]))]
    pair_key = e[:, 0] * n_img + e[:, 1]
    keys, start = np.unique(pair_key, return_index=True)
    bounds = np.append(start, len(e))
    pairs = []
    for k in range(len(keys)):
        i, j = divmod(int(keys[k]), n_img)
        q, t = e[bounds[k]:bounds[k + 1], 2], e[bounds[k]:bounds[k + 1], 3]
        n_out = int(outlier * len(q))
        if n_out:
            q = np.concatenate([q, rng.integers(0, nfeat, n_out)])
            t = np.concatenate([t, rng.integers(0, nfeat, n_out)])
        pairs.append((i, j, q, t))
    return pairs, np.full(n_img, nfeat, np.int64)
