"""Match a libimx mesh to a reference mesh up to renumbering."""
import numpy as np


def match(mine, ref):
    """Return (node_map, edge_map, tri_map, edge_flip) mine -> ref (0-based).

    Nodes are matched by bit-identical coordinates, edges by end-node pairs,
    triangles by vertex sets; edge_flip marks edges stored with the opposite
    orientation.
    """
    rp = ref["v_p"].reshape(-1, 3)
    mp = mine["v_p"].reshape(-1, 3)
    key = {row.tobytes(): i for i, row in enumerate(rp)}
    nmap = np.array([key.get(row.tobytes(), -1) for row in mp])
    rev = ref["ed_v"].reshape(-1, 2) - 1
    mev = mine["ed_v"].reshape(-1, 2) - 1
    ekey = {(min(a, b), max(a, b)): i for i, (a, b) in enumerate(rev)}
    emap = np.array([ekey.get((min(nmap[a], nmap[b]), max(nmap[a], nmap[b])), -1) for a, b in mev])
    flip = np.array([nmap[a] != rev[e, 0] for (a, b), e in zip(mev, emap)])
    rtv = ref["tr_v"].reshape(-1, 3) - 1
    mtv = mine["tr_v"].reshape(-1, 3) - 1
    tkey = {tuple(sorted(t)): i for i, t in enumerate(rtv)}
    tmap = np.array([tkey.get(tuple(sorted(nmap[t])), -1) for t in mtv])
    return nmap, emap, tmap, flip
