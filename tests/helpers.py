"""Shared test helpers: point sets, oracle set-up and the relabelling that maps
the GPU's computational-particle order onto the oracle's (computational
leaves share the reference order; vacuum leaves and boundary nodes are
matched by their exact FP64 centre and cell length)."""
import numpy as np

from apcloud import ic


def cloud(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        x, y, z = rng.random((3, n))
    elif kind == "clustered":
        x, y, z = ic.clustered(n, seed, n_clusters=4, scale=0.01, fraction=0.5)
    elif kind == "plummer":
        x, y, z = ic.clustered(n, seed, n_clusters=3, scale=0.003, fraction=0.6)
    elif kind == "lattice":
        m = int(round(n ** (1 / 3)))
        x, y, z = ic.lattice(m)
    elif kind == "jitter":
        m = int(round(n ** (1 / 3)))
        x, y, z = ic.perturbed_lattice(m, seed, 0.1 / m)
    elif kind == "ball":
        x, y, z = ic.uniform_ball(n, seed)
    else:
        raise ValueError(kind)
    return np.ascontiguousarray(x), np.ascontiguousarray(y), np.ascontiguousarray(z)


def relabel(gpu, orc):
    """Permutation p with orc index p[g] for every GPU index g; asserts the
    node sets are identical."""
    key_o = {}
    S = len(orc["x"])
    for i in range(S):
        k = (orc["x"][i], orc["y"][i], orc["z"][i], orc["h"][i])
        assert k not in key_o, "duplicate node in oracle output"
        key_o[k] = i
    Sg = len(gpu["x"])
    assert Sg == S, "different node counts: gpu %d oracle %d" % (Sg, S)
    p = np.empty(S, np.int64)
    for g in range(S):
        k = (gpu["x"][g], gpu["y"][g], gpu["z"][g], gpu["h"][g])
        assert k in key_o, "GPU node %d (%r) not in oracle output" % (g, k)
        p[g] = key_o[k]
    assert len(set(p.tolist())) == S
    return p


def neigh_lists(c):
    out = []
    for i in range(len(c["nneigh"])):
        f, n = c["fneigh"][i], c["nneigh"][i]
        out.append(c["neigh24"][f:f + n])
    return out
