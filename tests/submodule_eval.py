"""Map equation of every parent module's induced subgraph under a labelling
(the codelength the reference gives a sub-module partition's network,
generate_network_from_module + calibrate, OmpRelaxmap.cpp:640-695 and
Module.cpp:474-501), computed directly in numpy: module flow, teleport and
dangling flow of the members, exits counting only arcs inside the parent."""
import numpy as np

ALPHA, BETA = 0.15, 1.0 - 0.15


def plogp(x):
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 0, x * np.log(np.where(x > 0, x, 1.0)), 0.0)


def instance_codelengths(p, tau, ptr, idx, f, part, labels):
    n = len(p)
    src = np.repeat(np.arange(n), np.diff(ptr))
    dangling = np.diff(ptr) == 0
    d = np.where(dangling, p, 0.0)
    inside = part[src] == part[idx]
    cross = inside & (labels[src] != labels[idx])
    X = np.zeros(n)
    np.add.at(X, labels[src[cross]], f[cross])
    P, T, D = np.zeros(n), np.zeros(n), np.zeros(n)
    np.add.at(P, labels, p)
    np.add.at(T, labels, tau)
    np.add.at(D, labels, d)
    q = (ALPHA * P + BETA * D) * (1.0 - T) + BETA * X
    out = {}
    for I in np.unique(part):
        mods = np.unique(labels[part == I])
        S = q[mods].sum()
        out[int(I)] = float(plogp(S) - 2 * plogp(q[mods]).sum() + plogp(q[mods] + P[mods]).sum()
                            - plogp(p[part == I]).sum())
    return out
