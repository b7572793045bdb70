"""Observables for long-run checks: kinetic temperature, virial pressure and the
radial distribution function. The reference has no pressure or RDF (it reports
Σv²/(3N), test_ordered_mesh.hpp:85-93); the definitions here are the textbook
ones for DPD with m = kT = 1."""
import numpy as np


def kinetic_temperature(v):
    """T = <v^2> / 3 (per-component equipartition, m = 1), zero net momentum."""
    v = np.asarray(v, np.float64)
    return float((v ** 2).sum() / (3 * v.shape[0]))


def _pairs(x, box, rmax):
    """All pairs closer than rmax under the minimum image: (d = x_j - x_i nearest
    image, i, j) with i < j. A periodic k-d tree finds the candidates; the
    distances are recomputed here in FP64 from the positions."""
    from scipy.spatial import cKDTree
    x = np.asarray(x, np.float64)
    box = np.asarray(box, np.float64)
    xt = np.mod(x, box)
    ij = cKDTree(xt, boxsize=box).query_pairs(r=rmax, output_type="ndarray")
    if ij.size == 0:
        return np.zeros((0, 3)), np.zeros(0, np.int64), np.zeros(0, np.int64)
    i, j = ij[:, 0], ij[:, 1]
    d = x[j] - x[i]
    d -= np.round(d / box) * box
    r = np.sqrt((d ** 2).sum(axis=1))
    m = r < rmax
    return d[m], i[m], j[m]


def virial_sums(x, v, box, a, types):
    """The engine's observables restated in FP64 on a downloaded state: momentum
    sum v, sum v.v, the conservative virial sum over pairs (r < 1) of
    a_ij (1 - r) r (a_ij as the engine holds it, FP32) and the pair count."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    d, i, j = _pairs(x, box, 1.0)
    r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    a32 = np.asarray(a, np.float32).astype(np.float64)
    vir = (a32[types[i], types[j]] * (1.0 - r) * r).sum()
    return v.sum(axis=0), (v * v).sum(), vir, len(r)


def conservative_pressure(x, v, box, a, types):
    """P = rho kT + (1 / 3V) sum_pairs r_ij . F^C_ij with F^C = a_ij (1 - r) r_hat
    (the Groot-Warren equation of state is p = rho kT + 0.101 a rho^2 at rho = 3)."""
    x = np.asarray(x, np.float64)
    V = float(np.prod(box))
    d, i, j = _pairs(x, np.asarray(box, np.float64), 1.0)
    r = np.sqrt((d ** 2).sum(axis=1))
    aij = a[types[i], types[j]]
    vir = (aij * (1.0 - r) * r).sum()
    n = x.shape[0]
    return n / V * kinetic_temperature(v) + vir / (3 * V)


def rdf(x, box, rmax=2.0, nbins=40):
    x = np.asarray(x, np.float64)
    box = np.asarray(box, np.float64)
    d, _, _ = _pairs(x, box, rmax)
    r = np.sqrt((d ** 2).sum(axis=1))
    hist, edges = np.histogram(r, bins=nbins, range=(0.0, rmax))
    n = x.shape[0]
    rho = n / np.prod(box)
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    g = hist / (0.5 * n * rho * shell)
    return 0.5 * (edges[1:] + edges[:-1]), g
