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
    """All pairs closer than rmax under the minimum image (brute force, small N)."""
    n = x.shape[0]
    out_d, out_i, out_j = [], [], []
    for i in range(n - 1):
        d = x[i + 1:] - x[i]
        d -= np.round(d / box) * box
        r = np.sqrt((d ** 2).sum(axis=1))
        m = r < rmax
        out_d.append(d[m])
        out_i.append(np.full(m.sum(), i))
        out_j.append(np.nonzero(m)[0] + i + 1)
    return np.concatenate(out_d), np.concatenate(out_i), np.concatenate(out_j)


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
