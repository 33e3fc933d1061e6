"""Closed forms and brute-force partition functions (test infrastructure only)."""
import itertools

import numpy as np
from scipy import integrate

BETA_C_2D_ISING = 0.5 * np.log(1.0 + np.sqrt(2.0))
BETA_C_3D_ISING = 0.2216546          # Ferrenberg, Xu, Landau (2018)
BETA_C_3D_POTTS3 = 0.550565          # first-order transition, q = 3


def onsager_lnZ(beta):
    """ln Z / N of the square-lattice Ising model at h = 0 (Onsager 1944),
    single-integral form: ln(2 cosh 2b) + 1/(2 pi) int_0^pi ln[(1+sqrt(1-k^2 sin^2 t))/2] dt,
    k = 2 sinh 2b / cosh^2 2b."""
    k = 2.0 * np.sinh(2 * beta) / np.cosh(2 * beta) ** 2
    f = lambda t: np.log(0.5 * (1.0 + np.sqrt(max(0.0, 1.0 - k * k * np.sin(t) ** 2))))
    val, _ = integrate.quad(f, 0.0, np.pi, limit=400, epsabs=1e-15, epsrel=1e-14)
    return float(np.log(2 * np.cosh(2 * beta)) + val / (2 * np.pi))


def onsager_lnZ_double(beta):
    """Same quantity from the double integral
    ln 2 + 1/(2 pi^2) int int ln[cosh^2 2b - sinh 2b (cos t1 + cos t2)]."""
    c2, s2 = np.cosh(2 * beta) ** 2, np.sinh(2 * beta)
    f = lambda t2, t1: np.log(c2 - s2 * (np.cos(t1) + np.cos(t2)))
    val, _ = integrate.dblquad(f, 0.0, np.pi, 0.0, np.pi, epsabs=1e-13, epsrel=1e-13)
    return float(np.log(2.0) + val / (2 * np.pi ** 2))


def _lnsumexp(a):
    m = np.max(a)
    return float(m + np.log(np.sum(np.exp(a - m))))


def brute_lnZ_ising2d(beta, Lx, Ly):
    """ln Z / V of the periodic Lx x Ly Ising torus by enumeration."""
    V = Lx * Ly
    states = ((np.arange(2 ** V)[:, None] >> np.arange(V)[None, :]) & 1) * 2 - 1
    s = states.reshape(-1, Ly, Lx)
    E = np.sum(s * np.roll(s, 1, axis=1), axis=(1, 2)) + np.sum(s * np.roll(s, 1, axis=2), axis=(1, 2))
    return _lnsumexp(beta * E) / V


def _torus_bonds(shape):
    """Forward bond of every site along every axis of a periodic box.  An
    axis of length 1 gives a self-loop and one of length 2 a doubled bond,
    exactly as the periodic trace of the tensor network does."""
    sites = list(itertools.product(*[range(L) for L in shape]))
    idx = {p: n for n, p in enumerate(sites)}
    bonds = []
    for p in sites:
        for ax, L in enumerate(shape):
            qn = list(p)
            qn[ax] = (qn[ax] + 1) % L
            bonds.append((idx[p], idx[tuple(qn)]))
    return len(sites), np.array(bonds)


def brute_lnZ_ising3d(beta, shape=(2, 2, 2)):
    """ln Z / V of the periodic Ising box ``shape`` (for L = 2 every pair of
    neighbours is joined by two bonds, as in the tensor network)."""
    V, b = _torus_bonds(shape)
    states = ((np.arange(2 ** V)[:, None] >> np.arange(V)[None, :]) & 1) * 2 - 1
    E = np.sum(states[:, b[:, 0]] * states[:, b[:, 1]], axis=1)
    return _lnsumexp(beta * E) / V


def brute_lnZ_potts3d(beta, q=3, shape=(2, 2, 2)):
    """ln Z / V of the periodic q-state Potts box ``shape``, weight exp(beta delta)."""
    V, b = _torus_bonds(shape)
    n = np.arange(q ** V)
    states = np.stack([(n // q ** k) % q for k in range(V)], axis=1)
    E = np.sum(states[:, b[:, 0]] == states[:, b[:, 1]], axis=1)
    return _lnsumexp(beta * E) / V


def brute_corr_ising(beta, shape, a, b):
    """<s_a s_b> of the periodic Ising box ``shape`` (sites a, b as coordinate
    tuples), by enumeration; same bond convention as brute_lnZ_ising3d."""
    V, bonds = _torus_bonds(shape)
    sites = list(itertools.product(*[range(L) for L in shape]))
    ia, ib = sites.index(tuple(a)), sites.index(tuple(b))
    states = ((np.arange(2 ** V)[:, None] >> np.arange(V)[None, :]) & 1) * 2 - 1
    E = np.sum(states[:, bonds[:, 0]] * states[:, bonds[:, 1]], axis=1)
    w = beta * E
    p = np.exp(w - np.max(w))
    return float(np.sum(p * states[:, ia] * states[:, ib]) / np.sum(p))


def brute_ising_field(beta, h, shape=(2, 2, 2)):
    """(ln Z / V, <s>) of the periodic Ising box ``shape`` with coupling beta
    and field h (weight exp(beta sum s s' + h sum s)), by enumeration; same
    bond convention as brute_lnZ_ising3d."""
    V, bonds = _torus_bonds(shape)
    states = ((np.arange(2 ** V)[:, None] >> np.arange(V)[None, :]) & 1) * 2 - 1
    w = beta * np.sum(states[:, bonds[:, 0]] * states[:, bonds[:, 1]], axis=1) + h * np.sum(states, axis=1)
    p = np.exp(w - np.max(w))
    return _lnsumexp(w) / V, float(np.sum(p * states[:, 0]) / np.sum(p))


def brute_corr_ising_field(beta, h, shape, a, b):
    """<s_a s_b> of the periodic Ising box with field h (weight exp(beta sum s s' +
    h sum s)), by enumeration; same bond convention as brute_lnZ_ising3d."""
    V, bonds = _torus_bonds(shape)
    sites = list(itertools.product(*[range(L) for L in shape]))
    ia, ib = sites.index(tuple(a)), sites.index(tuple(b))
    states = ((np.arange(2 ** V)[:, None] >> np.arange(V)[None, :]) & 1) * 2 - 1
    w = beta * np.sum(states[:, bonds[:, 0]] * states[:, bonds[:, 1]], axis=1) + h * np.sum(states, axis=1)
    p = np.exp(w - np.max(w))
    return float(np.sum(p * states[:, ia] * states[:, ib]) / np.sum(p))
