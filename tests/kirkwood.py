"""Kirkwood's analytic solvation energy of point charges inside a dielectric
sphere in a linearised Poisson-Boltzmann solvent (Kirkwood 1934).

For charges q_i at x_i (|x_i| < a) in a sphere of radius a, interior
dielectric e_in, exterior e_out with Debye parameter kappa (the ion-exclusion
radius equals a), the reaction-field energy is

  E = 1/2 sum_ij q_i q_j / e_in * sum_n (r_i r_j)^n / a^(2n+1) F_n P_n(cos g_ij)
  F_n = (e_out g_n + (n+1) e_in) / (n e_in - e_out g_n),
  g_n = x k_n'(x) / k_n(x),  x = kappa a,

with k_n the modified spherical Bessel functions of the second kind.  At
kappa = 0, g_n = -(n+1) and F_n = (n+1)(e_in - e_out)/(n e_in + (n+1) e_out);
a single central charge gives Born with Debye screening,
E = q^2/(2a) (1/(e_out (1 + kappa a)) - 1/e_in).
Returned in kcal/mol for charges in e and lengths in Angstrom (332.0716).
"""
import numpy as np

COULOMB = 332.0716


def log_derivs(x, nmax):
    """g_n = x k_n'(x)/k_n(x), n = 0..nmax, by upward recurrence of k_n."""
    if x == 0.0:
        return -(np.arange(nmax + 1) + 1.0)
    k = np.zeros(nmax + 2)
    k[0] = np.exp(-x) / x
    k[1] = np.exp(-x) * (1.0 / x + 1.0 / x ** 2)
    for n in range(1, nmax + 1):
        k[n + 1] = k[n - 1] + (2 * n + 1) / x * k[n]
    g = np.zeros(nmax + 1)
    for n in range(nmax + 1):
        # k_n'(x) = -k_{n+1}(x) + n/x k_n(x)
        g[n] = x * (-k[n + 1] + n / x * k[n]) / k[n]
    return g


def energy(pos, q, a, eps_in, eps_out, kappa, nmax=None):
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    q = np.asarray(q, dtype=np.float64)
    r = np.linalg.norm(pos, axis=1)
    rmax = r.max() if len(r) else 0.0
    if nmax is None:
        ratio = (rmax / a) ** 2 if rmax > 0 else 0.0
        nmax = 8 if ratio == 0 else int(min(400, max(8, np.log(1e-17) / np.log(ratio) + 5)))
    g = log_derivs(kappa * a, nmax)
    n = np.arange(nmax + 1)
    F = (eps_out * g + (n + 1) * eps_in) / (n * eps_in - eps_out * g)
    u = np.where(r[:, None] > 0, pos / np.where(r > 0, r, 1.0)[:, None], 0.0)
    cosg = np.clip(u @ u.T, -1.0, 1.0)
    rr = np.outer(r, r) / a ** 2
    # Legendre recurrence over n
    P0 = np.ones_like(cosg)
    P1 = cosg.copy()
    tot = F[0] * P0 / a
    pw = rr.copy()
    if nmax >= 1:
        tot = tot + F[1] * pw * P1 / a
    Pm, Pc = P0, P1
    for k in range(1, nmax):
        Pn = ((2 * k + 1) * cosg * Pc - k * Pm) / (k + 1)
        pw = pw * rr
        tot = tot + F[k + 1] * pw * Pn / a
        Pm, Pc = Pc, Pn
    return 0.5 * COULOMB / eps_in * float(q @ tot @ q)


def born(q, a, eps_in, eps_out, kappa):
    return 0.5 * COULOMB * q * q / a * (1.0 / (eps_out * (1.0 + kappa * a)) - 1.0 / eps_in)
