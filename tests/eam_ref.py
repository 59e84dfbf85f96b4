"""Independent closed-form EAM sums for small periodic cells (test helper).

Evaluates E = sum_i F_i(sum_j rho_{j->i}(r_ij)) + 1/2 sum_{i != j} phi(r_ij) over
explicit periodic images with the analytic functions of
mapp4py_amd.potentials, and its exact gradient. No tables, no neighbor list:
it pins the oracle's density conventions (which element's function feeds which
atom), pair bookkeeping and image handling.
"""
import numpy as np


def pair_arrays(H, x, rc, kmax=2):
    H = np.asarray(H, float)
    x = np.asarray(x, float)
    n = len(x)
    rows = []
    rng = range(-kmax, kmax + 1)
    for a in rng:
        for b in rng:
            for c in rng:
                sh = np.array([a, b, c], float) @ H
                d = x[:, None, :] - (x[None, :, :] + sh)
                r = np.sqrt((d ** 2).sum(-1))
                ii, jj = np.nonzero((r < rc) & (r > 1e-12))
                rows.append((ii, jj, d[ii, jj], r[ii, jj]))
    i = np.concatenate([r[0] for r in rows])
    j = np.concatenate([r[1] for r in rows])
    d = np.concatenate([r[2] for r in rows])
    r = np.concatenate([r[3] for r in rows])
    return i, j, d, r


def energy_forces(an, fs, H, x, elem, kmax=2):
    """Closed-form energy and forces. an: EamKnots.analytic."""
    elem = np.asarray(elem)
    n = len(x)
    i, j, d, r = pair_arrays(H, x, an["rc"], kmax)
    ei, ej = elem[i], elem[j]
    ne = int(elem.max()) + 1 if len(elem) else 1
    rho = np.zeros(n)
    drho_in = np.zeros_like(r)   # d rho_{j->i} / dr
    drho_out = np.zeros_like(r)  # d rho_{i->j} / dr
    phi = np.zeros_like(r)
    dphi = np.zeros_like(r)

    def dens(a, b):  # density at a b-atom due to an a-atom
        return an["rho"][a][b] if fs else an["rho"][a][0]

    for a in range(len(an["F"])):
        for b in range(len(an["F"])):
            m = (ei == a) & (ej == b)
            if not m.any():
                continue
            fin = dens(b, a)   # at i (element a) due to j (element b)
            fout = dens(a, b)
            rho += np.bincount(i[m], weights=fin(r[m]), minlength=n)
            drho_in[m] = fin.d(r[m])
            drho_out[m] = fout.d(r[m])
            p = an["phi"][min(a, b)][max(a, b)]
            phi[m] = p(r[m])
            dphi[m] = p.d(r[m])
    F = np.zeros(n)
    dF = np.zeros(n)
    for a in range(len(an["F"])):
        m = elem == a
        F[m] = an["F"][a](rho[m])
        dF[m] = an["dF"][a](rho[m])
    E = F.sum() + 0.5 * phi.sum()
    # dE/dr_ij for the ordered pair (i, j): phi' + F'_i rho'_{j->i} + F'_j rho'_{i->j}, halved
    # because the full sum visits every pair twice
    g = 0.5 * (dphi + dF[i] * drho_in + dF[j] * drho_out) * 2.0 / r
    f = np.zeros((n, 3))
    # force on i from the ordered pair (i, j): -g * d ; on j: +g * d (Newton III)
    np.add.at(f, i, -0.5 * g[:, None] * d)
    np.add.at(f, j, 0.5 * g[:, None] * d)
    W = 0.5 * np.einsum("k,ka,kb->ab", -g, d, d)
    return E, f, W, rho
