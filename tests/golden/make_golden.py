"""Write tests/golden/tiny_cases.json: 50-digit values of the regularized
semi-discrete dual (functional F, gradient g, C_global, per-cell candidate
sets) on tiny seeded inputs.

The definition evaluated here is the one SotSolver::evaluate_functional
assembles (SotSolver.cc:217-377, local_assemble SotSolver.cc:769-869):
    e_qj = (psi_j - |x_q - y_j|^2 / 2) / eps,  j in C(cell(q))
    C(K) = { j : |y_j - center(K)| <= R }
    S_q  = sum_j nu_j exp(e_qj)                 (any shift cancels)
    F    = sum_q rho_q JxW_q eps log S_q - sum_j psi_j nu_j
    g_j  = sum_q rho_q JxW_q nu_j exp(e_qj) / S_q - nu_j
    C    = sum_q rho_q JxW_q / S_q^shifted,   S^shifted = S_q exp(-max_j e_qj)
with q-points whose double-precision S_q underflows to 0 skipped (plain mode,
SotSolver.cc:854).  Inputs are stored as the exact doubles used, so the
values are the mathematical ones for those inputs; quadrature points come
from textbook Gauss-Legendre nodes, not from the oracle.

refine_cases.json holds the same kind of values for the softmax potential
transfer between target levels (refine() below), bary_cases.json for the
weighted barycenters of the Lloyd step (barycenters() below), points_cases.json
for the conditional densities and the barycentric transport map
(point_queries() below).

Run: python tests/golden/make_golden.py   (needs mpmath and numpy only)
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))

# double underflow: exp(e) rounds to 0 below log(2^-1075)
UNDERFLOW = mp.log(mp.mpf(2) ** -1075)


def gauss2():
    a = 0.5 - 0.5 / np.sqrt(3.0)
    return [a, 1.0 - a], [0.5, 0.5]


def box_source(dim, n, seed):
    """n^dim cells on [0,1]^dim, 2-point Gauss per direction, bilinear /
    trilinear density from seeded nodal values."""
    rng = np.random.default_rng(seed)
    gp, gw = gauss2()
    h = 1.0 / n
    nodal = 0.5 + rng.random(tuple([n + 1] * dim))
    xs, rho, jxw, cen = [], [], [], []
    for cell in np.ndindex(*([n] * dim)):
        c = [(cell[d] + 0.5) * h for d in range(dim)]
        cen.append(c)
        for q in np.ndindex(*([2] * dim)):
            xi = [gp[q[d]] for d in range(dim)]
            x = [(cell[d] + xi[d]) * h for d in range(dim)]
            r = 0.0
            for v in np.ndindex(*([2] * dim)):
                phi = 1.0
                for d in range(dim):
                    phi *= xi[d] if v[d] else 1.0 - xi[d]
                r += nodal[tuple(cell[d] + v[d] for d in range(dim))] * phi
            w = 1.0
            for d in range(dim):
                w *= gw[q[d]]
            xs.append(x)
            rho.append(r)
            jxw.append(w * h ** dim)
    return (np.array(xs), np.array(rho), np.array(jxw), np.array(cen), 2 ** dim)


def evaluate(dim, nq, x, rho, jxw, cen, y, nu, psi, eps, R, lse):
    n_cells = len(cen)
    N = len(y)
    eps = mp.mpf(eps)
    F = mp.mpf(0)
    Cg = mp.mpf(0)
    g = [mp.mpf(0)] * N
    counts, members = [], []
    for c in range(n_cells):
        cand = []
        for j in range(N):
            d2 = sum((mp.mpf(y[j][k]) - mp.mpf(cen[c][k])) ** 2 for k in range(dim))
            if mp.sqrt(d2) <= mp.mpf(R):
                cand.append(j)
        counts.append(len(cand))
        members.append(cand)
        for q in range(nq):
            i = c * nq + q
            es = []
            for j in cand:
                d2 = sum((mp.mpf(x[i][k]) - mp.mpf(y[j][k])) ** 2 for k in range(dim))
                es.append((mp.mpf(psi[j]) - d2 / 2) / eps)
            if not es:
                continue
            terms = []
            for e in es:
                if not lse and e < UNDERFLOW + 5:
                    terms.append(mp.mpf(0))
                else:
                    terms.append(mp.exp(e))
            S = sum(mp.mpf(nu[j]) * t for j, t in zip(cand, terms))
            if S == 0:
                continue
            w = mp.mpf(rho[i]) * mp.mpf(jxw[i])
            F += w * eps * mp.log(S)
            M = max(es) if lse else mp.mpf(0)
            Cg += w / (S * mp.exp(-M))
            for j, t in zip(cand, terms):
                g[j] += w * mp.mpf(nu[j]) * t / S
    for j in range(N):
        F -= mp.mpf(psi[j]) * mp.mpf(nu[j])
        g[j] -= mp.mpf(nu[j])
    return F, g, Cg, counts, members


def case(name, dim, n, N, seed, eps, R, lse, cluster=None):
    x, rho, jxw, cen, nq = box_source(dim, n, seed)
    rng = np.random.default_rng(seed + 100)
    if cluster is None:
        y = rng.random((N, dim))
    else:
        y = cluster + 0.05 * rng.random((N, dim))
    nu = rng.random(N) + 0.2
    nu = nu / nu.sum()
    psi = 0.05 * rng.standard_normal(N)
    F, g, Cg, counts, members = evaluate(dim, nq, x, rho, jxw, cen, y, nu, psi, eps, R, lse)
    hashes = []
    for m in members:
        h = 0
        for j in m:
            v = ((j + 1) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
            v ^= v >> 29
            h = (h + v) & 0xFFFFFFFFFFFFFFFF
        hashes.append(str(h))
    return {
        "name": name, "dim": dim, "nq": nq, "eps": eps, "R": R, "use_lse": lse,
        "x": x.ravel().tolist(), "rho": rho.tolist(), "jxw": jxw.tolist(),
        "center": cen.ravel().tolist(), "y": y.ravel().tolist(), "nu": nu.tolist(),
        "psi": psi.tolist(),
        "F": mp.nstr(F, 40), "g": [mp.nstr(v, 40) for v in g], "C": mp.nstr(Cg, 40),
        "count": counts, "hash": hashes,
    }


def refine(dim, nq, x, rho, jxw, cen, yc, nuc, psic, yf, children, lam, R):
    """SoftmaxRefinement::compute_refinement (SoftmaxRefinement.cc:61-280):
        pot_f = sum_q w_q exp(-|x_q - y_f|^2 / (2 lam)) / Z_q,
        Z_q   = sum_c nu_c exp((psi_c - |x_q - y_c|^2 / 2) / lam),
    c over the coarse points within R of the center of the cell of q that are
    within R of x_q, f over the children of those coarse points (a child
    listed k times counts k times) within R of x_q; psi_f = -lam log pot_f,
    0 where nothing arrives."""
    lam = mp.mpf(lam)
    Rm = mp.mpf(R)
    pot = [mp.mpf(0)] * len(yf)
    for c in range(len(cen)):
        cand = [j for j in range(len(yc))
                if mp.sqrt(sum((mp.mpf(yc[j][k]) - mp.mpf(cen[c][k])) ** 2 for k in range(dim))) <= Rm]
        fine = [f for j in cand if j < len(children) for f in children[j] if 0 <= f < len(yf)]
        for q in range(nq):
            i = c * nq + q
            Z = mp.mpf(0)
            for j in cand:
                d2 = sum((mp.mpf(x[i][k]) - mp.mpf(yc[j][k])) ** 2 for k in range(dim))
                if mp.sqrt(d2) <= Rm:
                    Z += mp.mpf(nuc[j]) * mp.exp((mp.mpf(psic[j]) - d2 / 2) / lam)
            if Z == 0:
                continue
            w = mp.mpf(rho[i]) * mp.mpf(jxw[i])
            for f in fine:
                d2 = sum((mp.mpf(x[i][k]) - mp.mpf(yf[f][k])) ** 2 for k in range(dim))
                if mp.sqrt(d2) <= Rm:
                    pot[f] += w * mp.exp(-d2 / (2 * lam)) / Z
    return [(-lam * mp.log(v)) if v > 0 else mp.mpf(0) for v in pot]


def refine_case(name, dim, n, Nc, Nf, seed, lam, R, lse):
    x, rho, jxw, cen, nq = box_source(dim, n, seed)
    rng = np.random.default_rng(seed + 200)
    yc = rng.random((Nc, dim))
    nuc = rng.random(Nc) + 0.2
    nuc = nuc / nuc.sum()
    psic = 0.05 * rng.standard_normal(Nc)
    # fine points scattered around their parents, every coarse point a parent
    parent = np.concatenate([np.arange(Nc), rng.integers(0, Nc, Nf - Nc)])
    yf = yc[parent] + 0.08 * (rng.random((Nf, dim)) - 0.5)
    children = [[int(f) for f in np.flatnonzero(parent == c)] for c in range(Nc)]
    # one fine point listed under a second parent, one invalid index
    children[0].append(int(np.flatnonzero(parent == 1)[0]))
    children[-1].append(Nf + 3)
    psif = refine(dim, nq, x, rho, jxw, cen, yc, nuc, psic, yf, children, lam, R)
    return {
        "name": name, "dim": dim, "nq": nq, "lambda": lam, "R": R, "use_lse": lse,
        "x": x.ravel().tolist(), "rho": rho.tolist(), "jxw": jxw.tolist(),
        "center": cen.ravel().tolist(), "y_coarse": yc.ravel().tolist(), "nu_coarse": nuc.tolist(),
        "psi_coarse": psic.tolist(), "y_fine": yf.ravel().tolist(), "children": children,
        "psi_fine": [mp.nstr(v, 40) for v in psif],
    }


def barycenters(dim, nq, x, rho, jxw, cen, y, nu, psi, eps, R):
    """compute_weighted_barycenters_euclidean (SotSolver.cc:871-971):
        b_j = sum_q (rho_q JxW_q / S_q) nu_j exp(e_qj) x_q / nu_j,
    j over the candidates of the cell of q (cell rule by the center)."""
    eps = mp.mpf(eps)
    b = [[mp.mpf(0)] * dim for _ in range(len(y))]
    for c in range(len(cen)):
        cand = [j for j in range(len(y))
                if mp.sqrt(sum((mp.mpf(y[j][k]) - mp.mpf(cen[c][k])) ** 2 for k in range(dim))) <= mp.mpf(R)]
        for q in range(nq):
            i = c * nq + q
            es = {j: (mp.mpf(psi[j]) - sum((mp.mpf(x[i][k]) - mp.mpf(y[j][k])) ** 2
                                            for k in range(dim)) / 2) / eps for j in cand}
            S = sum(mp.mpf(nu[j]) * mp.exp(e) for j, e in es.items())
            if S == 0:
                continue
            w = mp.mpf(rho[i]) * mp.mpf(jxw[i])
            for j, e in es.items():
                for k in range(dim):
                    b[j][k] += w * mp.exp(e) / S * mp.mpf(x[i][k])
    return b


def bary_case(name, dim, n, N, seed, eps, R):
    x, rho, jxw, cen, nq = box_source(dim, n, seed)
    rng = np.random.default_rng(seed + 300)
    y = rng.random((N, dim))
    nu = rng.random(N) + 0.2
    nu = nu / nu.sum()
    psi = 0.05 * rng.standard_normal(N)
    b = barycenters(dim, nq, x, rho, jxw, cen, y, nu, psi, eps, R)
    return {"name": name, "dim": dim, "nq": nq, "eps": eps, "R": R,
            "x": x.ravel().tolist(), "rho": rho.tolist(), "jxw": jxw.tolist(),
            "center": cen.ravel().tolist(), "y": y.ravel().tolist(), "nu": nu.tolist(),
            "psi": psi.tolist(), "bary": [[mp.nstr(v, 40) for v in r] for r in b]}


def main_bary():
    cases = [
        bary_case("2d_untruncated", 2, 3, 6, 21, 0.05, 1.7976931348623157e308),
        bary_case("3d_truncated", 3, 2, 7, 22, 0.1, 0.6),
        bary_case("2d_truncated_small_eps", 2, 4, 9, 23, 0.01, 0.4),
    ]
    out = os.path.join(HERE, "bary_cases.json")
    with open(out, "w") as f:
        json.dump(cases, f)
    print("wrote", out)


def point_queries(dim, y, nu, psi, eps, R, pts, rho, idx, radius):
    """get_potential_conditioned_density (SotSolver.cc:973-1046) at the points
    with threshold R, and the barycentric map (OptimalTransportPlan.cc:218-310)
    with truncation radius `radius` (strict <, nearest target when no target
    is inside)."""
    eps = mp.mpf(eps)
    cond = [[mp.mpf(0)] * len(pts) for _ in idx]
    bary = []
    for p, x in enumerate(pts):
        d2 = [sum((mp.mpf(x[k]) - mp.mpf(y[j][k])) ** 2 for k in range(dim)) for j in range(len(y))]
        cand = [j for j in range(len(y)) if mp.sqrt(d2[j]) <= mp.mpf(R)]
        Z = sum(mp.mpf(nu[j]) * mp.exp((mp.mpf(psi[j]) - d2[j] / 2) / eps) for j in cand)
        if Z > 0:
            for a, i in enumerate(idx):
                if i in cand:
                    cond[a][p] = mp.mpf(rho[p]) * mp.exp((mp.mpf(psi[i]) - d2[i] / 2) / eps) / Z
        near = [j for j in range(len(y)) if mp.sqrt(d2[j]) < mp.mpf(radius)]
        if not near:
            jn = min(range(len(y)), key=lambda j: d2[j])
            bary.append([mp.mpf(y[jn][k]) for k in range(dim)])
            continue
        wts = {j: mp.mpf(nu[j]) * mp.exp((mp.mpf(psi[j]) - d2[j] / 2) / eps) for j in near}
        W = sum(wts.values())
        bary.append([sum(w * mp.mpf(y[j][k]) for j, w in wts.items()) / W for k in range(dim)])
    return cond, bary


def points_case(name, dim, N, P, seed, eps, R, radius):
    rng = np.random.default_rng(seed + 400)
    y = rng.random((N, dim))
    nu = rng.random(N) + 0.2
    nu = nu / nu.sum()
    psi = 0.05 * rng.standard_normal(N)
    pts = rng.random((P, dim)) * 1.2 - 0.1
    rho = 0.5 + rng.random(P)
    idx = [0, N // 2, N - 1]
    cond, bary = point_queries(dim, y, nu, psi, eps, R, pts, rho, idx, radius)
    return {"name": name, "dim": dim, "eps": eps, "R": R, "radius": radius,
            "y": y.ravel().tolist(), "nu": nu.tolist(), "psi": psi.tolist(),
            "points": pts.ravel().tolist(), "rho": rho.tolist(), "indices": idx,
            "cond": [[mp.nstr(v, 40) for v in r] for r in cond],
            "bary_map": [[mp.nstr(v, 40) for v in r] for r in bary]}


def main_points():
    cases = [
        points_case("2d", 2, 12, 40, 31, 0.02, 0.45, 0.3),
        points_case("3d", 3, 15, 30, 32, 0.05, 0.6, 0.35),
        points_case("2d_small_radius", 2, 10, 30, 33, 0.01, 1.7976931348623157e308, 0.12),
    ]
    out = os.path.join(HERE, "points_cases.json")
    with open(out, "w") as f:
        json.dump(cases, f)
    print("wrote", out)


def main_refine():
    cases = [
        refine_case("2d_truncated_lse", 2, 3, 6, 20, 11, 0.02, 0.35, True),
        refine_case("2d_untruncated_plain", 2, 2, 5, 12, 12, 0.05, 1.7976931348623157e308, False),
        refine_case("3d_truncated_lse", 3, 2, 5, 14, 13, 0.03, 0.45, True),
        refine_case("2d_tight_lse", 2, 3, 7, 18, 14, 0.01, 0.12, True),
    ]
    out = os.path.join(HERE, "refine_cases.json")
    with open(out, "w") as f:
        json.dump(cases, f)
    print("wrote", out)


def main():
    cases = [
        case("2d_untruncated_lse", 2, 2, 5, 1, 0.05, 1.7976931348623157e308, True),
        case("2d_untruncated_plain", 2, 2, 5, 2, 0.05, 1.7976931348623157e308, False),
        case("3d_truncated_lse", 3, 2, 7, 3, 0.1, 0.6, True),
        case("2d_truncated_lse", 2, 3, 9, 4, 0.02, 0.45, True),
        case("2d_plain_underflow", 2, 2, 3, 5, 2e-4, 1.7976931348623157e308, False,
             cluster=np.array([0.02, 0.02])),
    ]
    out = os.path.join(HERE, "tiny_cases.json")
    with open(out, "w") as f:
        json.dump(cases, f)
    print("wrote", out)


if __name__ == "__main__":
    main()
    main_refine()
    main_bary()
    main_points()
