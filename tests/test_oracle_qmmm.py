"""The restatement's QM/MM embedding routines (src/fmm.F90:1134-1336).
CPU only.

interpolate_electron_fields is pinned to the reference's own output
(tests/golden/water999_qmmm.npz).  The reference's get_boxes and
get_multipoles abort in finalize before returning (they deallocate
interaction lists they never allocate, src/fmm.F90:318-321), so their
restatement is pinned by brute force instead: the classification rule
evaluated independently per node from the exported tree, and every far
box's multipoles as direct moment sums over the sites it contains."""
from math import comb

import numpy as np
import pytest

import oracle
from conftest import block_rel, load_golden

CASES = [("t5_n64_p4", 0.5, 64, 4), ("t3_n16_p6", 0.3, 16, 6), ("t5_n32_p5", 0.5, 32, 5)]


@pytest.mark.parametrize("key,theta,ncrit,p", CASES)
def test_interpolate_vs_reference(key, theta, ncrit, p):
    g = load_golden("water999_qmmm.npz")
    F = int(g[key + "_field_order"])
    got = oracle.interpolate_electron_fields(g["coordinates"], g[key + "_local_expansion"], g[key + "_box_indices"],
                                             theta, ncrit, p, F, impl="oracle")
    for d in range(got.shape[2]):
        assert max(block_rel(got[:, :, d], g[key + "_field"][:, :, d], F)) < 1e-12


def _classify_brute(t, targets, theta):
    """Per node: close = any target with |target - c| theta <= rmax; visited =
    every ancestor close; far boxes and close leaves in node order."""
    nn = len(t["r"])
    c = t["centers"]
    close = np.array([any(oracle.norm2(tg - c[i]) * theta <= t["rmax"][i] for tg in targets) for i in range(nn)])
    near, far = [], []
    for i in range(nn):
        a, visited = t["parent"][i], True
        while a >= 0:
            visited &= bool(close[a])
            a = t["parent"][a]
        if not visited:
            continue
        if not close[i]:
            far.append(i)
        elif t["is_leaf"][i]:
            near.extend(t["members"][t["mem_off"][i]:t["mem_off"][i + 1]] + 1)
    return np.array(near, np.int32), np.array(far, np.int32)


@pytest.mark.parametrize("key,theta,ncrit,p", CASES)
def test_get_boxes_classification_brute_force(key, theta, ncrit, p):
    g = load_golden("water999_qmmm.npz")
    x, tg = g["coordinates"], g["targets"]
    near, bc, bi = oracle.get_boxes(x, tg, theta, ncrit, p, impl="oracle")
    t = oracle.tree(x, ncrit, theta)
    bnear, bfar = _classify_brute(t, tg, theta)
    np.testing.assert_array_equal(near, bnear)
    np.testing.assert_array_equal(bi, bfar + 1)
    np.testing.assert_array_equal(bc, t["centers"][bfar])
    assert len(bfar) > 0 and len(bnear) > 0
    np.testing.assert_array_equal(bi, g[key + "_box_indices"])


def _moments(x, m, c, p):
    """Cartesian multipoles about c up to order p of sites with charges and dipoles."""
    comps = [(a, b, k - a - b) for k in range(p + 1) for a in range(k, -1, -1) for b in range(k - a, -1, -1)]
    src = comps[:m.shape[1]]
    d = x - c
    out = np.zeros(len(comps))
    for ti, t in enumerate(comps):
        for si, s in enumerate(src):
            if all(s[q] <= t[q] for q in range(3)):
                e = [t[q] - s[q] for q in range(3)]
                out[ti] += comb(t[0], s[0]) * comb(t[1], s[1]) * comb(t[2], s[2]) * \
                    np.sum(d[:, 0] ** e[0] * d[:, 1] ** e[1] * d[:, 2] ** e[2] * m[:, si])
    return out


def test_get_multipoles_direct_moments():
    g = load_golden("water999_qmmm.npz")
    x, m2, tg = g["coordinates"], g["multipoles"], g["targets"]
    theta, ncrit, p = 0.5, 64, 4
    near, bc, bm = oracle.get_multipoles(x, m2, tg, theta, ncrit, p, impl="oracle")
    t = oracle.tree(x, ncrit, theta)
    _, far = _classify_brute(t, tg, theta)
    nl = t["nleaf"]
    for k, node in enumerate(far):
        # sites of the box: members of every leaf in its subtree (pre-order: node .. node + size)
        sub = [node]
        for j in range(node + 1, len(nl)):
            a = t["parent"][j]
            while a >= 0 and a != node:
                a = t["parent"][a]
            if a == node:
                sub.append(j)
        sites = np.concatenate([t["members"][t["mem_off"][j]:t["mem_off"][j + 1]] for j in sub])
        assert len(sites) == nl[node]
        for d in range(2):
            ref = _moments(x[sites], m2[sites, :, d], t["centers"][node], p)
            assert np.linalg.norm(bm[k, :, d] - ref) <= 1e-12 * np.linalg.norm(ref), (k, d)
