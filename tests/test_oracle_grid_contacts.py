"""Pins the oracle's cell list and contact counts by brute force."""
import numpy as np
import pytest

import pyoracle as po
from helpers import radius_for, relaxed, walk


def brute_cell_list(r, csl, cps):
    r = np.asarray(r, np.float32)
    ir = np.floor(r / np.float32(csl)).astype(np.int64)
    iofst = (cps // 2) * (1 + cps + cps * cps)
    uci = (iofst + ir[:, 0] + ir[:, 1] * cps + ir[:, 2] * cps * cps).astype(np.uint32)
    soi = np.argsort(uci, kind="stable").astype(np.uint32)
    sci = uci[soi]
    n_c = cps**3
    beg, end = np.zeros(n_c, np.uint32), np.zeros(n_c, np.uint32)
    cells, first, counts = np.unique(sci, return_index=True, return_counts=True)
    beg[cells] = first
    end[cells] = first + counts
    beg[sci[0]] = 0
    return uci, sci, soi, beg, end


@pytest.mark.parametrize("N,cvf,seed", [(1, 0.2, 1), (2, 0.2, 2), (1000, 0.2, 3), (20000, 0.3, 4)])
def test_cell_list_matches_stable_sort(N, cvf, seed):
    R_n = radius_for(N, cvf) if N > 2 else 5.0
    r = walk(N, R_n, seed).astype(np.float32)
    cps = int(2 * np.ceil(np.float32(R_n) / np.float32(2.0)))
    out = po.cell_list(r, 2.0, cps)
    uci, sci, soi, beg, end = brute_cell_list(r, 2.0, cps)
    assert out["bad"] == 0
    assert np.array_equal(out["uci"], uci)
    assert np.array_equal(out["sci"], sci)
    assert np.array_equal(out["soi"], soi)
    assert np.array_equal(out["beg"], beg)
    assert np.array_equal(out["end"], end)


def brute_contacts(r, rc, bs, nb):
    r = np.asarray(r, np.float32)
    m = np.zeros((nb, nb), np.uint32)
    rc2 = np.float32(rc) * np.float32(rc)
    for i in range(len(r)):
        d = r[i + 1:] - r[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        js = np.nonzero(d2 < rc2)[0] + i + 1
        for j in js:
            m[i // bs, j // bs] += 1
            m[j // bs, i // bs] += 1
    return m


@pytest.mark.parametrize("N,bs", [(300, 1), (2000, 16)])
def test_contacts_match_brute_force(N, bs):
    R_n = radius_for(N, 0.3)
    r = walk(N, R_n, 5).astype(np.float32)
    nb = (N + bs - 1) // bs
    want = brute_contacts(r, 1.5, bs, nb)
    assert np.array_equal(po.contacts(r, 1.5, bs, nb), want)
    assert np.array_equal(po.contacts_grid(r, 1.5, bs, nb), want)
    assert np.array_equal(want, want.T)


def test_ccp_contact_map_small():
    # 8 beads: 0-3 in one tight cluster, 4-7 spread out; cms = 2 pixels,
    # packed lower triangle [(0,0), (1,0), (1,1)] (chrana.cu:59-60)
    r = np.zeros((8, 3), np.float32)
    r[:4] = [[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]]
    r[4:] = [[10, 0, 0], [13, 0, 0], [10, 3, 0], [10, 0, 3]]
    cm = po.ccp_contact_map(r)
    assert cm.shape == (3,)
    assert cm[0] == pytest.approx(1000.0)  # 16 of 16 pairs incl. self
    assert cm[1] == 0.0
    assert cm[2] == pytest.approx(1000.0 * 4 / 16)  # only the 4 self pairs


def test_ccp_contact_map_index_layout():
    # pixel i_a maps to (i_x, i_y) = np.tril_indices order, which is how
    # the reference's viewer unpacks contact-map.bin (Scripts/viewana.py:127-129)
    N = 64
    r = relaxed(N, radius_for(N, 0.3), 3).astype(np.float32)
    cm = po.ccp_contact_map(r)
    cms = N // 4
    ix, iy = np.tril_indices(cms)
    ref = np.zeros(len(cm), np.float32)
    d = np.linalg.norm(r[:, None, :] - r[None, :, :], axis=2) < 2.0
    for a, (x, y) in enumerate(zip(ix, iy)):
        ref[a] = 1000.0 * d[4 * x:4 * x + 4, 4 * y:4 * y + 4].sum() / 16
    assert np.allclose(cm, ref)
