"""ORACLE — numpy restatement of the lookup-table uncertainty loop of
Lookup/lookup.py:347-376 (makelookup) with find_nearest (lookup.py:296-299).

Test infrastructure only (tests/ may import it as the checker).  It draws
the same normals as ismc_table_uncertainty (Philox4x32-10, key = seed,
counter = (cell, 0, pair, 0x7ab1e), u = (w + 0.5) 2^-32, Box-Muller) and
then runs the reference's per-cell loop literally: every sample is looked
up with a full argmin over the grid axis.

Pinned by: Philox against the Random123 known answers (the same vectors
as the C oracle's), and the loop against closed forms in the tests (a
constant table has zero spread; a table linear in R and T on a wide grid
has std = slope x the draw's std).
"""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)
TAG = 0x7ab1e


def philox4x32(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 (Salmon et al. SC'11) on uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) & M32 for x in (c0, c1, c2, c3))
    k0 = np.uint64(k0) & M32
    k1 = np.uint64(k1) & M32
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c0
        p1 = np.uint64(0xCD9E8D57) * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ k0
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ k1
        c1 = p1 & M32
        c3 = p0 & M32
        c0, c2 = n0 & M32, n2 & M32
        k0 = (k0 + np.uint64(0x9E3779B9)) & M32
        k1 = (k1 + np.uint64(0xBB67AE85)) & M32
    return [x.astype(np.uint32) for x in (c0, c1, c2, c3)]


def normals(cell, M, seed):
    """(zR[M], zT[M]) standard normals of one cell."""
    pairs = (M + 1) // 2
    p = np.arange(pairs, dtype=np.uint64)
    w = philox4x32(np.full(pairs, cell), 0, p, TAG, seed & 0xFFFFFFFF, seed >> 32)
    u = [(x.astype(np.float64) + 0.5) * 2.0 ** -32 for x in w]
    zR, zT = [], []
    for h in range(2):
        r = np.sqrt(-2.0 * np.log(u[2 * h]))
        ang = 6.283185307179586 * u[2 * h + 1]
        zR.append(r * np.cos(ang))
        zT.append(r * np.sin(ang))
    # sample 2p + h
    zR = np.stack(zR, 1).ravel()[:M]
    zT = np.stack(zT, 1).ravel()[:M]
    return zR, zT


def find_nearest(array, value):
    """lookup.py:296-299."""
    return (np.abs(np.asarray(array) - value)).argmin()


def uncertainty(mua, mus, rstd, tstd, Rvals, Tvals, M, seed, cells=None):
    """Percent uncertainty per cell [i (R)][j (T)], the reference's loop;
    `cells` restricts it to a list of (i, j) (others stay NaN)."""
    P = len(Rvals)
    ua = np.full((P, P), np.nan)
    us = np.full((P, P), np.nan)
    todo = cells if cells is not None else [(i, j) for i in range(P) for j in range(P)]
    for i, j in todo:
        R, T = Rvals[i], Tvals[j]
        R_std, T_std = rstd[i][j], tstd[i][j]
        if np.isnan(R_std) or np.isnan(T_std) or R_std < 0 or T_std < 0:
            continue
        zR, zT = normals(i * P + j, M, seed)
        R_dist = R + R_std * zR
        T_dist = T + T_std * zT
        muas, muss = [], []
        for k in range(M):
            R_idx = find_nearest(Rvals, R_dist[k])
            T_idx = find_nearest(Tvals, T_dist[k])
            muas.append(mua[R_idx][T_idx])
            muss.append(mus[R_idx][T_idx])
        R_idx = find_nearest(Rvals, R)
        T_idx = find_nearest(Tvals, T)
        ua[i, j] = 100 * np.std(np.array(muas)) / mua[R_idx][T_idx]
        us[i, j] = 100 * np.std(np.array(muss)) / mus[R_idx][T_idx]
    return ua, us
