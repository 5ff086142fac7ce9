"""Block statistics of an AFQMC run (host side, numpy).

Mirrors ad_afqmc/stat_utils.py: ``blocking_analysis`` (:8-52) and ``reject_outliers``
(:55-59).  The reference stores the re-blocked weights and energies in float32 arrays
(stat_utils.py:25-26) and forms the error from them; ``block_dtype`` keeps that as the
default.  The reference's committed run (examples/samples_raw.dat -> ene_err.txt) was
analysed with float64 blocks: ``block_dtype=np.float64`` reproduces its error estimate
bit for bit, float32 agrees to ~1e-8 relative.
"""

import numpy as np


def blocking_analysis(weights, energies, neql=0, printQ=False, writeBlockedQ=False,
                      block_dtype=np.float32):
    """Weighted mean and plateau error of re-blocked samples.

    For block sizes b = 1, 2, ... < n/2 the samples are grouped into floor(n/b) blocks;
    each block's weight is the sum of its weights and its energy the weighted mean; the
    error is the weighted standard error of the block energies,
    sqrt(sum_j w_j (e_j - mean)^2 / (v1 - v2/v1) / (nblocks - 1)), v1 = sum w,
    v2 = sum w^2.  The first block size whose error is no more than 1.05 times the
    previous one ends the search: the plateau error is the larger of the two.
    Returns (mean energy over all samples, plateau error or None).
    """
    weights = np.asarray(weights)[neql:]
    energies = np.asarray(energies)[neql:]
    n = weights.shape[0]
    we = weights * energies
    mean_energy = we.sum() / weights.sum()
    if printQ:
        print(f"#\n# Mean: {mean_energy:.8e}", flush=True)
        print("# Block size    # of blocks         Mean                Error", flush=True)
    prev = 0.0
    plateau = None
    b = 1
    while b < n / 2.0:
        nb = n // b
        # block sums in float64, stored in block_dtype
        bw = np.empty(nb, dtype=block_dtype)
        be = np.empty(nb, dtype=block_dtype)
        for j in range(nb):
            bw[j] = weights[j * b:(j + 1) * b].sum()
            be[j] = we[j * b:(j + 1) * b].sum() / bw[j]
        v1 = bw.sum()
        v2 = (bw ** 2).sum()
        mean = np.multiply(bw, be).sum() / v1
        # the block count is an int64 there, so the last division and the root are
        # float64 whatever block_dtype is
        err = (np.multiply(bw, (be - mean) ** 2).sum() / (v1 - v2 / v1)
               / np.int64(nb - 1)) ** 0.5
        if writeBlockedQ:
            np.savetxt(f"samples_blocked_{b}.dat", np.stack((bw, be)).T)
        if printQ:
            print(f"  {b:5d}           {nb:6d}       {mean:.8e}       {err:.6e}", flush=True)
        if err <= 1.05 * prev and plateau is None:
            plateau = max(err, prev)
        prev = err
        b += 1
    if printQ and plateau is not None:
        print(f"# Stochastic error estimate: {plateau:.6e}\n#", flush=True)
    return mean_energy, plateau


def reject_outliers(data, obs, m=10.0):
    """Rows of `data` whose column `obs` lies within m median absolute deviations of
    its median; returns (kept rows, boolean mask)."""
    d = np.abs(data[:, obs] - np.median(data[:, obs]))
    mdev = np.median(d) + 1.0e-10
    s = d / mdev if mdev else 0.0
    keep = s < m
    return data[keep], keep
