"""Pin the cull oracle (oracle/orc_cull.c) against brute force, hand-made
known answers and independent library routines."""
import ctypes as C

import numpy as np
import pytest
import scipy.ndimage as ndi

from tests.helpers import cull_params, make_cull_grid

NODATA = np.float32(-2e9)


def _seq_cull_bruteforce(offR, offA, b=9, nGood=17, maxA=1.0, maxR=0.75):
    """Literal sequential pass: raster order, in place, median = the
    floor(n/2)-th smallest (1-based) of the valid box values."""
    R, A = offR.copy(), offA.copy()
    nA, nR = R.shape
    for i in range(nA):
        for j in range(nR):
            box = (slice(max(0, i - b // 2), min(nA, i + b // 2 + 1)), slice(max(0, j - b // 2), min(nR, j + b // 2 + 1)))
            v = (R[box] > 1 - 2e9) & (A[box] > 1 - 2e9)
            n = int(v.sum())
            if n > nGood:
                k = max(n // 2, 1)
                mr = np.sort(R[box][v])[k - 1]
                ma = np.sort(A[box][v])[k - 1]
                if abs(float(np.float32(A[i, j] - ma))) > maxA or abs(float(np.float32(R[i, j] - mr))) > maxR:
                    R[i, j] = A[i, j] = NODATA
            else:
                R[i, j] = A[i, j] = NODATA
    return R, A


def test_cull_pass_matches_literal_sequential_pass(oracle32):
    offR, offA, corr, typ = make_cull_grid(40, 50, seed=1)
    cp, arrs = cull_params(offR, offA, corr, typ)
    oracle32.cull("st_cull_st_data", cp)
    R, A = _seq_cull_bruteforce(offR, offA)
    assert np.array_equal(arrs["offR"], R) and np.array_equal(arrs["offA"], A)


def _jacobi_cull(offR, offA, b=9, nGood=17, maxA=1.0, maxR=0.75):
    """Every decision from the pass-start state (what a naive parallel pass
    computes)."""
    R, A = offR.copy(), offA.copy()
    nA, nR = R.shape
    for i in range(nA):
        for j in range(nR):
            box = (slice(max(0, i - b // 2), min(nA, i + b // 2 + 1)), slice(max(0, j - b // 2), min(nR, j + b // 2 + 1)))
            v = (offR[box] > 1 - 2e9) & (offA[box] > 1 - 2e9)
            n = int(v.sum())
            cull = True
            if n > nGood:
                k = max(n // 2, 1)
                mr = np.sort(offR[box][v])[k - 1]
                ma = np.sort(offA[box][v])[k - 1]
                cull = abs(float(np.float32(offA[i, j] - ma))) > maxA or abs(float(np.float32(offR[i, j] - mr))) > maxR
            if cull:
                R[i, j] = A[i, j] = NODATA
    return R, A


def test_cull_in_place_order_matters(oracle32):
    """Across a noisy step edge the raster-order in-place pass and a
    parallel pass disagree; the oracle follows the in-place order."""
    rng = np.random.default_rng(6)
    offR = np.zeros((24, 24), np.float32)
    offA = np.zeros((24, 24), np.float32)
    offR[:, 12:] = 1.0
    offR += (0.3 * rng.standard_normal(offR.shape)).astype(np.float32)
    cp, arrs = cull_params(offR, offA, np.ones_like(offR), np.ones(offR.shape, np.uint8))
    oracle32.cull("st_cull_st_data", cp)
    Rs, _ = _seq_cull_bruteforce(offR, offA)
    Rj, _ = _jacobi_cull(offR, offA)
    assert np.array_equal(arrs["offR"], Rs)
    assert not np.array_equal(Rs, Rj)


def test_plane_fit_matches_truncated_svd(oracle64):
    """cullStats' svdfit: least squares with singular values below 1e-5 of
    the largest dropped == numpy pinv with rcond = 1e-5."""
    lib = oracle64.lib
    lib.orc_lsq3.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    rng = np.random.default_rng(3)
    for c0, i0, j0 in ((5.0, 3, 4), (0.5, 400, 700), (2e-4, 40, 70), (30.0, 800, 20)):
        ii, jj = np.mgrid[i0 - 4:i0 + 5, j0 - 4:j0 + 5]
        keep = rng.random(81) > 0.2
        xi = ii.ravel()[keep].astype(np.int32)
        xj = jj.ravel()[keep].astype(np.int32)
        y = (0.3 + 0.01 * xi - 0.02 * xj + 0.05 * rng.standard_normal(xi.size))
        x = np.zeros(3)
        lib.orc_lsq3(c0, xi.ctypes.data, xj.ctypes.data, y.ctypes.data, xi.size, x.ctypes.data)
        Amat = np.stack([np.full(xi.size, c0), xi, xj], 1).astype(np.float64)
        ref = np.linalg.pinv(Amat, rcond=1e-5) @ y
        assert np.allclose(Amat @ x, Amat @ ref, rtol=1e-7, atol=1e-7)


def test_smooth_closed_form(oracle32):
    """cullSmooth on a field with no holes: interior points are the
    weighted box mean with weights 1 and 0.5 at +/- s/2 for even s."""
    nA, nR = 20, 24
    ii, jj = np.mgrid[0:nA, 0:nR]
    R = (0.1 * ii + 0.2 * jj).astype(np.float32)
    A = (0.05 * ii * jj).astype(np.float32)
    cp, arrs = cull_params(R, A, np.ones_like(R), np.ones(R.shape, np.uint8))
    arrs["sigmaA"][:] = 1.0
    arrs["sigmaR"][:] = 1.0
    oracle32.cull("st_cull_smooth", cp)
    wa = np.array([0.5, 1, 1, 1, 1, 1, 1, 1, 0.5])
    wr = np.array([0.5, 1, 0.5])
    w = np.outer(wa, wr)
    for (i, j) in ((10, 10), (6, 5)):
        ref = (w * R[i - 4:i + 5, j - 1:j + 2]).sum() / w.sum()
        assert abs(arrs["offRS"][i, j] - ref) < 1e-5
        # effective looks: sum of weights x 1 (type 1) -> sigma / sqrt(looks)
        assert abs(arrs["sigmaA"][i, j] - 1 / np.sqrt(w.sum())) < 1e-6


def test_islands_bounding_box_rule(oracle32):
    """cullIslands removes 4-connected components whose bounding box spans
    <= islandThresh in both axes, except the last-labelled component."""
    rng = np.random.default_rng(7)
    valid = rng.random((30, 40)) < 0.45
    RS = np.where(valid, 1.0, -2e9).astype(np.float32)
    cp, arrs = cull_params(RS, RS, np.ones_like(RS), np.ones(RS.shape, np.uint8), islandThresh=3)
    arrs["offRS"][:] = RS
    arrs["offAS"][:] = RS
    oracle32.cull("st_cull_islands", cp)
    lab, n = ndi.label(valid)  # 4-connectivity, labels in raster order of first pixel
    keep = valid.copy()
    for l in range(1, n):
        ys, xs = np.nonzero(lab == l)
        if xs.max() - xs.min() <= 3 and ys.max() - ys.min() <= 3:
            keep[lab == l] = False
    assert np.array_equal(arrs["offRS"] > -1e9, keep)
