"""Shared helpers of the test suite: oracle construction and error metrics."""
import numpy as np

from simbody_amd import models as M
from oracle_ctypes import OracleModel


def rel_inf(a, b):
    """||a-b||_inf / max(||b||_inf, 1e-300): the per-instance relative error."""
    a = np.asarray(a, np.float64).ravel(); b = np.asarray(b, np.float64).ravel()
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) if b.size else 0.0


def simtk_eq(a, b, n=1):
    """SimTK_TEST_EQ tolerance (Testing.h:102-135): |a-b| <= n*SignificantReal*max(|a|,|b|,1)."""
    sig = np.finfo(np.float64).eps ** 0.875
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return bool(np.all(np.abs(a - b) <= n * sig * np.maximum(np.maximum(np.abs(a), np.abs(b)), 1.0)))


def mixed_trees(seed, n, nmin=4, nmax=16, free_every=3):
    """Small random trees exercising all five mobilizers with general frames."""
    trees = []
    for i in range(n):
        t = M.random_tree(seed, i, nmin=nmin, nmax=nmax, general_frames=True)
        if i % free_every == 0:
            t.mobilizer[0] = M.FREE
        trees.append(t)
    return trees


def ext_trees(seed, n, nmin=4, nmax=16):
    """Random trees (general joint frames) whose non-root bodies draw all ten
    mobilizers but Free uniformly -- Pin, Slider, Universal, Ball, Translation,
    Planar, Cylinder, Screw, Gimbal -- with Screw pitches in [-0.5, 0.5];
    every third root is Free."""
    types = [M.PIN, M.SLIDER, M.UNIVERSAL, M.BALL, M.TRANSLATION, M.PLANAR, M.CYLINDER, M.SCREW, M.GIMBAL]
    trees = []
    for i in range(n):
        t = M.random_tree(seed, i, nmin=nmin, nmax=nmax, general_frames=True)
        rng = np.random.default_rng(seed * 7919 + i)
        t.mobilizer[:] = rng.choice(types, size=t.nb)
        if i % 3 == 0:
            t.mobilizer[0] = M.FREE
        t.pitch = rng.uniform(-0.5, 0.5, size=t.nb)
        trees.append(t)
    return trees


def rev_trees(seed, n, nmin=4, nmax=16, p_rev=0.5):
    """ext_trees with each non-Free mobilizer reversed with probability p_rev."""
    trees = ext_trees(seed, n, nmin, nmax)
    for i, t in enumerate(trees):
        rng = np.random.default_rng(seed * 104729 + i)
        t.reversed = ((rng.uniform(size=t.nb) < p_rev) & (t.mobilizer != M.FREE)).astype(np.int32)
    return trees


def rot_to_quat(R):
    """Unit quaternion (w, x, y, z) of a rotation matrix (Shepperd's method)."""
    t = np.trace(R)
    k = int(np.argmax([t, R[0, 0], R[1, 1], R[2, 2]]))
    if k == 0:
        w = 0.5 * np.sqrt(1 + t); f = 0.25 / w
        q = [w, (R[2, 1] - R[1, 2]) * f, (R[0, 2] - R[2, 0]) * f, (R[1, 0] - R[0, 1]) * f]
    else:
        i, j, l = k - 1, k % 3, (k + 1) % 3
        x = 0.5 * np.sqrt(1 + R[i, i] - R[j, j] - R[l, l]); f = 0.25 / x
        q = [0.0] * 4
        q[0] = (R[l, j] - R[j, l]) * f
        q[1 + i] = x; q[1 + j] = (R[j, i] + R[i, j]) * f; q[1 + l] = (R[l, i] + R[i, l]) * f
    q = np.array(q)
    return q / np.linalg.norm(q)


def weld_trees(seed, n, nmin=4, nmax=16, p_weld=0.3):
    """ext_trees with each non-root body welded to its parent with
    probability p_weld (the root keeps its mobilizer, so every tree moves)."""
    trees = ext_trees(seed, n, nmin, nmax)
    for i, t in enumerate(trees):
        rng = np.random.default_rng(seed * 15485863 + i)
        w = rng.uniform(size=t.nb) < p_weld
        w[0] = False
        t.mobilizer = np.where(w, M.WELD, t.mobilizer).astype(np.int32)
    return trees


def bs_trees(seed, n, nmin=4, nmax=16, p_bs=0.3):
    """ext_trees with each non-root body a BendStretch with probability p_bs
    (and body 2 always one, when the tree has it)."""
    trees = ext_trees(seed, n, nmin, nmax)
    for i, t in enumerate(trees):
        rng = np.random.default_rng(seed * 32452843 + i)
        w = rng.uniform(size=t.nb) < p_bs
        w[0] = False
        w[1:2] = True
        t.mobilizer = np.where(w, M.BENDSTRETCH, t.mobilizer).astype(np.int32)
    return trees
