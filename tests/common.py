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
