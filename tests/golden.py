"""Loader for the committed golden fixtures (converted from the reference's JAX golden vectors)."""
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_NPZ = None
MANIFEST = json.load(open(os.path.join(HERE, "nbody_manifest.json")))
KNOWN = json.load(open(os.path.join(HERE, "known_answers.json")))
N_PART, BOXSIZE = 16, 100.0
# configurations of export_jax_nbody_reference.py; the raw linear-resampling one records an upstream
# bug and is not a parity target (NR/test/nbody_parity_tests.jl:38)
CONFIGS = sorted(g for g in MANIFEST["groups"] if g not in ("cosmo", "ics", "bullfrog_resample_linear"))


def arr(key):
    global _NPZ
    if _NPZ is None:
        _NPZ = np.load(os.path.join(HERE, "nbody_reference.npz"))
    return _NPZ[key]


def particles(key):
    """golden (n,n,n,3) [i,j,k,c] -> (3,n,n,n) [c,k,j,i] (column-major (n,n,n,3) of the reference)."""
    return np.ascontiguousarray(arr(key).transpose(3, 2, 1, 0))


def fphi():
    """golden (n,n,n//2+1) [i,j,kz] -> (n//2+1,n,n) [kz,j,i]."""
    f = arr("ics/fphi_re") + 1j * arr("ics/fphi_im")
    return np.ascontiguousarray(f.transpose(2, 1, 0))


def kwargs(cfg):
    return MANIFEST["groups"][cfg]["kwargs"]


def rel(x, y):
    """max|x - y| / max|y| (NR/test/nbody_parity_tests.jl:11)."""
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    return float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
