"""Snapshots of the particle state at chosen scale factors, through run_nbody's step callback.

The reference writes its N-body snapshots from `step_callback(k, a, X, P)` gated by `callback_when`
(NR/src/nbody/run_nbody.jl:151-159): the displacement psi = X - q at each requested a as an .npy file
(studies/lpt_pdf/copula/djn_snapshots.jl:67-76).  Here a snapshot is one .npz file holding

    psi   (3, n, n, n)  displacement X - q, component-major [c, k, j, i]: the memory layout of the
                        reference's column-major (n, n, n, 3) array, so np.asfortranarray(psi.transpose
                        (3, 2, 1, 0)) is that array element for element
    P     (3, n, n, n)  canonical momentum P = Pi F+(a) (run_nbody.jl:149, 164-171), same layout
    a, step, boxsize, n (scalars)

The state is unwrapped (psi is not reduced modulo the box), as the reference keeps it.  The same files
serve as restart points: read_snapshot returns what run_nbody needs to continue from them.
"""
import os

import numpy as np


def snapshot_name(directory, a, prefix="snapshot"):
    return os.path.join(directory, f"{prefix}_a{a:.6f}.npz")


def write_snapshot(path, psi, P, a, boxsize, step=-1):
    psi = np.asarray(psi)
    np.savez(path, psi=psi, P=np.asarray(P, psi.dtype), a=np.float64(a), step=np.int64(step),
             boxsize=np.float64(boxsize), n=np.int64(psi.shape[1]))


def read_snapshot(path):
    """-> dict(psi, P, a, step, boxsize, n)."""
    with np.load(path, allow_pickle=False) as f:
        return {k: (f[k] if f[k].ndim else f[k].item()) for k in f.files}


def lagrangian_q(n, boxsize, dtype):
    return (np.arange(n) * (boxsize / n)).astype(dtype)


def remove_q(X, boxsize):
    """psi = X - q for (3,n,n,n) positions [c,k,j,i] (X unwrapped, as run_nbody passes it)."""
    n = X.shape[1]
    q = lagrangian_q(n, boxsize, X.dtype)
    psi = np.array(X, copy=True)
    psi[0] -= q[None, None, :]
    psi[1] -= q[None, :, None]
    psi[2] -= q[:, None, None]
    return psi


class SnapshotWriter:
    """callback_when / step_callback pair for run_nbody that writes a snapshot at each of a_targets
    (matched to the step grid within rtol), as the reference's snapshot callback does."""

    def __init__(self, directory, a_targets, boxsize, rtol=1e-8, prefix="snapshot"):
        self.directory, self.boxsize, self.rtol, self.prefix = directory, float(boxsize), rtol, prefix
        self.targets = np.asarray(sorted(a_targets), np.float64)
        self.written = []
        os.makedirs(directory, exist_ok=True)

    def _hit(self, a):
        return bool(np.any(np.abs(self.targets - a) <= self.rtol * np.maximum(1.0, np.abs(self.targets))))

    def when(self, k, a):
        return self._hit(a)

    def __call__(self, k, a, X, P):
        if not self._hit(a):
            return
        path = snapshot_name(self.directory, a, self.prefix)
        write_snapshot(path, remove_q(X, self.boxsize), P, a, self.boxsize, k)
        self.written.append(path)
