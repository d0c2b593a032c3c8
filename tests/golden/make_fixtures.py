"""Convert the reference's JAX golden vectors to committed fixtures.

Source: native/DiscoDJNative/test/reference/jax_nbody_reference.h5 (FP64, 16^3 particles,
L = 100 Mpc/h, written by export_jax_nbody_reference.py of the reference).  h5py is not
installed here, so every dataset is extracted with `h5dump -b LE` and every attribute is
read from `h5dump -a` text output.  Run once in a container that holds the reference:

    python tests/golden/make_fixtures.py

Output (this directory):
    nbody_reference.npz   every dataset, keyed "<group>/<name>", C-order exactly as in the file
                          (particle arrays (n,n,n,3) indexed [i,j,k,c], fphi (n,n,n//2+1))
    nbody_manifest.json   per group: shape/dtype of each dataset and the run_nbody keyword set;
                          cosmo attributes (Dplus_unnormed_at_1, Omega_m)
    known_answers.json    the N-GenIC / ranlxd1 constants of the reference's runtests.jl:406-421
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

H5 = "/root/reference/native/DiscoDJNative/test/reference/jax_nbody_reference.h5"
H5DUMP = "/opt/conda/bin/h5dump"
HERE = os.path.dirname(os.path.abspath(__file__))


def _run(args):
    return subprocess.run([H5DUMP] + args, check=True, capture_output=True, text=True).stdout


def _objects():
    groups, datasets = [], []
    for line in _run(["-n", H5]).splitlines():
        m = re.match(r"\s*(group|dataset)\s+(\S+)", line)
        if m:
            (groups if m.group(1) == "group" else datasets).append(m.group(2))
    return groups, datasets


def _dataset(path):
    hdr = _run(["-H", "-d", path, H5])
    dims = re.search(r"SIMPLE \{ \( ([0-9, ]+) \)", hdr).group(1)
    shape = tuple(int(s) for s in dims.split(","))
    is_float = "H5T_IEEE_F64LE" in hdr
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "d.bin")
        _run(["-d", path, "-b", "LE", "-o", out, H5])
        arr = np.fromfile(out, dtype="<f8" if is_float else "<i8")
    return arr.reshape(shape)


def _attribute(path):
    txt = _run(["-m", "%.17g", "-a", path, H5])
    body = txt.split("DATA {", 1)[1]
    val = re.search(r"\(0\):\s*(.*)", body).group(1).strip()
    if "H5T_STRING" in txt:
        return val.strip('"')
    if "H5T_ENUM" in txt:
        return val == "TRUE"
    if "H5T_IEEE_F64" in txt:
        return float(val)
    return int(val)


def main():
    groups, datasets = _objects()
    arrays, manifest = {}, {"source": "native/DiscoDJNative/test/reference/jax_nbody_reference.h5",
                            "groups": {}}
    for d in datasets:
        a = _dataset(d)
        arrays[d.lstrip("/")] = a
        g, name = d.lstrip("/").split("/", 1)
        manifest["groups"].setdefault(g, {"datasets": {}})["datasets"][name] = list(a.shape)
    for g in groups:
        if g == "/":
            continue
        gname = g.lstrip("/")
        entry = manifest["groups"].setdefault(gname, {"datasets": {}})
        if gname == "cosmo":
            entry["attrs"] = {k: _attribute(f"{g}/{k}") for k in ("Dplus_unnormed_at_1", "Omega_m")}
        elif gname != "ics":
            raw = _run(["-a", f"{g}/kwargs", H5])
            m = re.search(r'\(0\):\s*"(.*)"\s*$', raw, re.M)
            kw = json.loads(m.group(1).replace('\\"', '"'))
            entry["kwargs"] = kw
    np.savez_compressed(os.path.join(HERE, "nbody_reference.npz"), **arrays)
    with open(os.path.join(HERE, "nbody_manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    # runtests.jl:406-421 (N-GenIC bit-exact block): ranlxd1 seed 42 first five uniforms,
    # the norm of ngenic_field_gsl(42, 8) and the shape checks.
    known = {
        "ranlxd1_seed42_first5": [0.66962007120990563, 0.26813969602963539, 0.0948083864381708,
                                  0.31262174009276222, 0.34292274337927253],
        "ranlxd1_atol": 1e-15,
        "ngenic_field_gsl_42_8_norm": 13.448527,
        "ngenic_field_gsl_42_8_norm_atol": 1e-3,
        "ngenic_field_gsl_42_8_shape": [8, 8, 5],
        "source": "native/DiscoDJNative/test/runtests.jl:406-421",
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(known, f, indent=1)
    print("wrote", len(arrays), "datasets")


if __name__ == "__main__":
    sys.exit(main())
