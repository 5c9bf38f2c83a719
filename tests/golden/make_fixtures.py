"""Turn the reference's committed JLD2 simulation outputs into small fixtures.

Run in a container that has the reference checkout and libhdf5 (the
container this repository is developed in):

    python tests/golden/make_fixtures.py

It compiles ``jld2_dump.c`` against libhdf5, dumps the numeric content of
each JLD2 file and writes ``.npz`` archives next to this script.  Arrays are
stored in Julia index order: ``X[i1, i2, ...]`` equals the Julia array element
``X[i1+1, i2+1, ...]`` (HDF5 stores the reversed dimensions).

Fixtures produced (all data, no code, from simulation/ in the reference):
* ``gs_loopgas_D2_X32.npz``  -- simulation/test_test_D2_X32/gs.jld2:
  energies, gradnorms, the optimised A (= Optim state x), the Optim state
  gradient g(x), xk.value, xk.g_norm.
* ``env_loopgas_D2_X32.npz`` -- simulation/hb_test_D2_X32/basis.jld2:
  converged environment Cs/Es, A, shifted H (honeycomb XX minus E_gs/2),
  tangent basis.
* ``env_ising_D2_X32.npz``   -- simulation/test_test_D2_X32/basis.jld2:
  Ising h=2 converged environment with direction-dependent chi, basis.
* ``es_ising_D2_X32.npz``    -- simulation/test_test_D2_X32/es_*.jld2:
  effH, effN for every momentum on the path (keys ``effH_<px>_<py>``).
"""
import glob
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/simulation"
HDF5 = os.environ.get("HDF5_ROOT", "/opt/conda")


def build_dumper(tmp):
    exe = os.path.join(tmp, "jld2_dump")
    subprocess.check_call([
        "gcc", "-O1", "-o", exe, os.path.join(HERE, "jld2_dump.c"),
        f"-I{HDF5}/include", f"-L{HDF5}/lib", "-lhdf5", f"-Wl,-rpath,{HDF5}/lib"])
    return exe


def dump(exe, path, tmp):
    out = tempfile.mkdtemp(dir=tmp)
    txt = subprocess.check_output([exe, path, out]).decode()
    arrays = {}
    for line in txt.splitlines():
        rec = json.loads(line)
        dt = {"c128": np.complex128, "f64": np.float64, "i64": np.int64}[rec["dtype"]]
        a = np.fromfile(os.path.join(out, rec["file"]), dtype=dt)
        a = a.reshape(rec["shape"]) if rec["shape"] else a.reshape(())
        arrays[rec["name"]] = np.ascontiguousarray(a.T)  # Julia index order
    return arrays


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = build_dumper(tmp)

        g = dump(exe, f"{REF}/test_test_D2_X32/gs.jld2", tmp)
        np.savez(os.path.join(HERE, "gs_loopgas_D2_X32.npz"),
                 energies=g["energies"], gradnorms=g["gradnorms"], A=g["A"],
                 gx=g["xk.metadata.kvvec.2.second"], x=g["xk.metadata.kvvec.3.second"],
                 value=g["xk.value"], g_norm=g["xk.g_norm"])

        h = dump(exe, f"{REF}/hb_test_D2_X32/basis.jld2", tmp)
        np.savez(os.path.join(HERE, "env_loopgas_D2_X32.npz"),
                 A=h["ts.A"], Ad=h["ts.Ad"],
                 **{f"C{i+1}": h[f"ts.Cs.{i}"] for i in range(4)},
                 **{f"E{i+1}": h[f"ts.Es.{i}"] for i in range(4)},
                 H1=h["H.0"], H2=h["H.1"], basis=h["basis"])

        t = dump(exe, f"{REF}/test_test_D2_X32/basis.jld2", tmp)
        np.savez(os.path.join(HERE, "env_ising_D2_X32.npz"),
                 A=t["ts.A"], Ad=t["ts.Ad"],
                 **{f"C{i+1}": t[f"ts.Cs.{i}"] for i in range(4)},
                 **{f"E{i+1}": t[f"ts.Es.{i}"] for i in range(4)},
                 basis=t["basis"])

        es = {}
        for p in sorted(glob.glob(f"{REF}/test_test_D2_X32/es_*.jld2")):
            tag = os.path.basename(p)[3:-5]
            e = dump(exe, p, tmp)
            es[f"effH_{tag}"] = e["effH"]
            es[f"effN_{tag}"] = e["effN"]
            es[f"envB_{tag}"] = e["envB"]
        np.savez(os.path.join(HERE, "es_ising_D2_X32.npz"), **es)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
