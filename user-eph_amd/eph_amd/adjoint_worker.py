"""Process that owns the GPU trajectory for the torch binding (eph_amd.adjoint).

Run as `python -m eph_amd.adjoint_worker`: it imports the library and numpy
only -- never torch -- so the HIP runtime of libeph_amd.so is the only one in
the process.  Pickled requests arrive on stdin, replies leave on stdout; the
first request is the trajectory description.
    ("forward", beta, v0)  -> ("ok", loss, Ee, grid_term)
    ("backward",)          -> ("ok", g_beta, g_v0)
    ("info",)              -> ("ok", dict)
    ("close",)             -> process ends
Errors come back as ("error", message).
"""
import os
import pickle
import sys
import tempfile

import numpy as np


def build_trajectory(spec):
    """Harness system + fix + Adjoint from a plain-data description."""
    import eph_amd as ea
    from eph_amd import inputs
    tmp = spec.get("tmpdir") or tempfile.mkdtemp(prefix="ephadj")
    beta = os.path.join(tmp, "adjoint.beta")
    names = inputs.adjoint_beta(beta, n_beta=spec.get("n_beta", 64))
    cells, a = spec["cells"], spec.get("a", 3.615)
    md = ea.MD(cells, a=a, ntypes=1, seed=spec.get("md_seed", 7), T=spec.get("T", 600.0),
               pka_eV=spec.get("pka_eV", 0.0))
    inf = md.info()
    g = spec.get("grid", cells)
    args = inputs.fix_args(beta, names, seed=spec.get("fix_seed", 4711), flags=spec.get("flags", "7"),
                           grid=(g, g, g), T_infile=spec.get("T_infile", "NULL"))
    fix = ea.FixEPH(args, 1, inf["boxlo"], inf["boxhi"], spec["dt"], inputs.BOLTZ, inputs.FTM2V,
                    masses=[inputs.MASS_CU])
    fix.init()
    md.attach(fix)
    target = None
    if spec.get("target_seed") is not None:
        rng = np.random.default_rng(spec["target_seed"])
        target = 300.0 + spec.get("target_amp", 50.0) * rng.uniform(size=g * g * g)
    adj = ea.Adjoint(md, fix, spec["nsteps"], spec.get("every", 20), target)
    meta = dict(beta_file=beta, names=names, args=args, target=target, boxlo=inf["boxlo"],
                boxhi=inf["boxhi"])
    return md, fix, adj, meta


class _Pipe:
    def __init__(self, rfile, wfile):
        self.r, self.w = rfile, wfile

    def recv(self):
        return pickle.load(self.r)

    def send(self, obj):
        pickle.dump(obj, self.w, protocol=pickle.HIGHEST_PROTOCOL)
        self.w.flush()


def serve(conn, spec):
    try:
        md, fix, adj, _ = build_trajectory(spec)
        conn.send(("ok", dict(n_elem=adj.n_elem, n_beta=adj.n_beta, nlocal=adj.nlocal,
                              ncells=adj.ncells, beta0=adj.beta0(), v0=adj.v0())))
    except Exception as e:  # reported to the parent, which raises
        conn.send(("error", repr(e)))
        return
    while True:
        msg = conn.recv()
        try:
            if msg[0] == "forward":
                loss, ee, sq = adj.forward(msg[1], msg[2])
                conn.send(("ok", loss, ee, sq))
            elif msg[0] == "backward":
                gb, gv = adj.backward()
                conn.send(("ok", gb, gv))
            elif msg[0] == "close":
                break
            else:
                conn.send(("error", f"unknown request {msg[0]!r}"))
        except Exception as e:
            conn.send(("error", repr(e)))
    adj.close()


def main():
    conn = _Pipe(sys.stdin.buffer, sys.stdout.buffer)
    spec = conn.recv()
    serve(conn, spec)


if __name__ == "__main__":
    main()
