"""Worker process that owns the HIP library for callers that load PyTorch.

PyTorch bundles its own HIP runtime; two HIP runtimes in one process do not share the device,
so a process that imports torch drives the library through this worker instead. Messages are
(command, payload) tuples pickled over the worker's stdin/stdout; both ends are this package.

Commands: ("init", (params, geometry, options, kwargs)), ("energy", coords) -> (energy,
charges), ("gradients", None) -> gradients, ("close", None).
"""
import pickle
import sys


def serve(inp, out) -> None:
    from . import api
    calc = None
    latvecs = None
    while True:
        try:
            cmd, payload = pickle.load(inp)
        except EOFError:
            break
        try:
            if cmd == "init":
                params, geo, options, kwargs = payload
                calc = api.DftbPlus(params, geo, options, **kwargs)
                latvecs = geo.latvecs
                res = calc.get_nr_orbitals()
            elif cmd == "energy":
                calc.set_geometry(payload, latvecs)
                res = (calc.get_energy(), calc.get_gross_charges(), calc.n_iter, calc.converged)
            elif cmd == "gradients":
                res = calc.get_gradients()
            elif cmd == "close":
                if calc is not None:
                    calc.close()
                pickle.dump(("ok", None), out)
                out.flush()
                break
            else:
                raise ValueError(f"unknown command {cmd!r}")
            pickle.dump(("ok", res), out)
        except Exception as exc:   # reported to the caller, the worker stays alive
            pickle.dump(("error", f"{type(exc).__name__}: {exc}"), out)
        out.flush()


if __name__ == "__main__":
    import os
    # the protocol owns the original stdout; anything else printed goes to stderr
    proto = os.fdopen(os.dup(1), "wb")
    os.dup2(2, 1)
    serve(sys.stdin.buffer, proto)
