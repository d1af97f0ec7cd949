"""Input/output vectors of the reference's assess_nlconvergence
(assess_steady_state.py:10-74), written to tests/golden/nl_vectors.json.

The reference function uses NumPy only, so it is imported from the reference
tree and called on small seeded arrays; nothing else of the reference runs.
Run in the build container, where the reference is mounted:
    python tests/golden/make_nl_vectors.py /root/reference
"""
import contextlib
import io
import json
import os
import sys

import numpy as np


def main(ref):
    sys.path.insert(0, ref)
    from assess_steady_state import assess_nlconvergence
    rng = np.random.Generator(np.random.PCG64(31))
    n, npres = 12, 5
    cases = []
    # (solve_Stokes, solve_T, size of the change, inside on entry, tol)
    for st, sT, eps, inside, tol in ((True, True, 1e-5, True, 1e-3),
                                     (True, True, 1e-2, True, 1e-3),
                                     (True, True, 1e-2, False, 1e-3),
                                     (True, False, 1e-4, True, 1e-3),
                                     (True, True, 2e-3, True, 1e-3),
                                     (False, True, 1e-4, True, 1e-3)):
        u, w, T = (rng.standard_normal(n) for _ in range(3))
        p = rng.standard_normal(npres)
        um, wm, Tm = (a + eps * rng.standard_normal(n) for a in (u, w, T))
        pm = p + eps * rng.standard_normal(npres)
        buf = io.StringIO()
        with contextlib.redirect_stdout(io.StringIO()):
            out = assess_nlconvergence(3, 2, st, sT, u, w, p, T, um.copy(), wm.copy(), pm.copy(),
                                       Tm.copy(), tol, inside, buf)
        cases.append({"solve_Stokes": st, "solve_T": sT, "inside_in": inside, "tol": tol,
                      "u": u.tolist(), "w": w.tolist(), "p": p.tolist(), "T": T.tolist(),
                      "u_mem": um.tolist(), "w_mem": wm.tolist(), "p_mem": pm.tolist(),
                      "T_mem": Tm.tolist(), "conv_line": buf.getvalue(),
                      "out_u_mem": out[0].tolist(), "out_w_mem": out[1].tolist(),
                      "out_p_mem": out[2].tolist(), "out_T_mem": out[3].tolist(),
                      "inside_out": bool(out[4])})
    fx = {"source": "assess_steady_state.py:10-74 (assess_nlconvergence), istep 3, iter_nl 2",
          "cases": cases}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "nl_vectors.json"),
              "w") as f:
        json.dump(fx, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
