"""Writes tests/golden/stencils.npz: inputs and outputs of the reference's own
stencil code (berild/spdepy ``src/spdepy/spdes/ccode/*.cpp``), compiled from its
sources by ``oracle/Makefile`` into ``oracle/_ref``.  Run in the CPU container:

    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ref_native as R  # noqa: E402


def cases():
    rng = np.random.default_rng(20240611)
    out = {}
    hx, hy = 0.7, 1.3
    for (M, N) in [(7, 5), (6, 6), (4, 3)]:
        for bc in (1, 2, 3):
            tag = "%dx%d_b%d" % (M, N, bc)
            Hv = rng.normal(size=(M * N, 4, 2, 2))
            Hv = Hv + Hv.transpose(0, 1, 3, 2)
            out["AHv_" + tag] = (Hv, R.ah(M, N, Hv, hx, hy, bc, True))
            iso = np.zeros((M * N, 4, 2, 2))
            iso[:, :, 0, 0] = iso[:, :, 1, 1] = np.exp(rng.normal(size=(M * N, 4)))
            out["AHiso_" + tag] = (iso, R.ah(M, N, iso, hx, hy, bc, True))
            if bc != 2:   # the reference's bc=2 constant stencil has an ABI mismatch
                Hc = rng.normal(size=(2, 2))
                Hc = Hc + Hc.T
                out["AHc_" + tag] = (Hc, R.ah(M, N, Hc, hx, hy, bc, False))
                out["AHciso_" + tag] = (0.5 * np.eye(2), R.ah(M, N, 0.5 * np.eye(2), hx, hy, bc, False))
            for diff in (1, 2, 3):
                G = rng.normal(size=2)
                out["Awc_%s_d%d" % (tag, diff)] = (G, R.aw(M, N, G, hx, hy, diff, bc, False))
                G0 = np.array([0.0, 0.4])   # G = 0 gives NaN in the derivative stencils
                out["Awc0_%s_d%d" % (tag, diff)] = (G0, R.aw(M, N, G0, hx, hy, diff, bc, False))
                Gv = rng.normal(size=(M * N, 4))
                dG = rng.normal(size=(M * N, 4))
                Gv[3, 1] = 0.0
                out["Awv_%s_d%d" % (tag, diff)] = ((Gv, dG), R.aw(M, N, Gv, hx, hy, diff, bc, True, dG))
    return out, hx, hy


def main():
    assert R.available(), "build oracle/_ref first: make -C oracle"
    out, hx, hy = cases()
    flat = {"hx": hx, "hy": hy}
    for name, (inp, (row, col, val)) in out.items():
        if isinstance(inp, tuple):
            flat[name + "/G"], flat[name + "/dG"] = inp
        else:
            flat[name + "/in"] = inp
        flat[name + "/row"], flat[name + "/col"], flat[name + "/val"] = row, col, val
    np.savez_compressed(os.path.join(HERE, "stencils.npz"), **flat)
    print("wrote", len(out), "cases")


if __name__ == "__main__":
    main()
