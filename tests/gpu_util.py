"""Helpers shared by the GPU tests (no test cases here)."""
import sys

import numpy as np


def step(msg):
    """Name the case before each GPU step so a failed run can be read."""
    print("[gpu-step] " + msg, flush=True)
    sys.stdout.flush()


def product_case(c, kind=None):
    """Build the drop-in model for an oracle fixture c (same grid, same seeds)."""
    import spdepy_amd as sp
    from spdepy_amd import spdes
    g = c["grid"]
    x, y, t = g.x, g.y, g.t
    ne = g.Ne if g.Ne else None
    pg = sp.grid(x=x, y=y, t=t, extend=ne, Nbs=g.Nbs)
    g0 = sp.grid(x=x, y=y, extend=ne, Nbs=g.Nbs)
    mod0 = spdes.WhittleMatern2D(grid=g0, par=np.array(c["mod0"].par), bc=c["model"].bc)
    om = c["model"]
    if kind is None:
        kind = getattr(om, "kind", "var")
    if kind == "var":
        m = spdes.VarAdvectionVarDiffusion2D(grid=pg, mod0=mod0, bc=om.bc)
    else:
        cls = {"iso": spdes.AdvectionIDiffusion2D, "ani": spdes.AdvectionDiffusion2D,
               "ha": spdes.AdvectionHaDiffusion2D}[kind]
        m = cls(grid=pg, mod0=mod0, bc=om.bc)
    return m
