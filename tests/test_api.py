"""Host-side API of the drop-in (no GPU): the model dispatcher and the
initial parameter vectors of each family class, against the reference's
constructors (src/spdepy/spdes/__init__.py:1-103 and each class's default
``par``, e.g. var_advection_idiffusion2D.py:22)."""
import numpy as np
import pytest

import spdepy_amd as sp
from spdepy_amd import spdes

LOG100 = np.log(100.0)
X = (np.arange(8) + 0.5) / 8
Y = (np.arange(7) + 0.5) / 7

# (spde, anisotropic, ha) -> (class name, default own parameters incl. log sigma and log tau)
FAMILIES = {
    ("advection-diffusion", False, False): ("AdvectionIDiffusion2D", None),
    ("advection-var-diffusion", False, False):
        ("AdvectionVarIDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, 0.1, 0.1, 0, LOG100])),
    ("advection-var-diffusion", True, False):
        ("AdvectionVarDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 9, [0.1] * 9, 0.1, 0.1, 0, LOG100])),
    ("advection-var-diffusion", True, True):
        ("AdvectionVarHaDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 9, [0.1] * 9, 0.1, 0.1, 0, LOG100])),
    ("var-advection-diffusion", False, False):
        ("VarAdvectionIDiffusion2D", np.hstack([-1, -1, [0.01] * 18, 0, LOG100])),
    ("var-advection-diffusion", True, False):
        ("VarAdvectionDiffusion2D", np.hstack([-1, -1, 0.1, 0.1, [0.1] * 18, 0, LOG100])),
    ("var-advection-diffusion", True, True):
        ("VarAdvectionHaDiffusion2D", np.hstack([-1, -1, 0.1, 0.1, [0.1] * 18, 0, LOG100])),
    ("var-advection-var-diffusion", False, False):
        ("VarAdvectionVarIDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 18, 0, LOG100])),
    ("var-advection-var-diffusion", True, False):
        ("VarAdvectionVarDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 18, [0.1] * 18, 0, LOG100])),
    ("var-advection-var-diffusion", True, True):
        ("VarAdvectionVarHaDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 18, [0.1] * 18, 0, LOG100])),
    ("cov-advection-diffusion", False, False): ("CovAdvectionIDiffusion2D", np.array([-1, -1, 0.1, 0, LOG100])),
    ("cov-advection-diffusion", True, False):
        ("CovAdvectionDiffusion2D", np.array([-1, -1, 0.1, 0.1, 0.1, 0, LOG100])),
    ("cov-advection-diffusion", True, True):
        ("CovAdvectionHaDiffusion2D", np.array([-1, -1, 0.01, 0.01, 0.01, 0, LOG100])),
    ("cov-advection-var-diffusion", False, False):
        ("CovAdvectionVarIDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, 0.1, 0, LOG100])),
    ("cov-advection-var-diffusion", True, False):
        ("CovAdvectionVarDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 9, [0.1] * 9, 0.1, 0, LOG100])),
    ("cov-advection-var-diffusion", True, True):
        ("CovAdvectionVarHaDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 9, [0.1] * 9, 0.1, 0, LOG100])),
    ("seperable-spatial-temporal", False, False):
        ("SeperableSpatialTemporalIDiffusion2D", np.hstack([[-1] * 9, [-1] * 9, 0.1, 1, LOG100])),
    ("seperable-spatial-temporal", True, False):
        ("SeperableSpatialTemporal2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 9, [0.1] * 9, -1, LOG100])),
    ("seperable-spatial-temporal", True, True):
        ("SeperableSpatialTemporalHa2D", np.hstack([[-1] * 9, [-1] * 9, [0.1] * 9, [0.1] * 9, 0.1, 1, LOG100])),
}


@pytest.mark.parametrize("key", sorted(FAMILIES))
def test_dispatch_and_default_parameters(key):
    spde, ani, ha = key
    name, default = FAMILIES[key]
    g = sp.grid(x=X, y=Y, t=np.arange(4.0), Nbs=3)
    g0 = sp.grid(x=X, y=Y, Nbs=3)
    mod0 = spdes.spde_init("whittle-matern", g0, ani=False, ha=False, bc=3)
    m = spdes.spde_init(spde, g, ani=ani, ha=ha, bc=3, mod0=mod0)
    assert type(m).__name__ == name
    if default is not None and spde.startswith("seperable"):
        assert np.array_equal(m.getPars(), default)
        assert np.array_equal(m.initFit(np.zeros(10), idx=np.arange(10)), default)
    elif default is not None:
        assert np.array_equal(m.getPars(), default)
        # initFit appends the initial-state parameters before log tau
        data = np.zeros(10)
        kw = dict(ww=np.zeros((g.Ns, 4))) if spde.startswith("cov") else {}
        par = m.initFit(data, idx=np.arange(10), **kw)
        assert par.size == default.size + mod0.getPars().size - 1
        assert np.array_equal(par[:default.size - 1], default[:-1]) and par[-1] == LOG100


def test_numeric_model_codes():
    g = sp.grid(x=X, y=Y, t=np.arange(4.0), Nbs=3)
    mod0 = spdes.spde_init("whittle-matern", sp.grid(x=X, y=Y, Nbs=3), ani=False, ha=False, bc=3)
    for code, name in ((3, "AdvectionVarHaDiffusion2D"), (6, "VarAdvectionHaDiffusion2D"),
                       (7, "VarAdvectionVarHaDiffusion2D")):
        assert type(spdes.spde_init(code, g, mod0=mod0)).__name__ == name


def test_reference_abi_stencil_libraries_export_the_reference_names():
    """One library per stencil kind and boundary condition, each exporting
    the reference's C names (AH_new/AH_Row/AH_Col/AH_Val/AH_delete or the
    Aw_ set), loadable without a GPU."""
    import ctypes
    import os
    d = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spdepy_amd", "ccode")
    for kind, pre in (("AcH", "AH"), ("AH", "AH"), ("Acw", "Aw"), ("Aw", "Aw")):
        for bc in (1, 2, 3):
            lib = ctypes.CDLL(os.path.join(d, "lib_%s_2D_b%d.so" % (kind, bc)))
            for nm in ("_new", "_Row", "_Col", "_Val", "_delete"):
                assert hasattr(lib, pre + nm), (kind, bc, nm)
