"""MI355X-native GMRF hot path of berild/spdepy: ``grid`` and ``model``
factories with the reference's signatures (src/spdepy/__init__.py:13-53,
src/spdepy/grids/__init__.py:3-49)."""
from .grids import grid
from .model import Model
from .spdes import spde_init

__version__ = "0.1.0"


def model(**kwargs):
    """Factory with the reference's defaults (ha=True, bc=3, anisotropic=True)."""
    assert kwargs.get("grid") is not None and "grid" in kwargs.get("grid").type, "Grid is not defined"
    ha = True if type(kwargs.get("ha")) is not bool else kwargs.get("ha")
    bc = 3 if type(kwargs.get("bc")) is not int else kwargs.get("bc")
    ani = True if type(kwargs.get("anisotropic")) is not bool else kwargs.get("anisotropic")
    mesh = kwargs.get("grid")
    if mesh.type == "gridST" and kwargs.get("mod0") is None:
        mesh0 = grid(x=mesh.x, y=mesh.y, extend=mesh.Ne if mesh.Ne else None, Nbs=mesh.Nbs)
        mod0 = spde_init(model="whittle-matern", grid=mesh0, ani=ani, ha=ha, bc=bc)
    elif mesh.type == "gridST":
        m0 = kwargs.get("mod0")
        mod0 = m0.mod if hasattr(m0, "mod") else m0
    else:
        mod0 = None
    return Model(spde=kwargs.get("spde"), grid=mesh, parameters=kwargs.get("parameters"), ani=ani, ha=ha, bc=bc,
                 mod0=mod0)
