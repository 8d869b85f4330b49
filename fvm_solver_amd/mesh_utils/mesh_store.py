"""Boundary-facet descriptors -- mirror of ``time_fvm/mesh_utils/mesh_store.py:5-59``."""
from dataclasses import dataclass
from enum import Flag, auto


class FacetBCTypes(Flag):
    Dirich = auto()
    Neuman = auto()
    Both = Dirich | Neuman
    Farfield = auto()
    Inlet = auto()


@dataclass
class Facet:
    """Per-component BC types of one boundary facet with Dirichlet values / Neumann
    gradients; unset entries become NaN (mesh_store.py:24-41)."""
    edge_type: list
    U: list = None
    dUdn: list = None
    euler_wall: bool = False
    tag: str = None

    def __post_init__(self):
        if self.U is None:
            self.U = [None] * len(self.edge_type)
        if self.dUdn is None:
            self.dUdn = [None] * len(self.edge_type)
        for e, u, dudn in zip(self.edge_type, self.U, self.dUdn, strict=True):
            if FacetBCTypes.Dirich in e:
                assert u is not None, "Dirichlet BC requires a value."
                assert dudn is None, "Dirichlet BC does not require a gradient."
            if FacetBCTypes.Neuman in e:
                assert dudn is not None, "Neumann BC requires a gradient."
                assert u is None, "Neumann BC does not require a value."
        self.U = [float("nan") if u is None else u for u in self.U]
        self.dUdn = [float("nan") if d is None else d for d in self.dUdn]

    def dirichlet(self):
        return [FacetBCTypes.Dirich in e for e in self.edge_type]

    def neumann(self):
        return [FacetBCTypes.Neuman in e for e in self.edge_type]

    def farfield(self) -> bool:
        ff = [FacetBCTypes.Farfield in e for e in self.edge_type]
        assert all(ff) or not any(ff), f"All boundary must be farfield if any are: {ff}"
        return all(ff)

    def inlet(self) -> bool:
        il = [FacetBCTypes.Inlet in e for e in self.edge_type]
        assert all(il) or not any(il), f"All boundary must be inlet if any are: {il}"
        return all(il)


def channel_bc_tags(tags, dim):
    """Boundary tag -> Facet mapping of the reference drivers (run_fvm.py:264-274,
    run_fvm_3d.py:231-239): NavierWall = no-slip (Dirichlet V = 0, Neumann rho, T),
    Left = inlet, Right = farfield."""
    E = FacetBCTypes
    out = {}
    for i, t in enumerate(tags):
        if t == "NavierWall":
            out[i] = Facet([E.Dirich] * dim + [E.Neuman, E.Neuman], [0.0] * dim + [None, None],
                           [None] * dim + [0, 0], tag=t)
        elif t == "Left":
            out[i] = Facet([E.Inlet] * (dim + 2), tag=t)
        elif t == "Right":
            out[i] = Facet([E.Farfield] * (dim + 2), tag=t)
        else:
            raise ValueError(f"Unknown edge tag {t}")
    return out
