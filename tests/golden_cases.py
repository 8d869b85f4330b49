"""Helpers that load the golden fixtures written from the reference (tests/golden)."""
import os

import numpy as np

from oracle.fvm_oracle import BCSpec, Phys

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def phys_of(d):
    kw = dict(zip(list(d["phys_keys"]), [float(v) for v in d["phys"]]))
    kw["lim_p"] = int(kw["lim_p"])
    u = float(d["u_inf"])
    inlet = BCSpec(v_n_inf=u, v_t_inf=0.0, rho_inf=1.0, T_inf=100.0, v_inf=(u, 0.0, 0.0))
    exit_ = BCSpec(v_n_inf=-u, v_t_inf=0.0, rho_inf=1.0, T_inf=100.0, v_inf=(u, 0.0, 0.0))
    return Phys(**kw, inlet=inlet, exit=exit_)


def mesh_tuple(d):
    return d["points"], d["cells"], d["facets_int"], d["facets_bc"], [str(t) for t in d["tags"]]
