"""Loading oracle-built graphs into the HIP index (shared by the GPU tests)."""
import numpy as np

from pickle_amd import abi


def index_from_oracle(g, capacity=None, slack=True):
    """An abi.Index holding exactly the oracle graph g (its layers, the slack
    row written past a full last row, and the per-layer vector flags)."""
    idx = abi.Index(g.base.dtype, g.metric, g.dim, g.M, g.efc, capacity or g.base.shape[0])
    idx.set_base(g.base)
    for l in range(g.nlayers):
        L = g.layer(l)
        cap = L["cap"]
        adj = L["adj"] if slack else L["adj"][:cap]
        idx.load_layer(l, L["ext"], L["deg"].astype(np.int32), adj, L["has_data"], heur=L["heur"],
                       dist=L["dist"][:cap])
    idx.set_enter_level(g.enter_level)
    return idx
