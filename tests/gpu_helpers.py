"""Build libdfx objects from an oracle problem (same arrays on both sides)."""
import numpy as np

import dfx


def enriched_mass_block(P, d):
    """Assembled consistent-mass enriched block M^ee (column-major)."""
    nE, nC = P.nEnriched, P.nClassical
    Mee = np.zeros((nE, nE))
    off = moff = 0
    for n in d["cellNdofs"]:
        ids = d["cellMap"][off:off + n]
        Me = d["M"][moff:moff + n * n].reshape(n, n, order="F")
        sel = ids >= nC
        if sel.any():
            gi = ids[sel] - nC
            Mee[np.ix_(gi, gi)] += Me[np.ix_(np.where(sel)[0], np.where(sel)[0])]
        off += n
        moff += n * n
    return Mee


class Ops:
    def __init__(self, ctx, P, keep_fp32=True):
        d = P.arrays()
        self.d = d
        self.P = P
        self.cons = dfx.Constraints(ctx, P.nDofs, d["rows"], d["rowSizes"], d["rowOffsets"],
                                    d["cols"], d["weights"], d["inhom"])
        mk = lambda key: dfx.CellOperator(ctx, d["cellNdofs"], d["cellMap"], d[key], P.nDofs,
                                          P.nClassical, self.cons, keep_fp32=keep_fp32)
        self.H, self.M, self.K = mk("H"), mk("M"), mk("K")
        nE = P.nEnriched
        enrMinv = d["enrMinv"].reshape(nE, nE, order="F") if nE else None
        self.Minv = dfx.DiagonalOperator(ctx, d["dinv"], nE,
                                         enrMinv.ravel(order="F") if nE else None, self.cons)
        Mee = enriched_mass_block(P, d) if nE else None
        diag = d["mgll"].copy()
        self.Mlump = dfx.DiagonalOperator(ctx, diag, nE,
                                          Mee.ravel(order="F") if nE else None, self.cons)
