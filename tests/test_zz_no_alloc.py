"""The allocation-free apply (dfx_op_apply_noalloc) against dfx_op_apply on
the same inputs: bit for bit, two calls in a row with the second smaller
(the per-GPU scratch keeps the first call's contents), FP64 and FP32,
enriched (scratch used) and classical operators; applies whose enriched
partial rows exceed the scratch are refused."""
import numpy as np
import pytest

import dfx
from gpu_helpers import Ops
from problems import c1_like, enriched, guess

pytestmark = pytest.mark.gpu


def mv(ctx, a, dtype=dfx.F64):
    return dfx.MultiVector.from_numpy(ctx, a, dtype)


@pytest.mark.parametrize("name", ["enr", "c1"])
@pytest.mark.parametrize("dtype", [dfx.F64, dfx.F32])
def test_noalloc_apply_equals_apply(ctx, name, dtype):
    P = enriched(p=2, n=4) if name == "enr" else c1_like(n=3)
    ops = Ops(ctx, P)
    npd = np.float64 if dtype == dfx.F64 else np.float32
    for B, seed in ((24, 3), (8, 4)):           # the second call smaller
        X = guess(P, B, seed=seed).astype(npd)
        for op in (ops.H, ops.M, ops.Minv):
            Xa, Xb = mv(ctx, X, dtype), mv(ctx, X, dtype)
            Ya, Yb = (dfx.MultiVector(ctx, P.nDofs, B, dtype) for _ in range(2))
            op.apply(Xa, Ya)
            op.apply_noalloc(Xb, Yb)
            assert np.array_equal(Ya.to_numpy(), Yb.to_numpy())
            assert np.array_equal(Xa.to_numpy(), Xb.to_numpy())   # the same P2C on X


def test_noalloc_apply_refuses_applies_above_the_scratch(ctx):
    P = enriched(p=2, n=4)
    ops = Ops(ctx, P)
    nslots = int((P.arrays()["cellMap"] >= P.nClassical).sum())
    assert nslots > 0
    B = dfx.NOALLOC_SCRATCH_BYTES // (8 * nslots) + 1
    X = dfx.MultiVector(ctx, P.nDofs, B)
    X.set(0.0)
    Y = dfx.MultiVector(ctx, P.nDofs, B)
    with pytest.raises(dfx.DfxError, match="scratch"):
        ops.H.apply_noalloc(X, Y)
    # a classical operator uses no scratch: any width is accepted
    Q = c1_like(n=2)
    q = Ops(ctx, Q)
    Xq = dfx.MultiVector(ctx, Q.nDofs, 2)
    Xq.set(1.0)
    Yq, Yr = dfx.MultiVector(ctx, Q.nDofs, 2), dfx.MultiVector(ctx, Q.nDofs, 2)
    q.H.apply_noalloc(Xq, Yq)
    q.H.apply(Xq, Yr)
    assert np.array_equal(Yq.to_numpy(), Yr.to_numpy())
