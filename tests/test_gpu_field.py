"""Device field arithmetic (Fq, Fr, Fq2 on gfx950) vs the CPU oracle, bit-exact.

Edge values 0, 1, p-1, p-2, (p-1)/2, 2^253 and R mod p sit next to seeded
random values; every op of include/hzk.h's hzk_field_op is compared byte for
byte with oracle/orc_core.cpp."""
import numpy as np
import pytest

import pymodel as pm

pytestmark = pytest.mark.gpu

OPS = {"mul": 0, "add": 1, "sub": 2, "neg": 3, "sqr": 4, "inv": 5, "from_mont": 6, "to_mont": 7}
ORACLE_OP = {"mul": 0, "add": 1, "sub": 2, "neg": 3, "sqr": 4, "inv": 5, "from_mont": 6, "to_mont": 7}


def _inputs(orc, kind, n):
    m = pm.P if kind == "fq" else pm.R
    edge = [0, 1, 2, m - 1, m - 2, (m - 1) // 2, 1 << 253, pm.MONT % m, (pm.MONT * pm.MONT) % m]
    rnd = pm.elems(orc.gen_fq(7, 0, n) if kind == "fq" else orc.gen_fr(7, 0, n))
    a = edge + rnd + edge
    b = list(reversed(a))
    enc = lambda v: np.concatenate([pm.int_to_limbs(x) for x in v])
    return enc(a), enc(b)


@pytest.mark.parametrize("kind", ["fq", "fr"])
@pytest.mark.parametrize("op", list(OPS))
def test_prime_field_ops(ctx, orc, kind, op):
    a, b = _inputs(orc, kind, 4000)
    n = a.size // 4
    da, db, dout = ctx.upload(a), ctx.upload(b), ctx.alloc(a.nbytes)
    ctx.field_op(kind, op, da, db, dout, n)
    got = dout.download(np.uint64, a.size)
    want = orc.field_op(kind, ORACLE_OP[op], a, b)
    bad = np.nonzero((got.reshape(-1, 4) != want.reshape(-1, 4)).any(axis=1))[0]
    assert bad.size == 0, f"{kind}.{op}: {bad.size} mismatches, first at {bad[:5]}"


@pytest.mark.parametrize("op", ["mul", "add", "sub", "neg", "sqr", "inv"])
def test_fq2_ops(ctx, orc, op):
    n = 2000
    a = orc.gen_fq(8, 0, 2 * n)
    b = orc.gen_fq(8, 1, 2 * n)
    a[:8] = 0   # zero element
    b[8:16] = a[8:16]
    da, db, dout = ctx.upload(a), ctx.upload(b), ctx.alloc(a.nbytes)
    ctx.field_op("fq2", op, da, db, dout, n)
    got = dout.download(np.uint64, a.size)
    want = orc.field_op("fq2", OPS[op], a, b)
    assert np.array_equal(got, want), op
