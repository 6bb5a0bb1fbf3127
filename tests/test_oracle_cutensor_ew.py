"""The oracle's cutensor elementwise / reduction family (libraries/cutensor/{ew, ewb,
ewb_castable, reduction}.h), pinned two independent ways:
  * numpy: transposes, label-placed broadcasting, ufuncs and float64 reductions;
  * the reference cpu_ops' own algorithms restated here in Python: permute_t (through the
    oracle's permute, itself pinned against numpy.transpose), the operand scaling, the
    stride_increment_t walk of ewb (ewb.h:55-114, 255-326) and the left fold of reduction
    (reduction.h:137-171), compared bit for bit.
"""
import itertools
import math

import numpy as np
import pytest

DIMS = (2, 3, 4, 5)


def _cm(x, dims):
    """column-major view of a flat buffer"""
    return np.asarray(x).reshape(tuple(dims)[::-1]).transpose()


def _flat(X):
    return np.ascontiguousarray(np.asarray(X).transpose()).reshape(-1)


def _data(n, seed):
    return np.random.default_rng(seed).standard_normal(n).astype(np.float32)


# ---------------------------------------------------------------- ew
@pytest.mark.parametrize("rank", [1, 2, 3, 4])
def test_ew_every_ordering_is_numpy_transpose(orc, rank):
    dims = DIMS[:rank]
    x = _data(int(np.prod(dims)), rank)
    for ordering in itertools.permutations(range(rank)):
        y, od = orc.cu_ew(x, dims, 2, 0.75, ordering)
        Y = np.transpose(_cm(x, dims), ordering)
        assert od == Y.shape
        assert np.array_equal(_flat((Y * Y) * np.float32(0.75)), y), ordering


def test_ew_self_inverse_orderings_equal_the_cpu_op_layout(orc):
    # cpu_op applies f, scales, then permutes in place with permute_t (ew.h:174-177); on
    # orderings that are their own inverse that layout equals the one chosen here
    for rank in (2, 3, 4):
        dims = DIMS[:rank]
        x = _data(int(np.prod(dims)), 7 + rank)
        for ordering in itertools.permutations(range(rank)):
            if any(ordering[ordering[i]] != i for i in range(rank)):
                continue
            y, _ = orc.cu_ew(x, dims, 3, 1.5, ordering)
            ref = orc.permute(np.maximum(x, np.float32(0)) * np.float32(1.5), dims, ordering)
            assert np.array_equal(y, ref), ordering


def test_ew_three_cycle_differs_from_permute_t(orc):
    # documented divergence: for a 3-cycle permute_t lays the data out by the inverse
    # ordering while set_out_meta (and the gpu_op's cuTENSOR call) use the ordering itself
    dims, ordering = (2, 3, 4), (1, 2, 0)
    x = _data(24, 3)
    y, od = orc.cu_ew(x, dims, 2, 1.0, ordering)
    assert od == (3, 4, 2)
    assert not np.array_equal(y, orc.permute(x * x, dims, ordering))


def test_ew_closed_forms(orc):
    x = np.array([0.0, 1.0, -1.0, 2.5, -2.5, 30.0, -30.0, 88.0, 89.0, 100.0, -100.0, -0.0, 1e-30, -1e-30,
                  4.0, 9.0, 0.25, math.log(2.0)], dtype=np.float32)
    n = (x.size,)
    one = lambda f, a=1.0, c=0.0: orc.cu_ew(x, n, f, a, (0,), c)[0]
    sig = one(0)
    assert sig[0] == 0.5 and sig[11] == 0.5
    for i, j in ((1, 2), (3, 4), (5, 6), (12, 13)):
        assert abs(sig[i] + sig[j] - 1.0) <= 2 ** -23
    assert sig[8] == 1.0 and sig[9] == 1.0  # exp overflows FP32: the limit
    assert abs(sig[10] / math.exp(-100.0) - 1.0) < 1e-6
    e = one(1)
    assert e[0] == 1.0 and abs(e[17] - 2.0) <= 2 ** -22
    ref = np.exp(x.astype(np.float64))
    fin = np.isfinite(ref) & (ref < 3e38)
    assert np.all(np.abs(e[fin] - ref[fin]) <= 2.0 ** -22 * ref[fin] + 1e-45)
    assert np.array_equal(one(2), x * x)
    assert np.array_equal(one(3), np.maximum(x, np.float32(0)))
    d = one(4)
    assert d[0] == 1.0 and d[11] == 1.0 and d[13] == 0.0 and d[12] == 1.0 and d[2] == 0.0
    s = one(5)
    assert s[14] == 2.0 and s[15] == 3.0 and s[16] == 0.5 and np.isnan(s[2])
    assert np.array_equal(one(6, 0.5, 1.25), (x + np.float32(1.25)) * np.float32(0.5))


# ---------------------------------------------------------------- ewb
_NP_OPS = {0: np.add, 1: np.fmax, 2: np.fmin, 3: np.multiply, 4: np.subtract, 5: np.divide}


def _ewb_numpy(orc, l, dl, ol, r, dr, orr, op, alpha, beta):
    od = orc.cu_ewb_dims(dl, ol, dr, orr)

    def place(x, d, o):
        X = _cm(x, d)
        if not o:
            return X.reshape((1,) * len(od))
        order = np.argsort(o)
        shape = [1] * len(od)
        for lab, i in zip(sorted(o), order):
            shape[lab] = d[i]
        return np.transpose(X, order).reshape(shape)

    with np.errstate(all="ignore"):
        res = _NP_OPS[op](np.float32(alpha) * place(l, dl, ol), np.float32(beta) * place(r, dr, orr))
    return _flat(np.broadcast_to(res, od)), od


def _ewb_cpu_op(orc, l, dl, ol, r, dr, orr, op, alpha, beta):
    """ewb.h:228-329 restated: permute_t each operand to ascending labels, scale by alpha /
    beta when not 1, then vec_op over output mode 0 for every position of stride_increment_t."""
    od = orc.cu_ewb_dims(dl, ol, dr, orr)

    def prep(x, d, o, a):
        p = orc.permute(x, d, o) if len(o) > 1 else np.asarray(x, np.float32).copy()
        if a != 1.0:
            p = p * np.float32(a)
        return p, sorted(o)

    lp, sl = prep(l, dl, ol, alpha)
    rp, sr = prep(r, dr, orr, beta)
    f = _NP_OPS[op]
    out = np.zeros(int(np.prod(od)), np.float32)
    if len(od) == 0:
        with np.errstate(all="ignore"):
            out[0] = f(lp[0], rp[0])
        return out
    inc_l = 1 if sl and sl[0] == 0 else 0
    inc_r = 1 if sr and sr[0] == 0 else 0
    n0 = od[0]

    def vec(lo, ro, oo):
        with np.errstate(all="ignore"):
            out[oo:oo + n0] = f(lp[lo + inc_l * np.arange(n0)], rp[ro + inc_r * np.arange(n0)])

    if len(od) == 1:
        vec(0, 0, 0)
        return out
    ds = od[1:]

    def stride_ordering(xs):
        ys = xs[1:] if xs[0] == 0 else xs
        return [y - 1 for y in ys]

    def strides(ordering, start):
        st, it, v = [0] * len(ds), 0, start
        for s in range(len(ds)):
            if it < len(ordering) and ordering[it] == s:
                st[s] = v
                v *= ds[s]
                it += 1
        return st

    sl_, sr_ = strides(stride_ordering(sl), 1 if inc_l == 0 else n0), strides(stride_ordering(sr), 1 if inc_r == 0 else n0)
    so, v = [], n0
    for s in range(len(ds)):
        so.append(v)
        v *= ds[s]
    for idx in itertools.product(*[range(d) for d in reversed(ds)]):
        idx = idx[::-1]
        vec(sum(i * s for i, s in zip(idx, sl_)), sum(i * s for i, s in zip(idx, sr_)),
            sum(i * s for i, s in zip(idx, so)))
    return out


EWB_CASES = [
    # (dims_l, ordering_l, dims_r, ordering_r)
    ((4, 5), (0, 1), (4, 5), (0, 1)),
    ((4, 5), (0, 1), (5,), (1,)),
    ((4, 5), (0, 1), (4,), (0,)),
    ((4, 5), (1, 0), (5, 4), (0, 1)),
    ((3, 4, 5), (0, 1, 2), (3, 5), (0, 2)),
    ((3, 4, 5), (2, 0, 1), (4, 5), (0, 1)),
    ((3, 4), (1, 3), (2, 5), (0, 2)),
    ((2, 3, 4, 5), (0, 1, 2, 3), (3, 5), (1, 3)),
    ((2, 3, 4, 5), (3, 1, 0, 2), (4, 3), (0, 1)),
    ((6,), (0,), (6,), (0,)),
    ((6,), (0,), (7,), (1,)),
]


@pytest.mark.parametrize("case", EWB_CASES)
@pytest.mark.parametrize("op", [0, 1, 2, 3, 4, 5])
def test_ewb_against_numpy_and_the_cpu_op(orc, case, op):
    dl, ol, dr, orr = case
    l = _data(int(np.prod(dl)), op + 1)
    r = _data(int(np.prod(dr)), op + 2)
    alpha, beta = (1.0, 1.0) if op % 2 else (1.5, -0.75)
    y, od = orc.cu_ewb(l, dl, ol, r, dr, orr, op, alpha, beta)
    ref, od2 = _ewb_numpy(orc, l, dl, ol, r, dr, orr, op, alpha, beta)
    assert od == od2
    np.testing.assert_array_equal(y, ref)
    np.testing.assert_array_equal(y, _ewb_cpu_op(orc, l, dl, ol, r, dr, orr, op, alpha, beta))


def test_ewb_scalars_and_shape_errors(orc):
    y, od = orc.cu_ewb(np.float32([3.0]), (), (), np.float32([4.0]), (), (), 3, 2.0, 0.5)
    assert od == () and y[0] == 12.0
    with pytest.raises(ValueError):
        orc.cu_ewb_dims((4,), (0,), (5,), (0,))  # mode 0 with two extents
    with pytest.raises(ValueError):
        orc.cu_ewb_dims((4,), (0,), (5,), (2,))  # mode 1 in neither input


# ---------------------------------------------------------------- ewb_castable
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_ewb_castable_is_the_mkl_op(orc, op):
    l = _data(1000, 1)
    r = _data(1000, 2)
    l[3], r[5] = np.nan, np.nan
    y = orc.cu_ewb_castable(op, l, r)
    np.testing.assert_array_equal(y, {0: np.add, 1: np.fmax, 2: np.fmin, 3: np.multiply}[op](l, r))


# ---------------------------------------------------------------- reduction
def _red_cpu_op(orc, x, dims, op, alpha, ordering):
    """reduction.h:137-171 restated: label every aggregated mode 0.. in input order and every
    output mode after them in ordering order, permute_t to that layout, left-fold each output's
    contiguous run with scalarAdd / Max / Min / Mul, then scale by alpha."""
    rank, ro = len(dims), len(ordering)
    labels = [-1] * rank
    for r, m in enumerate(ordering):
        labels[m] = rank - ro + r
    j = 0
    for i in range(rank):
        if labels[i] == -1:
            labels[i], j = j, j + 1
    xp = orc.permute(x, dims, labels) if rank > 1 else np.asarray(x, np.float32)
    num_out = int(np.prod([dims[m] for m in ordering])) if ordering else 1
    F = xp.reshape(num_out, -1)
    acc = F[:, 0].copy()
    for k in range(1, F.shape[1]):
        v = F[:, k]
        if op == 0:
            acc = acc + v
        elif op == 1:
            acc = np.where(acc > v, acc, v)
        elif op == 2:
            acc = np.where(acc < v, acc, v)
        else:
            acc = acc * v
    return acc * np.float32(alpha)


def _orderings(rank):
    for k in range(rank + 1):
        for sub in itertools.permutations(range(rank), k):
            yield sub


@pytest.mark.parametrize("rank", [1, 2, 3, 4])
@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_reduction_is_the_cpu_op_left_fold(orc, rank, op):
    dims = DIMS[:rank]
    x = _data(int(np.prod(dims)), 10 * rank + op)
    if op == 3:
        x = np.float32(1.0) + np.float32(0.05) * x  # keep products in range
    for ordering in _orderings(rank):
        y, od = orc.cu_reduction(x, dims, op, 1.25, ordering)
        assert od == tuple(dims[m] for m in ordering)
        np.testing.assert_array_equal(y, _red_cpu_op(orc, x, dims, op, 1.25, ordering), str(ordering))


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_reduction_against_numpy(orc, op):
    dims = (37, 20, 9, 3)
    x = _data(int(np.prod(dims)), 99)
    if op == 3:
        x = np.float32(1.0) + np.float32(0.01) * x
    X = _cm(x, dims).astype(np.float64)
    f = {0: np.sum, 1: np.max, 2: np.min, 3: np.prod}[op]
    for ordering in [(0,), (1,), (3, 0), (2, 1), (0, 2, 3), (1, 3), ()]:
        agg = tuple(i for i in range(4) if i not in ordering)
        ref = np.transpose(f(X, axis=agg, keepdims=True), list(ordering) + list(agg)).reshape(
            [dims[m] for m in ordering]) * 0.5
        y, _ = orc.cu_reduction(x, dims, op, 0.5, ordering)
        ref = _flat(ref)
        if op in (1, 2):
            assert np.array_equal(y, ref.astype(np.float32))
        else:
            assert np.allclose(y, ref, rtol=3e-6 * np.prod([dims[a] for a in agg]) ** 0.5, atol=1e-4)
