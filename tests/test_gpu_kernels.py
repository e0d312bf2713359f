"""GPU parity of the C-ABI kernels against the oracle (run with -m gpu).

Bars: permutes are bit-exact against the oracle's restatement of the
reference loops; the FP64 GEMM is within 1e-12 relative Frobenius error of
numpy's dgemm; FP32-MFMA and split-BF16 modes carry their own stated
tolerances; the SVD/eigh match SciPy's dgesvd/dsyev (the routines the
reference calls) to 1e-10 and up to the sign of each singular pair.
"""
import ctypes
import itertools
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rand(seed, shape):
    return np.random.default_rng(seed).uniform(-1, 1, size=shape)


# ---------------------------------------------------------------------------
# dense permute
DENSE_CASES = [
    ((256, 2, 2, 256), (3, 1, 0, 2)),   # contraction-and-permute workload, output permute
    ((256, 2, 2, 256), (1, 0, 2, 3)),   # its internal B permute (row copy)
    ((7,), (0,)),
    ((5, 3), (1, 0)),
    ((64, 64), (1, 0)),
    ((65, 130), (1, 0)),
    ((1, 1, 1), (2, 0, 1)),
    ((3, 1, 4, 1, 5), (4, 2, 0, 3, 1)),
    ((2, 2, 2, 2, 2, 2), (5, 4, 3, 2, 1, 0)),
    ((33, 17, 9), (2, 0, 1)),
    ((33, 17, 9), (1, 2, 0)),
    ((100, 3, 70), (2, 1, 0)),
    ((4, 300, 5), (0, 2, 1)),
    ((8, 8, 8, 8), (0, 1, 3, 2)),
    ((2, 1000, 3), (1, 0, 2)),
    ((129, 2, 3, 2, 67), (4, 1, 3, 0, 2)),
]


@pytest.mark.parametrize("dims,perm", DENSE_CASES)
def test_permute_dense_bitexact(tn, ora, dims, perm):
    _say("permute_dense dims=%s perm=%s" % (dims, perm))
    n = int(np.prod(dims))
    x = _rand(1, n)
    ref = np.empty(n)
    d64 = np.asarray(dims, np.int64)
    p32 = np.asarray(perm, np.int32)
    ora.lib().ora_permute_dense(ora._ptr(x), ora._ptr(ref), len(dims), ora._ptr(d64), ora._ptr(p32))
    assert np.array_equal(ref, np.transpose(x.reshape(dims), perm).reshape(-1))
    src = tn.DeviceArray.from_host(x)
    dst = tn.DeviceArray((n,))
    tn.permute_dense(src, dst, dims, perm)
    out = dst.to_host()
    assert out.tobytes() == ref.tobytes()


def test_permute_dense_random_sweep(tn, ora):
    rng = np.random.default_rng(7)
    for case in range(60):
        r = int(rng.integers(1, 7))
        dims = [int(v) for v in rng.integers(1, 9, size=r)]
        if case % 5 == 0:
            dims[int(rng.integers(0, r))] = int(rng.integers(60, 200))
        perm = [int(v) for v in rng.permutation(r)]
        n = int(np.prod(dims))
        _say("permute sweep %d dims=%s perm=%s" % (case, dims, perm))
        x = rng.standard_normal(n)
        src = tn.DeviceArray.from_host(x)
        dst = tn.DeviceArray((n,))
        tn.permute_dense(src, dst, dims, perm)
        ref = np.transpose(x.reshape(dims), perm).reshape(-1)
        assert dst.to_host().tobytes() == ref.tobytes()


def test_permute_dense_hbm_size_roundtrip(tn):
    """Full-size (4096,2,2,4096) FP64 permute (512 MiB each way): checked by
    applying the inverse permutation on the device and by a host transpose."""
    dims = (4096, 2, 2, 4096)
    perm = (3, 1, 0, 2)
    inv = tuple(int(v) for v in np.argsort(perm))
    n = int(np.prod(dims))
    _say("hbm-size permute n=%d" % n)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    src = tn.DeviceArray.from_host(x)
    mid = tn.DeviceArray((n,))
    back = tn.DeviceArray((n,))
    tn.permute_dense(src, mid, dims, perm)
    out_dims = tuple(dims[p] for p in perm)
    tn.permute_dense(mid, back, out_dims, inv)
    assert back.to_host().tobytes() == x.tobytes()
    ref = np.ascontiguousarray(np.transpose(x.reshape(dims), perm)).reshape(-1)
    assert mid.to_host().tobytes() == ref.tobytes()


def test_reshape_elem_reference_table(tn, ora):
    """tn_reshape_elem_f64 with the offset table the reference builds
    (UniTensorReal.cpp:586-595), against the oracle's _reshapeElem formula."""
    for dims, perm in DENSE_CASES[:8]:
        r = len(dims)
        n = int(np.prod(dims))
        new_dims = [dims[p] for p in perm]
        new_acc = [1] * r
        src_acc = [1] * r
        for b in range(r - 1, 0, -1):
            new_acc[b - 1] = new_acc[b] * new_dims[b]
            src_acc[b - 1] = src_acc[b] * dims[b]
        off = src_acc + [0] * r
        for b in range(r):
            off[r + perm[b]] = new_acc[b]
        x = _rand(11, n)
        ref = np.empty(n)
        ora.lib().ora_reshape_elem(ora._ptr(x), ora._ptr(ref), r, ora._ptr(np.asarray(dims, np.int64)),
                                   ora._ptr(np.asarray(perm, np.int32)))
        src = tn.DeviceArray.from_host(x)
        dst = tn.DeviceArray((n,))
        tn.reshape_elem(src, dst, off)
        assert dst.to_host().tobytes() == ref.tobytes()


def test_transpose(tn):
    for M, N in [(1, 1), (3, 7), (64, 65), (300, 17)]:
        a = _rand(M * 31 + N, (M, N))
        A = tn.DeviceArray.from_host(a)
        AT = tn.DeviceArray((N, M))
        tn.call("tn_transpose_f64", tn._p(A), M, N, tn._p(AT), None)
        assert AT.to_host().tobytes() == np.ascontiguousarray(a.T).tobytes()


# ---------------------------------------------------------------------------
# symmetric block permute
def _u1_bond(ora, tp, charges_degs):
    qs = []
    for c, d in charges_degs:
        qs += [ora.Qnum(c)] * d
    return ora.Bond(tp, qs)


def _u1_tensor(ora, seed):
    rng = np.random.default_rng(seed)
    phys = _u1_bond(ora, ora.BD_IN, [(1, 1), (-1, 1)])
    chi_l = _u1_bond(ora, ora.BD_IN, [(c, int(rng.integers(1, 5))) for c in (-3, -1, 1, 3)])
    chi_r = _u1_bond(ora, ora.BD_OUT, [(c, int(rng.integers(1, 5))) for c in (-3, -1, 1, 3)])
    phys_o = _u1_bond(ora, ora.BD_OUT, [(1, 1), (-1, 1)])
    T = ora.UniTensor([chi_l, phys, chi_r, phys_o])
    T.set_raw_elem(rng.uniform(-1, 1, size=[b.dim for b in T.bonds]))
    return T


@pytest.mark.parametrize("perm,rb", [((2, 0, 3, 1), 1), ((3, 2, 1, 0), 2), ((0, 1, 2, 3), 3),
                                     ((1, 3, 0, 2), 2), ((0, 2, 1, 3), 0), ((2, 3, 0, 1), 4)])
def test_block_permute_bitexact(tn, ora, perm, rb):
    T = _u1_tensor(ora, 5)
    new_labels = [T.labels[p] for p in perm]
    _say("block permute perm=%s rb=%d blocks=%d" % (perm, rb, len(T.layout.blocks)))
    out = T.copy().permute(new_labels, rb)   # oracle result
    src_T = T.copy()
    # moves as the oracle computes them (the product derives the same from its own layout)
    rsp = list(perm)
    tgt = ora.UniTensor([src_T.bonds[p].copy().change(ora.BD_IN if i < rb else ora.BD_OUT)
                         for i, p in enumerate(perm)])
    moves = ora.permute_moves(src_T, tgt, rsp, new_labels)
    plan = tn.BlockPermPlan(4, moves[0], moves[1], moves[2], moves[3], tgt.elem_num())
    src = tn.DeviceArray.from_host(src_T.elem)
    dst = tn.DeviceArray((tgt.elem_num(),))
    plan(src, dst)
    got = dst.to_host()
    assert got.tobytes() == out.elem.tobytes()
    # densify invariant
    dense = np.transpose(T.get_raw_elem().reshape([b.dim for b in T.bonds]), perm)
    tgt.set_elem(got)
    assert np.array_equal(tgt.get_raw_elem().reshape(dense.shape), dense)


# ---------------------------------------------------------------------------
# grouped GEMM
GEMM_TOL = {0: 1e-12, 1: 2e-6, 2: 5e-5, 3: 2e-6, 4: 1e-11}   # relative Frobenius vs dgemm


def _group_case(seed, shapes):
    rng = np.random.default_rng(seed)
    descs = []
    a_parts, b_parts = [], []
    ao = bo = co = 0
    for (M, N, K) in shapes:
        a_parts.append(rng.uniform(-1, 1, M * K))
        b_parts.append(rng.uniform(-1, 1, K * N))
        descs.append((M, N, K, ao, bo, co, max(K, 1), max(N, 1), max(N, 1)))
        ao += M * K
        bo += K * N
        co += M * N
    A = np.concatenate(a_parts) if a_parts else np.zeros(0)
    B = np.concatenate(b_parts) if b_parts else np.zeros(0)
    return descs, A, B, co


def _check_gemm(tn, mode, shapes, seed=0, tile=None):
    descs, A, B, nc = _group_case(seed, shapes)
    plan = tn.GemmPlan(mode, np.array(descs, dtype=tn.GEMM_DESC))
    if tile is not None:
        plan.set_tile(tile)
    dA = tn.DeviceArray.from_host(A if A.size else np.zeros(1))
    dB = tn.DeviceArray.from_host(B if B.size else np.zeros(1))
    dC = tn.DeviceArray((max(nc, 1),))
    plan(dA, dB, dC)
    C = dC.to_host()
    for (M, N, K, ao, bo, co, *_r) in descs:
        ref = A[ao:ao + M * K].reshape(M, K) @ B[bo:bo + K * N].reshape(K, N)
        got = C[co:co + M * N].reshape(M, N)
        den = max(np.linalg.norm(ref), 1e-300)
        err = np.linalg.norm(got - ref) / den if M * N else 0.0
        assert err <= GEMM_TOL[mode], (mode, M, N, K, err)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_gemm_row_scaled_second_output(tn, mode):
    """The row-scaled second output (d_off, rs_off, rs_div) on every tile
    shape: C is bit-identical to the same plan without it, and D(m, n) is
    exactly C(m, n) * S[rs_off + m // rs_div] (the product numpy forms of
    the stored C), also through C index maps and for a K = 0 group, whose C
    and D are zeros; the bits of a separate row-scaling pass."""
    _say("gemm row-scaled second output mode=%d" % mode)
    shapes = [(40, 24, 17), (7, 5, 0), (66, 33, 64), (1, 1, 1), (96, 64, 48)]
    descs, A, B, nc = _group_case(71 + mode, shapes)
    rng = np.random.default_rng(9)
    ex = np.zeros(len(descs), dtype=tn.GEMM_DESC_EX)
    ns, nd, divs = 0, 0, []
    for g, (M, N, K, ao, bo, co, lda, ldb, ldc) in enumerate(descs):
        ex[g]["M"], ex[g]["N"], ex[g]["K"] = M, N, K
        ex[g]["a_off"], ex[g]["b_off"], ex[g]["c_off"] = ao, bo, co
        ex[g]["lda"], ex[g]["ldb"], ex[g]["ldc"] = lda, ldb, ldc
        if g == 2:   # C written transposed through index maps
            ex[g]["c_row"] = tn.index_map([(M, 1)])
            ex[g]["c_col"] = tn.index_map([(N, M)])
        div = [3, 1, 11, 1, 8][g]
        divs.append(div)
    # the buffer handed as C holds C, then D, then the scales S
    S = rng.uniform(0.1, 2.0, sum((M + d - 1) // d for (M, *_r), d in zip(descs, divs)))
    so = 0
    for g, ((M, *_r), div) in enumerate(zip(descs, divs)):
        ex[g]["d_off"] = nc + ex[g]["c_off"]
        ex[g]["rs_off"] = 2 * nc + so
        ex[g]["rs_div"] = div
        so += (M + div - 1) // div
    plain = ex.copy()
    plain["rs_div"] = 0
    dA, dB = tn.DeviceArray.from_host(A), tn.DeviceArray.from_host(B)
    ref = tn.DeviceArray((nc,))
    tn.GemmPlan(mode, plain)(dA, dB, ref)
    Cref = ref.to_host()
    for tile in range(NTILES[mode]):
        plan = tn.GemmPlan(mode, ex)
        plan.set_tile(tile)
        dC = tn.DeviceArray.from_host(np.concatenate([np.full(2 * nc, np.nan), S]))
        plan(dA, dB, dC)
        out = dC.to_host()
        C, D = out[:nc], out[nc:2 * nc]
        assert C.tobytes() == Cref.tobytes(), (mode, tile)
        for g, (M, N, K, ao, bo, co, *_r) in enumerate(descs):
            if g == 2:
                Cg, Dg = C[co:co + M * N].reshape(N, M).T, D[co:co + M * N].reshape(N, M).T
            else:
                Cg, Dg = C[co:co + M * N].reshape(M, N), D[co:co + M * N].reshape(M, N)
            scale = S[int(ex[g]["rs_off"]) - 2 * nc + np.arange(M) // divs[g]]
            assert (Dg * 1.0).tobytes() == (Cg * scale[:, None]).tobytes(), (mode, tile, g)
            if K == 0:
                assert not np.any(Cg) and not np.any(Dg)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_gemm_single_512(tn, mode):
    _say("gemm 512^3 mode=%d" % mode)
    _check_gemm(tn, mode, [(512, 512, 512)])


NTILES = {0: 9, 1: 9, 2: 5, 3: 5, 4: 4}


@pytest.mark.parametrize("mode,tile", [(m, t) for m in (0, 1, 2, 3, 4) for t in range(NTILES[m])])
def test_gemm_grouped_uneven(tn, mode, tile):
    shapes = [(1, 1, 1), (17, 33, 5), (64, 64, 64), (130, 70, 200), (3, 200, 1), (200, 3, 300),
              (0, 5, 5), (5, 0, 5), (5, 5, 0), (96, 128, 17)]
    _say("gemm grouped mode=%d tile=%d" % (mode, tile))
    _check_gemm(tn, mode, shapes, seed=mode * 10 + tile, tile=tile)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_gemm_tile_choice_is_bitwise_neutral(tn, mode):
    """Every tile shape (and so the autotuner's choice) gives identical bits:
    each output element accumulates K in the same order."""
    shapes = [(200, 130, 300), (17, 33, 5), (64, 64, 640)]
    descs, A, B, nc = _group_case(77, shapes)
    plan = tn.GemmPlan(mode, np.array(descs, dtype=tn.GEMM_DESC))
    dA, dB = tn.DeviceArray.from_host(A), tn.DeviceArray.from_host(B)
    outs = []
    for tile in range(NTILES[mode]):   # FP64/FP32: three 512-thread tiles and a two-stage 32x32
        plan.set_tile(tile)
        dC = tn.DeviceArray((nc,))
        plan(dA, dB, dC)
        outs.append(dC.to_host().tobytes())
    assert all(o == outs[0] for o in outs)
    dC = tn.DeviceArray((nc,))
    ms = plan.autotune(dA, dB, dC, reps=2)
    assert ms > 0 and dC.to_host().tobytes() == outs[0]


@pytest.mark.parametrize("mode", [0, 1])
def test_gemm_tile_choice_is_bitwise_neutral_paired_loads(tn, mode):
    """The same with every K and N even, so the tiles read A and B pairs
    with 16-byte loads; K = 640 runs twenty 32-k chunks, K = 6 one partial
    chunk."""
    shapes = [(200, 130, 300), (18, 34, 6), (64, 64, 640), (50, 96, 256)]
    descs, A, B, nc = _group_case(78, shapes)
    plan = tn.GemmPlan(mode, np.array(descs, dtype=tn.GEMM_DESC))
    dA, dB = tn.DeviceArray.from_host(A), tn.DeviceArray.from_host(B)
    outs = []
    for tile in range(NTILES[mode]):
        plan.set_tile(tile)
        dC = tn.DeviceArray((nc,))
        plan(dA, dB, dC)
        outs.append(dC.to_host())
    assert all(o.tobytes() == outs[0].tobytes() for o in outs)
    ref = np.concatenate([(A[d[3]:d[3] + d[0] * d[2]].reshape(d[0], d[2]) @
                           B[d[4]:d[4] + d[2] * d[1]].reshape(d[2], d[1])).ravel() for d in descs])
    assert np.linalg.norm(outs[0] - ref) <= GEMM_TOL[mode] * np.linalg.norm(ref)


@pytest.mark.parametrize("mode", [0, 1])
def test_gemm_index_maps_every_tile_same_bits(tn, mode):
    """Multi-level index maps on every tile shape: identical bits."""
    rng = np.random.default_rng(15 + mode)
    a = rng.uniform(-1, 1, (6, 5, 7, 4))
    M, K, N = 6 * 7, 5 * 4, 9
    bt = rng.uniform(-1, 1, (N, K))
    st = [5 * 7 * 4, 7 * 4, 4, 1]
    desc = np.zeros(1, dtype=tn.GEMM_DESC_EX)
    desc["M"], desc["N"], desc["K"] = M, N, K
    desc["lda"], desc["ldb"], desc["ldc"] = K, N, N
    desc["a_row"] = tn.index_map([(6, st[0]), (7, st[2])])
    desc["a_col"] = tn.index_map([(5, st[1]), (4, st[3])])
    desc["b_row"] = tn.index_map([(K, 1)])
    desc["b_col"] = tn.index_map([(N, K)])
    desc["c_row"] = tn.index_map([(M, 1)])
    desc["c_col"] = tn.index_map([(N, M)])
    plan = tn.GemmPlan(mode, desc)
    dA, dB = tn.DeviceArray.from_host(a), tn.DeviceArray.from_host(bt)
    outs = []
    for tile in range(NTILES[mode]):
        plan.set_tile(tile)
        dC = tn.DeviceArray((N, M))
        plan(dA, dB, dC)
        outs.append(dC.to_host().tobytes())
    assert all(o == outs[0] for o in outs)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_gemm_index_maps(tn, mode):
    """Multi-stride operands equal the same GEMM on explicitly permuted
    copies: A read as the (0,2)x(1,3) matricisation of a rank-4 tensor, B
    read transposed, C written as the transpose of a (5,N) x (M/5) layout."""
    rng = np.random.default_rng(5 + mode)
    d = (6, 5, 7, 4)
    a = rng.uniform(-1, 1, d)
    M, K, N = 6 * 7, 5 * 4, 9
    bt = rng.uniform(-1, 1, (N, K))           # stored as B^T
    st = [5 * 7 * 4, 7 * 4, 4, 1]
    desc = np.zeros(1, dtype=tn.GEMM_DESC_EX)
    desc["M"], desc["N"], desc["K"] = M, N, K
    desc["lda"], desc["ldb"], desc["ldc"] = K, N, N
    desc["a_row"] = tn.index_map([(6, st[0]), (7, st[2])])
    desc["a_col"] = tn.index_map([(5, st[1]), (4, st[3])])
    desc["b_row"] = tn.index_map([(K, 1)])
    desc["b_col"] = tn.index_map([(N, K)])
    desc["c_row"] = tn.index_map([(M, 1)])     # C stored transposed (N x M)
    desc["c_col"] = tn.index_map([(N, M)])
    plan = tn.GemmPlan(mode, desc)
    dA, dB, dC = tn.DeviceArray.from_host(a), tn.DeviceArray.from_host(bt), tn.DeviceArray((N, M))
    plan(dA, dB, dC)
    ref = (np.transpose(a, (0, 2, 1, 3)).reshape(M, K) @ bt.T).T
    assert np.linalg.norm(dC.to_host() - ref) <= GEMM_TOL[mode] * np.linalg.norm(ref)


def test_gemm_index_map_rejects_bad_extent(tn):
    desc = np.zeros(1, dtype=tn.GEMM_DESC_EX)
    desc["M"], desc["N"], desc["K"] = 6, 4, 3
    desc["lda"], desc["ldb"], desc["ldc"] = 3, 4, 4
    desc["a_row"] = tn.index_map([(2, 1), (2, 2)])   # covers 4 rows, M is 6
    desc["a_col"] = tn.index_map([(3, 1)])
    with pytest.raises(tn.TNError):
        tn.GemmPlan(0, desc)


# ---------------------------------------------------------------------------
# SVD / eigh
def _svd_check(tn, ora, a, tol=1e-10):
    M, N = a.shape
    A = tn.DeviceArray.from_host(a)
    U, S, VT = tn.svd(A, M, N)
    u, s, vt = U.to_host(), S.to_host(), VT.to_host()
    ur, sr, vtr = ora.svd(a)
    smax = max(sr.max(), 1e-300)
    assert np.max(np.abs(s - sr)) <= tol * smax
    assert np.all(np.diff(s) <= 0)
    assert np.linalg.norm(u @ np.diag(s) @ vt - a) <= 1e-12 * max(np.linalg.norm(a), 1e-300) * 10
    k = min(M, N)
    gaps = np.abs(np.diff(sr)) if k > 1 else np.array([1.0])
    for j in range(k):
        sep = min(gaps[j - 1] if j > 0 else np.inf, gaps[j] if j < k - 1 else np.inf)
        if sr[j] < 1e-8 * smax or sep < 1e-6 * smax:
            continue
        sg = np.sign(np.dot(u[:, j], ur[:, j])) or 1.0
        assert np.max(np.abs(sg * u[:, j] - ur[:, j])) <= 1e-8
        assert np.max(np.abs(sg * vt[j] - vtr[j])) <= 1e-8


@pytest.mark.parametrize("M,N", [(64, 64), (4, 5), (5, 4), (64, 32), (32, 64), (100, 80), (1, 7),
                                 (7, 1), (1, 1), (128, 128), (200, 90), (300, 300)])
def test_svd_random(tn, ora, M, N):
    _say("svd %dx%d" % (M, N))
    _svd_check(tn, ora, _rand(M * 1000 + N, (M, N)))


@pytest.mark.parametrize("M,N,r", [(64, 64, 32), (128, 128, 64), (40, 90, 7)])
def test_svd_rank_deficient(tn, ora, M, N, r):
    """theta of an iTEBD update has rank <= chi < min(M, N): the non-zero
    singular triplets must match LAPACK and the sweeps must terminate."""
    rng = np.random.default_rng(M + N + r)
    a = rng.uniform(-1, 1, (M, r)) @ rng.uniform(-1, 1, (r, N))
    A = tn.DeviceArray.from_host(a)
    u, s, vt = (x.to_host() for x in tn.svd(A, M, N))
    ur, sr, vtr = ora.svd(a)
    assert np.max(np.abs(s[:r] - sr[:r])) <= 1e-12 * sr[0]
    assert np.max(np.abs(s[r:])) <= 1e-12 * sr[0]
    assert np.linalg.norm(u[:, :r] @ np.diag(s[:r]) @ vt[:r] - a) <= 1e-12 * np.linalg.norm(a)
    assert np.allclose(u[:, :r].T @ u[:, :r], np.eye(r), atol=1e-12)


def test_svd_golden_matrix(tn, ora, golden):
    """Rebuild the 4x5 matrix from the reference's saved U, Sigma, VT dump and
    check the device SVD returns the dumped Sigma (descending) and factors."""
    U = np.fromfile(golden + "/mat_U", dtype=np.float64).reshape(4, 4)
    S = np.fromfile(golden + "/mat_Sigma", dtype=np.float64)
    VT = np.fromfile(golden + "/mat_VT", dtype=np.float64).reshape(4, 5)
    a = U @ np.diag(S) @ VT
    A = tn.DeviceArray.from_host(a)
    u, s, vt = (x.to_host() for x in tn.svd(A, 4, 5))
    assert np.max(np.abs(s - S)) <= 1e-13
    for j in range(4):
        sg = np.sign(np.dot(u[:, j], U[:, j]))
        assert np.max(np.abs(sg * u[:, j] - U[:, j])) <= 1e-12
        assert np.max(np.abs(sg * vt[j] - VT[j])) <= 1e-12


def test_svd_batched_uneven(tn, ora):
    shapes = [(64, 64), (10, 3), (3, 10), (50, 50), (1, 1), (120, 60)]
    mats = [_rand(i + 40, s) for i, s in enumerate(shapes)]
    A = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [m.size for m in mats])[:-1].astype(np.int64)
    ks = [min(s) for s in shapes]
    u_off = np.cumsum([0] + [s[0] * k for s, k in zip(shapes, ks)])[:-1].astype(np.int64)
    s_off = np.cumsum([0] + ks)[:-1].astype(np.int64)
    vt_off = np.cumsum([0] + [k * s[1] for s, k in zip(shapes, ks)])[:-1].astype(np.int64)
    dA = tn.DeviceArray.from_host(A)
    dU = tn.DeviceArray((int(sum(s[0] * k for s, k in zip(shapes, ks))),))
    dS = tn.DeviceArray((int(sum(ks)),))
    dV = tn.DeviceArray((int(sum(k * s[1] for s, k in zip(shapes, ks))),))
    m = np.array([s[0] for s in shapes], np.int64)
    n = np.array([s[1] for s in shapes], np.int64)
    tn.call("tn_svd_f64_batched", len(shapes), tn._p(m), tn._p(n), tn._p(dA), tn._p(offs),
            tn._p(dU), tn._p(u_off), tn._p(dS), tn._p(s_off), tn._p(dV), tn._p(vt_off), None)
    Sall = dS.to_host()
    for i, mat in enumerate(mats):
        sr = ora.svd(mat)[1]
        assert np.max(np.abs(Sall[s_off[i]:s_off[i] + ks[i]] - sr)) <= 1e-12 * sr.max()


def _sweeps(tn):
    v = ctypes.c_int()
    tn.call("tn_svd_last_sweeps", ctypes.byref(v))
    return v.value


@pytest.mark.parametrize("n", [16, 64, 100, 200])
def test_svd_warm_start_nearby_matrix(tn, ora, n):
    """Warm start from the VT of a nearby matrix: the same SVD (LAPACK
    singular values, reconstruction, orthonormal factors) in fewer sweeps."""
    _say("svd warm n=%d" % n)
    rng = np.random.default_rng(n)
    a0 = rng.uniform(-1, 1, (n, n))
    a1 = a0 + 1e-3 * rng.uniform(-1, 1, (n, n))
    A0 = tn.DeviceArray.from_host(a0)
    _, _, VT0 = tn.svd(A0, n, n)
    A1 = tn.DeviceArray.from_host(a1)
    tn.svd(A1, n, n)
    cold = _sweeps(tn)
    U, S, VT = tn.DeviceArray((n, n)), tn.DeviceArray((n,)), tn.DeviceArray((n, n))
    tn.call("tn_svd_f64_warm", tn._p(A1), n, n, tn._p(VT0), tn._p(U), tn._p(S), tn._p(VT), None)
    warm = _sweeps(tn)
    u, s, vt = U.to_host(), S.to_host(), VT.to_host()
    sr = ora.svd(a1)[1]
    assert np.max(np.abs(s - sr)) <= 1e-12 * sr[0]
    assert np.linalg.norm(u @ np.diag(s) @ vt - a1) <= 1e-12 * np.linalg.norm(a1)
    assert np.allclose(u.T @ u, np.eye(n), atol=1e-12)
    assert np.allclose(vt @ vt.T, np.eye(n), atol=1e-12)
    assert warm < cold, (warm, cold)


def _svd_no_u(tn, a, vt0=None):
    """tn_svd_f64 / tn_svd_f64_warm with U = NULL: S and VT only."""
    M, N = a.shape
    k = min(M, N)
    A = tn.DeviceArray.from_host(a)
    S, VT = tn.DeviceArray((k,)), tn.DeviceArray((k, N))
    if vt0 is None:
        tn.call("tn_svd_f64", tn._p(A), M, N, None, tn._p(S), tn._p(VT), None)
    else:
        V0 = tn.DeviceArray.from_host(np.ascontiguousarray(vt0))
        tn.call("tn_svd_f64_warm", tn._p(A), M, N, tn._p(V0), None, tn._p(S), tn._p(VT), None)
    return S.to_host(), VT.to_host()


def _check_s_vt(ora, a, s, vt, rank=None):
    """LAPACK's singular values; VT with orthonormal rows such that A VT^T
    has orthogonal columns of norms S (so A = (A VT^T) VT when M >= N); the
    rows of separated singular values equal LAPACK's up to sign."""
    M, N = a.shape
    k = min(M, N)
    r = k if rank is None else rank
    _, sr, vtr = ora.svd(a)
    s0 = max(sr[0], 1e-300)
    assert np.max(np.abs(s[:r] - sr[:r])) <= 1e-12 * s0
    assert np.max(np.abs(s[r:]), initial=0.0) <= 1e-12 * s0
    assert np.all(np.diff(s) <= 0)
    assert np.max(np.abs(vt @ vt.T - np.eye(k))) <= 1e-12
    # columns orthogonal to the sweeps' tolerance: 10 eps times the rows of
    # the swept matrix (n for the QR-preconditioned path, max(M, N) otherwise)
    us = a @ vt.T
    gtol = max(1e-12, 20 * max(M, N) * np.finfo(float).eps)
    assert np.max(np.abs(us.T @ us - np.diag(s * s))) <= gtol * s0 * s0
    if M >= N:
        assert np.linalg.norm(us @ vt - a) <= 1e-12 * max(np.linalg.norm(a), 1e-300)
    gaps = np.abs(np.diff(sr)) if k > 1 else np.array([np.inf])
    for j in range(r):
        sep = min(gaps[j - 1] if j > 0 else np.inf, gaps[j] if j < k - 1 else np.inf)
        if sr[j] < 1e-8 * s0 or sep < 1e-6 * s0:
            continue
        sg = np.sign(np.dot(vt[j], vtr[j])) or 1.0
        assert np.max(np.abs(sg * vt[j] - vtr[j])) <= 1e-8


@pytest.mark.parametrize("M,N", [(64, 64), (5, 4), (100, 64), (128, 33), (7, 1), (1, 1), (64, 63),
                                 (33, 5), (64, 8), (64, 9), (40, 16), (64, 17), (50, 24), (64, 40),
                                 (64, 48), (64, 56), (120, 57), (32, 32), (32, 31), (16, 13),
                                 (20, 50), (96, 96), (200, 150), (300, 300), (700, 513),
                                 (1030, 1030), (1500, 1100), (2100, 300), (3072, 1024), (4096, 520),
                                 (4100, 70)])
def test_svd_without_u(tn, ora, M, N):
    """U not wanted: for M >= N the solver factorises A = Q R by Householder
    reflections (with the columns in registers for n <= 64; blocked for the
    block path: 32-column panels up to M = 1024, 16-column ones up to 2048,
    8-column ones up to 4096) and sweeps R^T, reading VT off its normalised
    columns; M < N, and M > 4096 on the block path, take the plain sweeps
    without writing U.  The n <= 64 shapes cover every count of 8-column
    blocks, partly and wholly empty blocks, and the three row chunk sizes
    (M <= 32, 64, 128)."""
    _say("svd no U %dx%d" % (M, N))
    a = _rand(M * 1000 + N + 3, (M, N))
    s, vt = _svd_no_u(tn, a)
    _check_s_vt(ora, a, s, vt)


@pytest.mark.parametrize("n", [64, 160])
@pytest.mark.parametrize("case", ["rank20", "zero columns", "zero", "graded"])
def test_svd_without_u_special(tn, ora, case, n):
    """Rank-deficient inputs on the QR-preconditioned path: noise columns are
    orthogonalised relative to their own norms (VT stays orthonormal); an
    exactly zero column cannot be normalised and sends the problem through
    the plain sweeps.  A spectrum over twelve decades keeps its small values
    to LAPACK's absolute accuracy."""
    _say("svd no U %s n=%d" % (case, n))
    rng = np.random.default_rng(11 + n)
    r = None
    if case == "rank20":
        a, r = rng.uniform(-1, 1, (n, 20)) @ rng.uniform(-1, 1, (20, n)), 20
    elif case == "zero columns":
        k = (5 * n) // 8
        a, r = np.hstack([rng.uniform(-1, 1, (n, k)), np.zeros((n, n - k))]), k
    elif case == "zero":
        a, r = np.zeros((n // 4, n // 4)), 0
    else:
        a = _orth(rng, n) @ np.diag(10.0 ** -np.linspace(0, 12, n)) @ _orth(rng, n).T
    s, vt = _svd_no_u(tn, a)
    if r == 0:
        k = a.shape[0]
        assert np.all(s == 0) and np.max(np.abs(vt @ vt.T - np.eye(k))) <= 1e-12
        return
    _check_s_vt(ora, a, s, vt, rank=r)


def test_svd_without_u_warm_theta_like(tn, ora):
    """Warm start on a theta-like matrix (rows scaled by a lambda decaying
    over nine decades, as theta = lambda C) after a 1e-3 change of C: the
    QR-preconditioned sweeps give the same factors as the plain warm sweeps,
    in fewer sweeps (4 against 18 in a CPU emulation)."""
    rng = np.random.default_rng(5)
    n = 64
    lam = np.exp(-np.arange(n) / 3.0)
    c0 = rng.uniform(-1, 1, (n, n))
    a0 = lam[:, None] * c0
    a1 = lam[:, None] * (c0 + 1e-3 * rng.uniform(-1, 1, (n, n)))
    vt0 = ora.svd(a0)[2]
    s, vt = _svd_no_u(tn, a1, vt0)
    qr_sweeps = _sweeps(tn)
    _check_s_vt(ora, a1, s, vt)
    A1, V0 = tn.DeviceArray.from_host(a1), tn.DeviceArray.from_host(np.ascontiguousarray(vt0))
    U, S, VT = tn.DeviceArray((n, n)), tn.DeviceArray((n,)), tn.DeviceArray((n, n))
    tn.call("tn_svd_f64_warm", tn._p(A1), n, n, tn._p(V0), tn._p(U), tn._p(S), tn._p(VT), None)
    plain_sweeps = _sweeps(tn)
    assert np.max(np.abs(S.to_host() - s)) <= 1e-12 * s[0]
    assert qr_sweeps < plain_sweeps, (qr_sweeps, plain_sweeps)


@pytest.mark.parametrize("n", [64, 160])
def test_svd_without_u_chained_warm_starts(tn, ora, n):
    """Fifty warm starts in a row, each from the VT the previous call wrote
    (in place, as the TEBD engines do): the rows stay orthonormal to
    rounding (no drift from call to call) and S stays LAPACK's."""
    _say("svd no U chained n=%d" % n)
    rng = np.random.default_rng(3 + n)
    lam = np.exp(-np.arange(n) / 6.0)
    c = rng.uniform(-1, 1, (n, n))
    VT = tn.DeviceArray.from_host(np.eye(n))
    S = tn.DeviceArray((n,))
    for k in range(50):
        c = c + 1e-3 * rng.uniform(-1, 1, (n, n))
        a = lam[:, None] * c
        A = tn.DeviceArray.from_host(a)
        tn.call("tn_svd_f64_warm", tn._p(A), n, n, tn._p(VT), None, tn._p(S), tn._p(VT), None)
    vt, s = VT.to_host(), S.to_host()
    sr = ora.svd(a)[1]
    assert np.max(np.abs(s - sr)) <= 1e-12 * sr[0]
    assert np.max(np.abs(vt @ vt.T - np.eye(n))) <= 1e-13


def test_svd_plan_without_u(tn, ora):
    """A plan mixing problems that form U with problems that do not
    (u_off = -1) over the small, LDS and block paths, one warm-started; and
    a plan where none does, executed with U = NULL."""
    shapes = [(64, 64), (30, 20), (100, 90), (64, 64), (160, 150)]
    mats = [_rand(i + 170, s) for i, s in enumerate(shapes)]
    A = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [m.size for m in mats])[:-1].astype(np.int64)
    ks = [min(s) for s in shapes]
    s_off = np.cumsum([0] + ks)[:-1].astype(np.int64)
    vt_off = np.cumsum([0] + [k * s[1] for s, k in zip(shapes, ks)])[:-1].astype(np.int64)
    m = np.array([s[0] for s in shapes], np.int64)
    n = np.array([s[1] for s in shapes], np.int64)
    v0t = ora.svd(mats[0] + 1e-3 * _rand(5, (64, 64)))[2]
    v0_off = np.array([-1, -1, -1, 0, -1], np.int64)
    dA = tn.DeviceArray.from_host(A)
    dV0 = tn.DeviceArray.from_host(v0t)
    dS = tn.DeviceArray((int(sum(ks)),))
    dV = tn.DeviceArray((int(sum(k * s[1] for s, k in zip(shapes, ks))),))
    for with_u in (True, False):
        u_off = np.array([-1, 0, -1, -1, -1] if with_u else [-1] * 5, np.int64)
        dU = tn.DeviceArray((30 * 20,)) if with_u else None
        h = ctypes.c_void_p()
        tn.call("tn_svd_set_block_min", 128)
        try:
            tn.call("tn_svd_plan_create", ctypes.byref(h), len(shapes), tn._p(m), tn._p(n), tn._p(offs),
                    tn._p(v0_off), tn._p(u_off), tn._p(s_off), tn._p(vt_off))
        finally:
            tn.call("tn_svd_set_block_min", 0)
        plan = ctypes.c_void_p(h.value)
        try:
            tn.call("tn_svd_plan_exec", plan, tn._p(dA), tn._p(dV0), tn._p(dU) if with_u else None, tn._p(dS),
                    tn._p(dV), None)
            nc, sw = ctypes.c_int(), ctypes.c_int()
            tn.call("tn_svd_plan_check", plan, ctypes.byref(nc), ctypes.byref(sw), None)
            assert nc.value == 0
        finally:
            tn.call("tn_svd_plan_destroy", plan)
        Sall, Vall = dS.to_host(), dV.to_host()
        for i, mat in enumerate(mats):
            k = ks[i]
            s = Sall[s_off[i]:s_off[i] + k]
            vt = Vall[vt_off[i]:vt_off[i] + k * shapes[i][1]].reshape(k, shapes[i][1])
            _check_s_vt(ora, mat, s, vt)
        if with_u:
            u = dU.to_host().reshape(30, 20)
            s = Sall[s_off[1]:s_off[1] + 20]
            vt = Vall[vt_off[1]:vt_off[1] + 400].reshape(20, 20)
            assert np.linalg.norm(u @ np.diag(s) @ vt - mats[1]) <= 1e-12 * np.linalg.norm(mats[1])


@pytest.mark.parametrize("M,N", [(100, 80), (80, 100), (96, 96), (70, 65), (130, 67)])
def test_svd_one_workgroup_lds_path(tn, ora, M, N):
    """Matrices beyond the small kernel that fit one workgroup's LDS, kept on
    the one-workgroup LDS kernel instead of block Jacobi: LAPACK's singular
    values and vectors."""
    _say("svd lds %dx%d" % (M, N))
    tn.call("tn_svd_set_block_min", 1 << 20)
    try:
        _svd_check(tn, ora, _rand(M * 1000 + N + 7, (M, N)))
    finally:
        tn.call("tn_svd_set_block_min", 0)


def _orth(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, n)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize("n", [192, 320])
def test_svd_block_graded_spectrum(tn, ora, n):
    """A theta-like spectrum decaying over ten decades, cold start: every
    singular value to 1e-12 of the largest, the factors reconstruct A, and
    the well-separated leading vectors match LAPACK's."""
    _say("svd graded n=%d" % n)
    rng = np.random.default_rng(n)
    s = np.exp(-np.arange(n) * (23.0 / n))
    a = _orth(rng, n) @ np.diag(s) @ _orth(rng, n).T
    A = tn.DeviceArray.from_host(a)
    u, sg, vt = (x.to_host() for x in tn.svd(A, n, n))
    ur, sr, vtr = ora.svd(a)
    assert np.max(np.abs(sg - sr)) <= 1e-12 * sr[0]
    assert np.linalg.norm(u @ np.diag(sg) @ vt - a) <= 1e-12 * np.linalg.norm(a)
    for j in range(n):
        if sr[j] < 1e-6:
            break
        sgn = np.sign(np.dot(u[:, j], ur[:, j]))
        assert np.max(np.abs(sgn * u[:, j] - ur[:, j])) <= 1e-8
        assert np.max(np.abs(sgn * vt[j] - vtr[j])) <= 1e-8


def test_svd_block_large_warm(tn, ora):
    """512 x 512 (the theta of bond dimension 256): cold, then warm-started
    from the right singular vectors of a nearby matrix, as the iTEBD update
    does; both give LAPACK's singular values, and the warm start needs fewer
    sweeps."""
    n = 512
    rng = np.random.default_rng(5)
    s = np.exp(-np.arange(n) * (20.0 / n))
    a0 = _orth(rng, n) @ np.diag(s) @ _orth(rng, n).T
    a1 = a0 + 1e-2 * (rng.uniform(-1, 1, (n, n)) / np.sqrt(n)) @ a0
    A0 = tn.DeviceArray.from_host(a0)
    _say("svd 512 cold")
    _, S0, VT0 = tn.svd(A0, n, n)
    cold = _sweeps(tn)
    assert np.max(np.abs(S0.to_host() - ora.svd(a0)[1])) <= 1e-12
    A1 = tn.DeviceArray.from_host(a1)
    U, S, VT = tn.DeviceArray((n, n)), tn.DeviceArray((n,)), tn.DeviceArray((n, n))
    _say("svd 512 warm")
    tn.call("tn_svd_f64_warm", tn._p(A1), n, n, tn._p(VT0), tn._p(U), tn._p(S), tn._p(VT), None)
    warm = _sweeps(tn)
    _say("sweeps cold %d warm %d" % (cold, warm))
    u, sg, vt = U.to_host(), S.to_host(), VT.to_host()
    sr = ora.svd(a1)[1]
    assert np.max(np.abs(sg - sr)) <= 1e-12 * sr[0]
    assert np.linalg.norm(u @ np.diag(sg) @ vt - a1) <= 1e-12 * np.linalg.norm(a1)
    assert np.allclose(vt @ vt.T, np.eye(n), atol=1e-12)
    k = int(np.sum(sr > 1e-6))
    assert np.allclose(u[:, :k].T @ u[:, :k], np.eye(k), atol=1e-10)
    assert warm < cold, (warm, cold)


def test_svd_plan_replay_and_graph(tn, ora):
    """A batch plan over every kernel path (small, LDS, block Jacobi), with one
    warm-started problem: direct execution, a second execution and graph
    replays all give the LAPACK singular values."""
    shapes = [(64, 64), (30, 20), (100, 90), (160, 150), (64, 64), (200, 200)]
    mats = [_rand(i + 70, s) for i, s in enumerate(shapes)]
    A = np.concatenate([m.reshape(-1) for m in mats])
    offs = np.cumsum([0] + [m.size for m in mats])[:-1].astype(np.int64)
    ks = [min(s) for s in shapes]
    u_off = np.cumsum([0] + [s[0] * k for s, k in zip(shapes, ks)])[:-1].astype(np.int64)
    s_off = np.cumsum([0] + ks)[:-1].astype(np.int64)
    vt_off = np.cumsum([0] + [k * s[1] for s, k in zip(shapes, ks)])[:-1].astype(np.int64)
    m = np.array([s[0] for s in shapes], np.int64)
    n = np.array([s[1] for s in shapes], np.int64)
    # problem 4 starts from the right singular vectors of problem 0, problem 5
    # (block Jacobi) from those of a nearby matrix
    v0t = ora.svd(mats[0])[2]
    v5 = ora.svd(mats[5] + 1e-3 * _rand(99, (200, 200)))[2]
    v0_off = np.array([-1, -1, -1, -1, 0, v0t.size], np.int64)
    dA = tn.DeviceArray.from_host(A)
    dV0 = tn.DeviceArray.from_host(np.concatenate([v0t.reshape(-1), v5.reshape(-1)]))
    dU = tn.DeviceArray((int(sum(s[0] * k for s, k in zip(shapes, ks))),))
    dS = tn.DeviceArray((int(sum(ks)),))
    dV = tn.DeviceArray((int(sum(k * s[1] for s, k in zip(shapes, ks))),))
    h = ctypes.c_void_p()
    tn.call("tn_svd_set_block_min", 128)   # 100 x 90 on the LDS kernel, 160 x 150 and up on block Jacobi
    try:
        tn.call("tn_svd_plan_create", ctypes.byref(h), len(shapes), tn._p(m), tn._p(n), tn._p(offs),
                tn._p(v0_off), tn._p(u_off), tn._p(s_off), tn._p(vt_off))
    finally:
        tn.call("tn_svd_set_block_min", 0)
    plan = ctypes.c_void_p(h.value)
    stream = tn.stream_create()

    def run():
        tn.call("tn_svd_plan_exec", plan, tn._p(dA), tn._p(dV0), tn._p(dU), tn._p(dS), tn._p(dV), stream)

    def check():
        nc, sw = ctypes.c_int(), ctypes.c_int()
        tn.call("tn_svd_plan_check", plan, ctypes.byref(nc), ctypes.byref(sw), stream)
        assert nc.value == 0 and sw.value > 0
        Sall = dS.to_host()
        for i, mat in enumerate(mats):
            sr = ora.svd(mat)[1]
            assert np.max(np.abs(Sall[s_off[i]:s_off[i] + ks[i]] - sr)) <= 1e-12 * sr.max()

    def clear():
        tn.call("tn_zero", tn._p(dS), dS.nbytes, stream)

    try:
        for _ in range(2):
            clear()
            run()
            check()
        g = tn.Graph(run, stream)
        for _ in range(3):
            clear()
            g.launch()
            check()
        del g
    finally:
        tn.call("tn_svd_plan_destroy", plan)


def test_eigh_heisenberg_gate(tn, ora):
    H = ora.heisenberg_two_site(1.0)
    dH = tn.DeviceArray.from_host(H)
    w, v = (x.to_host() for x in tn.eigh(dH, 4))
    assert np.allclose(w, [-0.75, 0.25, 0.25, 0.25], atol=1e-15)
    # rows are eigenvectors
    assert np.allclose(v @ H @ v.T, np.diag(w), atol=1e-14)
    # exph from these factors equals the closed form
    tau = 0.01
    G = v.T @ np.diag(np.exp(-tau * w)) @ v
    Gref = ora.exph(-tau, H)
    assert np.max(np.abs(G - Gref)) <= 1e-15


@pytest.mark.parametrize("n", [2, 9, 20, 64])
def test_eigh_random(tn, ora, n):
    a = _rand(n, (n, n))
    a = a + a.T
    dA = tn.DeviceArray.from_host(a)
    w, v = (x.to_host() for x in tn.eigh(dA, n))
    wr, vr = ora.eigh(a)
    assert np.max(np.abs(w - wr)) <= 1e-12 * np.abs(wr).max()
    assert np.max(np.abs(v @ a @ v.T - np.diag(w))) <= 1e-12 * np.abs(wr).max()


def test_eigh_batched_matches_single_and_oracle(tn, ora):
    """One launch over matrices of uneven size (n = 0 skipped) gives the
    bits of one launch per matrix, and each equals dsyev's eigenpairs."""
    ns = [4, 0, 9, 1, 16, 64, 90, 9]
    mats = []
    for i, n in enumerate(ns):
        a = _rand(40 + i, (n, n)) if n else np.zeros((0, 0))
        mats.append(a + a.T)
    ao, eo, vo = [], [], []
    flat, e_n, v_n = [], 0, 0
    for a in mats:
        ao.append(sum(x.size for x in flat))
        flat.append(a.ravel())
        eo.append(e_n)
        vo.append(v_n)
        e_n += a.shape[0]
        v_n += a.size
    dA = tn.DeviceArray.from_host(np.concatenate(flat))
    de, dv = tn.DeviceArray((max(e_n, 1),)), tn.DeviceArray((max(v_n, 1),))
    keep = [np.array(x, np.int64) for x in (ns, ao, eo, vo)]
    p64 = [x.ctypes.data_as(tn.PI64) for x in keep]
    tn.call("tn_eigh_f64_batched", len(ns), p64[0], tn._p(dA), p64[1], tn._p(de), p64[2], tn._p(dv), p64[3],
            None)
    e, v = de.to_host(), dv.to_host()
    for i, (n, a) in enumerate(zip(ns, mats)):
        if n == 0:
            continue
        w = e[eo[i]:eo[i] + n]
        vec = v[vo[i]:vo[i] + n * n].reshape(n, n)
        w1, v1 = (x.to_host() for x in tn.eigh(tn.DeviceArray.from_host(a), n))
        assert w.tobytes() == w1.tobytes() and vec.tobytes() == v1.tobytes()
        wr, _ = ora.eigh(a)
        assert np.max(np.abs(w - wr)) <= 1e-12 * np.abs(wr).max()
        assert np.max(np.abs(vec @ a @ vec.T - np.diag(w))) <= 1e-12 * np.abs(wr).max()


def test_eigh_batched_rejects_large_n(tn):
    a = tn.DeviceArray((91 * 91,))
    e, v = tn.DeviceArray((91,)), tn.DeviceArray((91 * 91,))
    z, n = np.zeros(1, np.int64), np.array([91], np.int64)
    pz, pn = z.ctypes.data_as(tn.PI64), n.ctypes.data_as(tn.PI64)
    with pytest.raises(tn.TNError):
        tn.call("tn_eigh_f64_batched", 1, pn, tn._p(a), pz, tn._p(e), pz, tn._p(v), pz, None)


def _crand(seed, shape):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def _c2d(z):
    """complex array -> interleaved float64 (device layout)"""
    return np.ascontiguousarray(z, dtype=np.complex128).view(np.float64).ravel()


def _d2c(x, shape):
    return np.ascontiguousarray(x).view(np.complex128).reshape(shape)


@pytest.mark.parametrize("shape", [(5, 3), (3, 5), (16, 16), (40, 24), (24, 40), (64, 64), (1, 7)])
def test_svd_c128_matches_lapack(tn, shape):
    """Complex SVD (zgesvd's A = U diag(S) VT): singular values against
    numpy's LAPACK to 1e-12, A rebuilt to 1e-12, U and VT orthonormal."""
    M, N = shape
    k = min(M, N)
    a = _crand(60 + M * N, (M, N))
    dA = tn.DeviceArray.from_host(_c2d(a))
    dU, dS, dV = tn.DeviceArray((2 * M * k,)), tn.DeviceArray((k,)), tn.DeviceArray((2 * k * N,))
    tn.call("tn_svd_c128", tn._p(dA), M, N, tn._p(dU), tn._p(dS), tn._p(dV), None)
    U, S, VT = _d2c(dU.to_host(), (M, k)), dS.to_host(), _d2c(dV.to_host(), (k, N))
    sr = np.linalg.svd(a, compute_uv=False)
    assert np.max(np.abs(S - sr)) <= 1e-12 * sr[0]
    assert np.all(np.diff(S) <= 0)
    assert np.linalg.norm(U @ np.diag(S) @ VT - a) <= 1e-12 * np.linalg.norm(a)
    assert np.max(np.abs(U.conj().T @ U - np.eye(k))) <= 1e-12
    assert np.max(np.abs(VT @ VT.conj().T - np.eye(k))) <= 1e-12


@pytest.mark.parametrize("n", [1, 4, 9, 33, 64])
def test_eigh_c128_matches_lapack(tn, n):
    """Hermitian eigen solver: eigenvalues ascending against numpy's LAPACK
    (zheev's absolute accuracy, 1e-12 of the largest), H = E^H diag(w) E to
    1e-11 (the positive shift 2 ||H||_F costs a few digits of absolute
    accuracy against zheev) and E E^H = I with the eigenvector rows E."""
    a = _crand(70 + n, (n, n))
    h = a + a.conj().T
    dH = tn.DeviceArray.from_host(_c2d(h))
    dw, dE = tn.DeviceArray((n,)), tn.DeviceArray((2 * n * n,))
    tn.call("tn_eigh_c128", tn._p(dH), n, tn._p(dw), tn._p(dE), None)
    w, E = dw.to_host(), _d2c(dE.to_host(), (n, n))
    wr = np.linalg.eigvalsh(h)
    scale = np.abs(wr).max()
    assert np.max(np.abs(w - wr)) <= 1e-12 * scale
    assert np.linalg.norm(E.conj().T @ np.diag(w) @ E - h) <= 1e-11 * np.linalg.norm(h)
    assert np.max(np.abs(E @ E.conj().T - np.eye(n))) <= 1e-12


def test_expand_c128(tn):
    z = _crand(80, 37)
    dz, dx = tn.DeviceArray.from_host(_c2d(z)), tn.DeviceArray((4 * 37,))
    tn.call("tn_expand_c128", tn._p(dz), 37, tn._p(dx), None)
    x = dx.to_host().reshape(37, 2, 2)
    assert np.array_equal(x[:, 0, 0], z.real) and np.array_equal(x[:, 0, 1], z.imag)
    assert np.array_equal(x[:, 1, 0], -z.imag) and np.array_equal(x[:, 1, 1], z.real)


# ---------------------------------------------------------------------------
def test_vector_ops(tn):
    x = _rand(9, 12345)
    dx = tn.DeviceArray.from_host(x)
    out = tn.DeviceArray((1,))
    tn.call("tn_norm2_f64", tn._p(dx), x.size, tn._p(out), None)
    assert abs(out.to_host()[0] - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    tn.call("tn_scale_f64", tn._p(dx), x.size, ctypes.c_double(0.5), None)
    assert np.array_equal(dx.to_host(), x * 0.5)
    M, N = 37, 53
    m = _rand(1, (M, N))
    d = _rand(2, M)
    dm = tn.DeviceArray.from_host(m)
    dd = tn.DeviceArray.from_host(d)
    tn.call("tn_diag_row_mul_f64", tn._p(dm), tn._p(dd), M, N, None)
    assert np.array_equal(dm.to_host(), m * d[:, None])
    s = np.sort(np.abs(_rand(3, 64)))[::-1].copy()
    ds = tn.DeviceArray.from_host(s)
    lam = tn.DeviceArray((32,))
    nrm = tn.DeviceArray((1,))
    err = tn.DeviceArray((1,))
    tn.call("tn_truncate_f64", tn._p(ds), 64, 32, tn._p(lam), tn._p(nrm), tn._p(err), None)
    ref_n = np.sqrt(np.sum(s[:32] ** 2))
    assert abs(nrm.to_host()[0] - ref_n) <= 1e-14 * ref_n
    assert np.allclose(lam.to_host(), s[:32] / ref_n, rtol=1e-14, atol=0)
    assert abs(err.to_host()[0] - np.sum(s[32:] ** 2)) <= 1e-14


# ---------------------------------------------------------------------------
# batched vector operations
def test_vecops_plan(tn):
    rng = np.random.default_rng(9)
    base = rng.uniform(-1, 1, 20000)
    base[100:140] = np.sort(np.abs(base[100:140]))[::-1]     # a descending S
    VEC = np.dtype([("op", "<i4"), ("pad", "<i4"), ("n", "<i8"), ("src", "<i8"), ("dst", "<i8"),
                    ("aux", "<i8"), ("rep", "<i8")])
    items = np.array([(0, 0, 9000, 0, 10000, 200, 300),     # rows of 300 scaled by base[200 + r]
                      (1, 0, 500, 9000, 19000, 0, 0),       # copy
                      (2, 0, 40, 300, 19600, 150, 0),       # / base[150]
                      (3, 0, 40, 100, 19700, 19800, 25)],   # truncate S to 25
                     dtype=VEC)
    d = tn.DeviceArray.from_host(base)
    h = ctypes.c_void_p()
    tn.call("tn_vecops_plan_create", ctypes.byref(h), len(items), items.ctypes.data_as(ctypes.c_void_p))
    try:
        tn.call("tn_vecops_plan_exec", ctypes.c_void_p(h.value), tn._p(d), None)
        out = d.to_host()
    finally:
        tn.call("tn_vecops_plan_destroy", ctypes.c_void_p(h.value))
    ref = base.copy()
    ref[10000:19000] = base[0:9000] * np.repeat(base[200:230], 300)
    ref[19000:19500] = base[9000:9500]
    ref[19600:19640] = base[300:340] / base[150]
    S = base[100:140]
    nrm = np.sqrt(np.sum(S[:25] ** 2))
    ref[19700:19725] = S[:25] / nrm
    ref[19800] = nrm
    ref[19801] = np.sum(S[25:] ** 2)
    assert np.max(np.abs(out - ref)) <= 1e-14


def test_lanczos_c_abi(tn, ora):
    # tn_lanczos_f64 on device buffers against the oracle's lanczosEV
    import ctypes
    n = 500
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, (n, n))
    S = X @ X.T + X.T @ X
    start = rng.uniform(-1, 1, n)
    A = tn.DeviceArray.from_host(S)
    psi = tn.DeviceArray.from_host(start)
    it = ctypes.c_int64(1000)
    e0 = ctypes.c_double()
    conv = ctypes.c_int()
    tn.call("tn_lanczos_f64", tn._p(A), n, tn._p(psi), ctypes.byref(it), 5e-15, ctypes.byref(e0),
            ctypes.byref(conv), None)
    e_o, psi_o, it_o, conv_o = ora.lanczos_ev(S, start)
    assert conv.value == 1 and conv_o
    assert abs(e0.value - e_o) <= 1e-12 * np.abs(np.linalg.eigvalsh(S)).max()
    assert abs(it.value - it_o) <= 3
    v = psi.to_host()
    assert np.dot(v, psi_o) / (np.linalg.norm(v) * np.linalg.norm(psi_o)) > 1 - 1e-8
    # the start vector's norm does not matter
    psi2 = tn.DeviceArray.from_host(start * 1e3)
    it2 = ctypes.c_int64(1000)
    tn.call("tn_lanczos_f64", tn._p(A), n, tn._p(psi2), ctypes.byref(it2), 5e-15, ctypes.byref(e0),
            ctypes.byref(conv), None)
    assert abs(e0.value - e_o) <= 1e-12 * np.abs(np.linalg.eigvalsh(S)).max()


def test_pool_reuse_across_streams_waits_for_the_freeing_stream(tn):
    """A block freed on stream A while A still writes it, then allocated on
    stream B and cleared there: B's clear lands after A's copy."""
    import ctypes
    n = 1 << 25                                   # 256 MiB of doubles
    tn.call("tn_pool_trim")                       # no other cached block of this size
    a, b = tn.stream_create(), tn.stream_create()
    try:
        src = tn.DeviceArray.from_host(np.ones(n))
        tn.pool_set_stream(a)
        x = tn.DeviceArray(n)
        ptr = x.ptr
        for _ in range(4):                        # ~0.3 ms of copies queued on A
            tn.call("tn_memcpy_d2d", ctypes.c_void_p(x.ptr), ctypes.c_void_p(src.ptr), 8 * n, a)
        x.free()
        tn.pool_set_stream(b)
        y = tn.DeviceArray(n)
        assert y.ptr == ptr                       # the same block, reused on B
        tn.call("tn_zero", ctypes.c_void_p(y.ptr), 8 * n, b)
        tn.call("tn_stream_sync", b)
        out = y.to_host(b)
        assert not out.any()
        tn.call("tn_stream_sync", a)
    finally:
        tn.pool_set_stream(None)
        tn.call("tn_stream_destroy", a)
        tn.call("tn_stream_destroy", b)


def test_lanczos_non_finite_is_an_error(tn):
    """A NaN or an overflow in the operator is reported as no convergence
    instead of a Krylov closure with a meaningless E0."""
    import ctypes
    n = 64
    rng = np.random.default_rng(12)
    X = rng.uniform(-1, 1, (n, n))
    for bad in (np.nan, np.inf):
        S = X + X.T
        S[3, 5] = S[5, 3] = bad
        A = tn.DeviceArray.from_host(S)
        psi = tn.DeviceArray.from_host(rng.uniform(-1, 1, n))
        it, e0, conv = ctypes.c_int64(100), ctypes.c_double(), ctypes.c_int()
        with pytest.raises(tn.TNError, match="non-finite"):
            tn.call("tn_lanczos_f64", tn._p(A), n, tn._p(psi), ctypes.byref(it), 5e-15, ctypes.byref(e0),
                    ctypes.byref(conv), None)


@pytest.mark.parametrize("dims,perm", [((7,), (0,)), ((3, 5), (1, 0)), ((4, 3, 5), (2, 0, 1)),
                                       ((256, 2, 2, 256), (3, 1, 0, 2)), ((2, 3, 1, 4, 5), (4, 2, 0, 3, 1))])
def test_complex_permute_bit_exact(tn, dims, perm):
    # complex dense permute: numpy's transpose of the complex array, bit for bit
    rng = np.random.default_rng(len(dims))
    n = int(np.prod(dims))
    z = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    src = tn.DeviceArray.from_host(z.view(np.float64))
    dst = tn.DeviceArray((2 * n,))
    d = np.array(dims, np.int64)
    p = np.array(perm, np.int32)
    tn.call("tn_permute_dense_c128", tn._p(src), tn._p(dst), len(dims), tn._p(d), tn._p(p), None)
    out = dst.to_host().view(np.complex128)
    want = np.ascontiguousarray(z.reshape(dims).transpose(perm)).reshape(-1)
    assert np.array_equal(out.view(np.float64), want.view(np.float64))


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 9), (128, 64, 256), (300, 200, 100), (5, 7, 0)])
def test_complex_matmul(tn, M, N, K):
    # complex product against numpy's (the textbook routine), 1e-12 relative
    rng = np.random.default_rng(M + N + K)
    a = rng.uniform(-1, 1, (M, K)) + 1j * rng.uniform(-1, 1, (M, K))
    b = rng.uniform(-1, 1, (K, N)) + 1j * rng.uniform(-1, 1, (K, N))
    A = tn.DeviceArray.from_host(np.ascontiguousarray(a).view(np.float64).reshape(-1)) if K else None
    B = tn.DeviceArray.from_host(np.ascontiguousarray(b).view(np.float64).reshape(-1)) if K else None
    C = tn.DeviceArray.from_host(np.full(2 * M * N, np.nan))
    tn.call("tn_matmul_c128", tn._p(A), tn._p(B), M, N, K, tn._p(C), None)
    c = C.to_host().view(np.complex128).reshape(M, N)
    want = a @ b
    if K == 0:
        assert np.array_equal(c, np.zeros((M, N), complex))
    else:
        assert np.linalg.norm(c - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("n", [10, 64, 300])
def test_lanczos_complex_c_abi(tn, ora, n):
    # tn_lanczos_c128 against the oracle's complex lanczosEV
    import ctypes
    rng = np.random.default_rng(60 + n)
    X = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    S = np.ascontiguousarray(X @ X.conj().T + X.conj().T @ X)
    start = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    A = tn.DeviceArray.from_host(S.view(np.float64).reshape(-1))
    psi = tn.DeviceArray.from_host(start.view(np.float64))
    it = ctypes.c_int64(1000)
    e0 = ctypes.c_double()
    conv = ctypes.c_int()
    tn.call("tn_lanczos_c128", tn._p(A), n, tn._p(psi), ctypes.byref(it), 5e-15, ctypes.byref(e0),
            ctypes.byref(conv), None)
    e_o, psi_o, it_o, conv_o = ora.lanczos_ev(S, start)
    scale = np.abs(np.linalg.eigvalsh(S)).max()
    assert conv.value == 1 and conv_o
    assert abs(e0.value - e_o) <= 1e-12 * scale
    assert abs(it.value - it_o) <= 3
    v = psi.to_host().view(np.complex128)
    assert np.vdot(start, v).real > 0
    cos = abs(np.vdot(v, psi_o)) / (np.linalg.norm(v) * np.linalg.norm(psi_o))
    assert cos > 1 - 1e-8
    assert np.linalg.norm(S @ v - e0.value * v) <= 1e-4 * scale * np.linalg.norm(v)


@pytest.mark.parametrize("shape", [(48, 40), (200, 160)])
def test_svd_plan_reports_failures_of_every_execution(tn, shape):
    """A problem that hits the sweep cap in one execution is reported by the
    next check even though a later execution converged; the check clears
    the counts.  Small kernel (48 x 40) and block path (200 x 160), S and VT
    only (QR-preconditioned sweeps).  The cap is forced from the host, so
    every kernel still stops normally."""
    M, N = shape
    print("svd status accumulation %dx%d" % shape, flush=True)
    hard = _rand(31, shape)
    easy = np.zeros(shape)
    easy[np.arange(N), np.arange(N)] = np.arange(N, 0, -1, dtype=np.float64)   # orthogonal columns
    dA, dE = tn.DeviceArray.from_host(hard), tn.DeviceArray.from_host(easy)
    dS, dV = tn.DeviceArray((N,)), tn.DeviceArray((N * N,))
    m, n = np.array([M], np.int64), np.array([N], np.int64)
    zero, none = np.zeros(1, np.int64), np.full(1, -1, np.int64)
    h = ctypes.c_void_p()
    tn.call("tn_svd_set_max_sweeps", 2)
    try:
        tn.call("tn_svd_plan_create", ctypes.byref(h), 1, tn._p(m), tn._p(n), tn._p(zero), tn._p(none),
                tn._p(none), tn._p(zero), tn._p(zero))
    finally:
        tn.call("tn_svd_set_max_sweeps", 0)
    plan = ctypes.c_void_p(h.value)
    nc, sw = ctypes.c_int(), ctypes.c_int()
    try:
        for src in (dA, dE):
            tn.call("tn_svd_plan_exec", plan, tn._p(src), None, None, tn._p(dS), tn._p(dV), None)
        tn.call("tn_svd_plan_check", plan, ctypes.byref(nc), ctypes.byref(sw), None)
        assert nc.value == 1 and sw.value >= 2
        tn.call("tn_svd_plan_check", plan, ctypes.byref(nc), ctypes.byref(sw), None)
        assert nc.value == 0 and sw.value == 0
        tn.call("tn_svd_plan_exec", plan, tn._p(dE), None, None, tn._p(dS), tn._p(dV), None)
        tn.call("tn_svd_plan_check", plan, ctypes.byref(nc), ctypes.byref(sw), None)
        assert nc.value == 0
        assert np.allclose(dS.to_host(), np.arange(N, 0, -1), rtol=1e-14, atol=0)
    finally:
        tn.call("tn_svd_plan_destroy", plan)


def _mem_stats(tn):
    v = [ctypes.c_size_t() for _ in range(4)]
    tn.call("tn_mem_stats", *[ctypes.byref(x) for x in v])
    return [x.value for x in v]


def test_budget_exhaustion_fails_cleanly_and_plans_stay_usable(tn):
    """An allocation past tn_set_budget fails with TN_ENOMEM and a message,
    nothing is leaked, and allocation works again once the budget allows it.
    A GEMM plan whose autotuner runs out of budget while building the tile
    lists of other shapes keeps its current shape and still computes the
    right product."""
    print("budget exhaustion", flush=True)
    M, N, K = 96, 80, 64
    a, b = _rand(41, (M, K)), _rand(42, (K, N))
    dA, dB, dC = tn.DeviceArray.from_host(a), tn.DeviceArray.from_host(b), tn.DeviceArray((M, N))
    plan = tn.GemmPlan(tn.TN_GEMM_F64, np.array([(M, N, K, 0, 0, 0, K, N, N)], dtype=tn.GEMM_DESC))
    before = plan.info()["tile_id"]
    in_use = _mem_stats(tn)[0]
    tn.call("tn_set_budget", in_use)          # no headroom: every new allocation fails
    try:
        assert _mem_stats(tn)[3] == in_use
        with pytest.raises(tn.TNError, match="budget exceeded"):
            tn.DeviceArray((1 << 20,))
        assert _mem_stats(tn)[0] == in_use
        with pytest.raises(tn.TNError, match="budget exceeded"):
            plan.autotune(dA, dB, dC, reps=1)
    finally:
        tn.call("tn_set_budget", 0)
    tn.sync()
    assert plan.info()["tile_id"] == before
    plan(dA, dB, dC)
    c = dC.to_host()
    assert np.linalg.norm(c - a @ b) <= 1e-13 * np.linalg.norm(a @ b)
    x = tn.DeviceArray((1 << 20,))       # the budget no longer binds
    x.free()
    plan.autotune(dA, dB, dC, reps=1)
    plan(dA, dB, dC)
    assert np.linalg.norm(dC.to_host() - a @ b) <= 1e-13 * np.linalg.norm(a @ b)


def test_gemm_exact_mode_is_exact_on_fixed_point_inputs(tn):
    """The exact split-BF16 mode on inputs that its first slices hold exactly
    (integers up to 1000 times 2^-10: 11 significant bits, two 7-bit slices
    after the row / column scaling): every slice product and sum is exact,
    so the result equals numpy's FP64 product (itself exact here: integer
    sums far below 2^53) bit for bit, for every tile shape, including uneven
    and K > 128 (several FP64 flushes) groups."""
    rng = np.random.default_rng(21)
    shapes = [(37, 29, 300), (64, 64, 64), (5, 130, 7), (100, 3, 513)]
    descs, A, B, co = [], [], [], 0
    ao = bo = 0
    for (M, N, K) in shapes:
        a = rng.integers(-1000, 1001, (M, K)) * 2.0 ** -10
        b = rng.integers(-1000, 1001, (K, N)) * 2.0 ** -10
        descs.append((M, N, K, ao, bo, co, K, N, N))
        A.append(a.reshape(-1))
        B.append(b.reshape(-1))
        ao += M * K
        bo += K * N
        co += M * N
    A, B = np.concatenate(A), np.concatenate(B)
    plan = tn.GemmPlan(tn.TN_GEMM_BF16_EXACT, np.array(descs, dtype=tn.GEMM_DESC))
    dA, dB = tn.DeviceArray.from_host(A), tn.DeviceArray.from_host(B)
    for tile in range(NTILES[4]):
        _say("exact mode fixed point, tile %d" % tile)
        plan.set_tile(tile)
        dC = tn.DeviceArray((co,))
        plan(dA, dB, dC)
        C = dC.to_host()
        for (M, N, K, a0, b0, c0, *_r) in descs:
            ref = A[a0:a0 + M * K].reshape(M, K) @ B[b0:b0 + K * N].reshape(K, N)
            assert C[c0:c0 + M * N].tobytes() == ref.reshape(-1).tobytes()
