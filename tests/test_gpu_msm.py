"""GPU parity tests: the HIP MSM engine (through the C ABI) against the CPU
oracle, the reference tests' closed forms and the committed golden vectors.
Comparison is bit-exact on the affine result in Montgomery form."""
import numpy as np
import pytest

import gpu_msm_amd as gm
import pymodel as m
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def aff_g1(rec):
    return m.g1_from_abi(rec)


def proj_to_aff(curve, proj):
    out = np.zeros(1, dtype=gm.G1_DTYPE if curve == "g1" else gm.G2_DTYPE)
    fn = "projective_to_affine_g1_wrapper" if curve == "g1" else "projective_to_affine_g2_wrapper"
    getattr(gm.lib(), fn)(out.ctypes.data, proj.ctypes.data)
    return out


def golden_point(curve, rec):
    if rec is None:
        return None
    if curve == "g1":
        return (int(rec[0], 16), int(rec[1], 16))
    return ((int(rec[0][0], 16), int(rec[0][1], 16)), (int(rec[1][0], 16), int(rec[1][1], 16)))


def as_point(curve, rec):
    return m.g1_from_abi(rec) if curve == "g1" else m.g2_from_abi(rec)


def same(a, b):
    return a.tobytes() == b.tobytes()


# ----------------------------------------------------------------- ABI wrappers

def test_g1_msm_wrapper_u64_vs_oracle():
    n = 300
    ks = o.gen_u64(77, n)
    bases = o.bases_from_k("g1", ks)
    bases[17]["inf"] = 1
    bases[17]["x"] = 0
    bases[17]["y"] = 0
    sc = o.gen_u64(78, n)
    sc[3] = 0
    sc[4] = 1
    std = np.zeros_like(bases)
    for i in range(n):
        std[i] = o.from_mont("g1", bases[i:i + 1])[0]
    res = gm.G1Projective.msm(std, sc)
    got = res.to_affine().rec
    want = o.msm("g1", bases, sc.reshape(-1, 1))
    assert same(got, want)


@pytest.mark.parametrize("curve,n", [("g1", 70001), ("g1", 300007), ("g2", 40009)])
def test_host_buffer_staging_sizes(curve, n):
    """Host-buffer calls of 8 MiB and more go through pinned staging slots
    split over worker threads: sizes that fill two workers, eight workers
    with a short last slot, and G2 records. The result equals the same MSM
    over resident copies of the same data."""
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    ks = o.gen_u64(130, n)
    ss = o.gen_scalars_modr(131, n)
    b = gm.Bases.from_k(cv, ks)
    s = gm.Scalars.upload(ss)
    ref = gm.msm_resident(b, s, scalar_bits=299)
    proj = gm.G1Projective if curve == "g1" else gm.G2Projective
    got = proj.msm_multi(b.download(0, n), ss, montgomery=True, scalars_lt_r=True).to_affine().rec
    assert same(got, ref)
    # uploads through the same staging: the bases and scalars round-trip
    b2 = gm.Bases.upload(cv, b.download(0, n))
    assert same(b2.download(0, n), b.download(0, n))
    s2 = gm.Scalars.upload(ss)
    assert same(gm.msm_resident(b2, s2, scalar_bits=299), ref)
    for h in (b, s, b2, s2):
        h.free()


def test_g2_msm_wrapper_u64_vs_oracle():
    n = 100
    ks = o.gen_u64(79, n)
    bases = o.bases_from_k("g2", ks)
    sc = o.gen_u64(80, n)
    res = gm.G2Projective.msm(bases, sc)  # G2 points taken as Montgomery form
    got = res.to_affine().rec
    want = o.msm("g2", bases, sc.reshape(-1, 1))
    assert same(got, want)


@pytest.mark.parametrize("curve,n", [("g1", 10), ("g1", 100), ("g2", 10)])
def test_reference_triangular_closed_form(golden, curve, n):
    """reference tests/test_msm.cu:132-592: bases G, scalars 1..N."""
    g = o.g1_generator(mont=(curve == "g2")) if curve == "g1" else o.g2_generator(mont=True)
    pts = np.repeat(g, n)
    sc = np.arange(1, n + 1, dtype=np.uint64)
    cls = gm.G1Projective if curve == "g1" else gm.G2Projective
    res = cls.msm(pts, sc).to_affine().rec
    assert as_point(curve, res[0]) == golden_point(curve, golden["triangular"][f"{curve}_{n}"])


def test_c1_host_config_g1_1024(golden):
    """C1: G1, N=1024, bases k_i*G, scalars in [0, r) with 8 zeros and 8 ones."""
    c = golden["c1_g1_1024"]
    n = c["n"]
    ks = o.gen_u64(c["seed_k"], n)
    ss = o.gen_scalars_modr(c["seed_s"], n)
    ss[c["zero_idx"]] = 0
    for i in c["one_idx"]:
        ss[i] = [1, 0, 0, 0, 0]
    bases = o.bases_from_k("g1", ks)
    res = gm.G1Projective.msm_multi(bases, ss, montgomery=True, scalars_lt_r=True).to_affine().rec
    assert as_point("g1", res[0]) == golden_point("g1", c["result"])
    assert same(res, o.msm("g1", bases, ss))


def test_edge_cases_g1():
    g = o.g1_generator()
    ng = g.copy()
    ng[0]["y"] = o.fp_unop("or_fp_neg", g[0]["y"].copy())
    two = o.scalar_mul("g1", g, 2)
    # P + (-P) = O; all-zero scalars = O; duplicate points take the doubling path
    pts = np.concatenate([g, ng])
    r = gm.G1Projective.msm_multi(pts, np.array([[5], [5]], dtype=np.uint64), montgomery=True).rec
    assert not r.view(np.uint64).any()
    r = gm.G1Projective.msm_multi(pts, np.zeros((2, 1), dtype=np.uint64), montgomery=True).rec
    assert not r.view(np.uint64).any()
    pts = np.concatenate([g, g, g, two])
    r = gm.G1Projective.msm_multi(pts, np.array([[1], [1], [1], [3]], dtype=np.uint64), montgomery=True)
    assert as_point("g1", r.to_affine().rec[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, 9)
    # n = 1, scalar r - 1 (-G)
    sc = np.array([o.int_to_limbs(m.R_ORDER - 1, 5)], dtype=np.uint64)
    r = gm.G1Projective.msm_multi(g, sc, montgomery=True, scalars_lt_r=True)
    assert as_point("g1", r.to_affine().rec[0]) == m.ec_neg(m.FP_OPS, m.G1_STD)
    # scalar bigger than the declared bound is rejected, not silently wrapped
    big = np.array([o.int_to_limbs(1 << 310, 5)], dtype=np.uint64)
    with pytest.raises(gm.MsmError):
        gm.G1Projective.msm_multi(g, big, montgomery=True, scalars_lt_r=True)
    # full 320-bit scalars are accepted without the flag
    r = gm.G1Projective.msm_multi(g, big, montgomery=True)
    assert as_point("g1", r.to_affine().rec[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, 1 << 310)


def test_edge_cases_g2_doubling_and_cancellation():
    """The G2 accumulation's own special branches (two lanes per Fp2 element):
    P + (-P) in one bucket gives O, equal points take the doubling path, and a
    bucket that returns to O mid-run continues from there."""
    g = o.g2_generator()
    ng = o.scalar_mul("g2", g, m.R_ORDER - 1)
    two = o.scalar_mul("g2", g, 2)
    pts = np.concatenate([g, ng])
    r = gm.G2Projective.msm_multi(pts, np.array([[5], [5]], dtype=np.uint64), montgomery=True).rec
    assert not r.view(np.uint64).any()
    pts = np.concatenate([g, g, g, two])
    r = gm.G2Projective.msm_multi(pts, np.array([[1], [1], [1], [3]], dtype=np.uint64), montgomery=True)
    assert as_point("g2", r.to_affine().rec[0]) == m.ec_mul(m.FP2_OPS, m.G2_STD, 9)
    # g - g cancels, then 2g + g + g: 4g per unit of scalar
    pts = np.concatenate([g, ng, two, g, g])
    r = gm.G2Projective.msm_multi(pts, np.full((5, 1), 7, dtype=np.uint64), montgomery=True)
    assert as_point("g2", r.to_affine().rec[0]) == m.ec_mul(m.FP2_OPS, m.G2_STD, 28)


def test_edge_cases_g2_infinity_and_zero():
    ks = o.gen_u64(81, 40)
    bases = o.bases_from_k("g2", ks)
    bases[0]["inf"] = 1
    bases[0]["x"] = 0
    bases[0]["y"] = 0
    ss = o.gen_scalars_modr(82, 40)
    ss[5] = 0
    r = gm.G2Projective.msm_multi(bases, ss, montgomery=True, scalars_lt_r=True).to_affine().rec
    assert same(r, o.msm("g2", bases, ss))


# -------------------------------------------------------------- resident engine

@pytest.fixture(scope="module")
def g1_pool_16(golden):
    w = golden["g1_2_16"]
    ks = o.gen_u64(w["seed_k"], w["n"])
    ss = o.gen_scalars_modr(w["seed_s"], w["n"])
    b = gm.Bases.from_k(gm.CURVE_G1, ks)
    s = gm.Scalars.upload(ss)
    return ks, ss, b, s


def test_generated_bases_match_oracle(g1_pool_16):
    ks, _, b, _ = g1_pool_16
    got = b.download(0, 512)
    want = o.bases_from_k("g1", ks[:512])
    assert same(got, want)
    got = b.download(60000, 64)
    assert same(got, o.bases_from_k("g1", ks[60000:60064]))


def test_g1_2_16_full_and_windows(golden, g1_pool_16):
    ks, ss, b, s = g1_pool_16
    w = golden["g1_2_16"]
    ref = gm.msm_resident(b, s, scalar_bits=299)
    assert as_point("g1", ref[0]) == golden_point("g1", w["result"])
    for c in (4, 9, 13, 16):
        assert same(gm.msm_resident(b, s, scalar_bits=299, window=c), ref), c
    assert same(gm.msm_resident(b, s, scalar_bits=299, device_combine=True), ref)
    # subset vs the oracle's own Pippenger
    sub = gm.msm_resident(b, s, n=1024, base_offset=4096, scalar_offset=4096, scalar_bits=299)
    assert same(sub, o.msm("g1", o.bases_from_k("g1", ks[4096:5120]), ss[4096:5120]))


@pytest.mark.parametrize("seg", [1, 3, 5, 6, 7, 11, 33, 100000])
def test_bucket_segments_any_length(g1_pool_16, seg):
    """Bucket-reduction segments of any length, including ones that do not
    divide 2^(c-1) (the window's last segment is shorter) and one longer than
    the window (capped at 2^(c-1)), give the same result as the default plan,
    for two windows."""
    ks, ss, b, s = g1_pool_16
    for c, n in ((9, 3000), (13, 65536)):
        ref = gm.msm_batch(b, s, [(n, 0, 0, c, 0)], scalar_bits=299)
        got = gm.msm_batch(b, s, [(n, 0, 0, c, seg)], scalar_bits=299)
        assert same(got, ref), (seg, c)


def test_g2_2_14_closed_form_with_zeros(golden):
    w = golden["g2_2_14"]
    n = w["n"]
    ks = o.gen_u64(w["seed_k"], n)
    ss = o.gen_scalars_modr(w["seed_s"], n)
    zs = []
    gz = m.MT64(w["zero_seed"])
    seen = set()
    while len(zs) < n // 100:
        v = gz.next() % n
        if v not in seen:
            seen.add(v)
            zs.append(v)
    ss[zs] = 0
    b = gm.Bases.from_k(gm.CURVE_G2, ks)
    s = gm.Scalars.upload(ss)
    assert same(b.download(0, 64), o.bases_from_k("g2", ks[:64]))
    r = gm.msm_resident(b, s, scalar_bits=299)
    assert as_point("g2", r[0]) == golden_point("g2", w["result"])
    sub = gm.msm_resident(b, s, n=256, base_offset=100, scalar_offset=100, scalar_bits=299)
    assert same(sub, o.msm("g2", o.bases_from_k("g2", ks[100:356]), ss[100:356]))


@pytest.mark.parametrize("curve,lg", [("g1", 18), ("g1", 20), ("g2", 16), ("g2", 20)])
def test_full_size_closed_form(golden, curve, lg):
    """Full-size configs: the closed form (sum s_i k_i mod r) * G pins the
    whole affine result; a 1024-point subset is recomputed by the oracle."""
    w = golden[f"{curve}_2_{lg}"]
    n = w["n"]
    ks = o.gen_u64(w["seed_k"], n)
    ss = o.gen_scalars_modr(w["seed_s"], n)
    if curve == "g2":
        zs, seen = [], set()
        gz = m.MT64(w["zero_seed"])
        while len(zs) < n // 100:
            v = gz.next() % n
            if v not in seen:
                seen.add(v)
                zs.append(v)
        ss[zs] = 0
    assert hex(o.limbs_to_int(ss[-1])) == w["s_last"]
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    b = gm.Bases.from_k(cv, ks)
    s = gm.Scalars.upload(ss)
    r = gm.msm_resident(b, s, scalar_bits=299)
    assert as_point(curve, r[0]) == golden_point(curve, w["result"])
    # the drop-in call with host buffers (copied through pinned staging by
    # several threads) gives the same point
    proj = gm.G1Projective if curve == "g1" else gm.G2Projective
    got = proj.msm_multi(b.download(0, n), ss, montgomery=True, scalars_lt_r=True).to_affine().rec
    assert same(got, r)
    off = n // 2 + 3
    sub = gm.msm_resident(b, s, n=1024, base_offset=off, scalar_offset=off, scalar_bits=299)
    assert same(sub, o.msm(curve, o.bases_from_k(curve, ks[off:off + 1024]), ss[off:off + 1024]))


# ------------------------------------------------------------------- batched

def _pool(curve, n, seed):
    ks = o.gen_u64(seed, n)
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    return ks, gm.Bases.from_k(cv, ks)


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_batch_matches_alone_and_oracle(curve):
    """Each MSM of a mixed-size batch equals the same MSM run alone, and the
    small ones equal the oracle."""
    pool_n = 1 << 14
    ks, b = _pool(curve, pool_n, 90 if curve == "g1" else 91)
    rng = np.random.default_rng(5)
    sizes = [1, 2, 7, 64, 300, 1000, 4096, 1 << 13, 17, 513]
    total = sum(sizes)
    ss = o.gen_scalars_modr(92, total)
    ss[3] = 0
    s = gm.Scalars.upload(ss)
    specs, so = [], 0
    for k, n in enumerate(sizes):
        bo = int(rng.integers(0, pool_n - n + 1))
        window = [0, 5, 0, 9, 0, 12, 0, 0, 3, 0][k]
        specs.append((n, bo, so, window))
        so += n
    got = gm.msm_batch(b, s, specs, scalar_bits=299)
    got_dev = gm.msm_batch(b, s, specs, scalar_bits=299, device_combine=True)
    assert same(got, got_dev)
    for k, (n, bo, so_, w) in enumerate(specs):
        alone = gm.msm_resident(b, s, n=n, base_offset=bo, scalar_offset=so_, scalar_bits=299, window=w)
        assert same(got[k:k + 1], alone), k
        if n <= 1000:
            want = o.msm(curve, o.bases_from_k(curve, ks[bo:bo + n]), ss[so_:so_ + n])
            assert same(got[k:k + 1], want), k


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_reduction_and_window_sum_special_cases(curve):
    """Equal and opposite points meet inside the bucket reduction's running
    sums and between a window's segment results, so the additions there take
    their doubling and infinity branches (on the split lane groups: G1 lane
    pairs, G2 pairs of Fp2 lane pairs). Window 4 (8 buckets, two segments of
    4), scalars below 16, points
    Q, 5Q, 4Q and -Q; each case against the oracle."""
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    k0 = 987654321
    base = o.bases_from_k(curve, np.array([k0, 5 * k0, 4 * k0], dtype=np.uint64))
    neg = base[0:1].copy()
    y = neg[0]["y"]
    for h in range(len(y) // 7):  # -y, component by component
        y[7 * h:7 * h + 7] = o.fp_unop("or_fp_neg", y[7 * h:7 * h + 7].copy())
    pts = np.concatenate([base, neg])  # Q, 5Q, 4Q, -Q
    cases = [
        ([0, 0], [2, 1]),  # S = Q + Q in the running sum (doubling)
        ([0], [2]),        # T = Q + Q after an empty bucket (doubling)
        ([0, 3], [2, 1]),  # S = Q + (-Q) (infinity), result Q
        ([1, 2], [4, 5]),  # segment results 20Q and 20Q (doubling in the window sum)
        ([1, 2], [4, 5 + 16]),  # the same in the low window, 16 * 4Q in the next
    ]
    for idx, sc in cases:
        sel = pts[idx]
        scal = np.array([[v, 0, 0, 0, 0] for v in sc], dtype=np.uint64)
        s = gm.Scalars.upload(scal)
        bb = gm.Bases.upload(cv, sel)
        got = gm.msm_batch(bb, s, [(len(idx), 0, 0, 4, 4)], scalar_bits=299)
        assert same(got, o.msm(curve, sel, scal)), (idx, sc)
        bb.free()
        s.free()


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_tail_split_and_unsplit_agree(curve):
    """The bucket reduction and window sums split each point operation's
    products over two lane groups (G1: two lanes; G2: two lane pairs) while
    the plan's segments fit one round of waves, and run unsplit past it; the
    fix-up splits while its split launch fits 256K threads. One MSM alone
    (25 windows x 512 segments of 4 buckets: 25,600 G1 or 51,200 G2 threads;
    51,200 buckets cut into chunks of 4 entries) takes the split kernels; the
    same MSM in a batch of 64 (819,200 segments, 3.3M buckets) the unsplit
    ones. Both equal the oracle."""
    pool_n = 1 << 14
    ks, b = _pool(curve, pool_n, 120)
    ss = o.gen_scalars_modr(121, pool_n)
    s = gm.Scalars.upload(ss)
    alone = gm.msm_batch(b, s, [(pool_n, 0, 0, 12, 4)], scalar_bits=299)
    many = gm.msm_batch(b, s, [(pool_n, 0, 0, 12, 4)] * 64, scalar_bits=299)
    for k in (0, 31, 63):
        assert same(many[k:k + 1], alone), k
    assert same(alone, o.msm(curve, o.bases_from_k(curve, ks), ss))


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_random_plans_against_oracle(curve):
    """Seeded random MSMs (sizes 1..3000, windows 2..13, bucket segments of
    any length or the plan's own), in one batch and each alone, against the
    oracle: small windows and short segments reach the direct window sums, a
    window of one segment and the last short segment."""
    rng = np.random.default_rng(140 if curve == "g1" else 141)
    pool_n = 4096
    ks, b = _pool(curve, pool_n, 142)
    ss = o.gen_scalars_modr(143, pool_n)
    s = gm.Scalars.upload(ss)
    specs = []
    for _ in range(8):
        n = int(rng.integers(1, 3001))
        bo = int(rng.integers(0, pool_n - n + 1))
        so = int(rng.integers(0, pool_n - n + 1))
        c = int(rng.integers(2, 14))
        seg = int(rng.choice([0, 1, 2, 3, 5, 7, 9, 16, 40]))
        specs.append((n, bo, so, c, seg))
    got = gm.msm_batch(b, s, specs, scalar_bits=299)
    for k, (n, bo, so, c, seg) in enumerate(specs):
        want = o.msm(curve, o.bases_from_k(curve, ks[bo:bo + n]), ss[so:so + n])
        assert same(got[k:k + 1], want), specs[k]
        alone = gm.msm_batch(b, s, [specs[k]], scalar_bits=299)
        assert same(alone, want), specs[k]
    s.free()


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_batch_many_windows_direct_window_sums(curve):
    """A batch of over 1024 windows sums windows of few segments with one
    lane each, and the rest with tiles; every MSM equals the same MSM run
    alone (which takes the tile path) and, for a sample, the oracle."""
    pool_n = 1 << 13
    ks, b = _pool(curve, pool_n, 110 if curve == "g1" else 111)
    rng = np.random.default_rng(12)
    sizes = [300] * 40 + [4096, 50, 2000]
    windows = [6] * 40 + [12, 5, 10]
    ss = o.gen_scalars_modr(112, sum(sizes))
    s = gm.Scalars.upload(ss)
    specs, so = [], 0
    for n, w in zip(sizes, windows):
        specs.append((n, int(rng.integers(0, pool_n - n + 1)), so, w))
        so += n
    got = gm.msm_batch(b, s, specs, scalar_bits=299)
    for k in (0, 17, 39, 40, 41, 42):
        n, bo, so_, w = specs[k]
        alone = gm.msm_resident(b, s, n=n, base_offset=bo, scalar_offset=so_, scalar_bits=299, window=w)
        assert same(got[k:k + 1], alone), k
    for k in (3, 41):
        n, bo, so_, _ = specs[k]
        want = o.msm(curve, o.bases_from_k(curve, ks[bo:bo + n]), ss[so_:so_ + n])
        assert same(got[k:k + 1], want), k


def test_batch_u64_scalars_and_reorder():
    """Permuting the MSMs of a batch does not change any result."""
    pool_n = 4096
    _, b = _pool("g1", pool_n, 93)
    sc = o.gen_u64(94, 6000).reshape(-1, 1)
    s = gm.Scalars.upload(sc)
    specs = [(100, 0, 0), (2000, 50, 100), (5, 4000, 2100), (3000, 1000, 2105), (895, 3000, 5105)]
    r1 = gm.msm_batch(b, s, specs)
    perm = [3, 0, 4, 2, 1]
    r2 = gm.msm_batch(b, s, [specs[k] for k in perm])
    for a, k in enumerate(perm):
        assert same(r2[a:a + 1], r1[k:k + 1])


def test_batch_graph_replays_match_eager():
    """A captured hipGraph of the batch pipeline gives the eager result on
    every replay, with eager work and synchronisation between replays."""
    pool_n = 1 << 13
    _, b = _pool("g1", pool_n, 95)
    sizes = [3000, 17, 900, 4096, 1]
    ss = o.gen_scalars_modr(96, sum(sizes))
    s = gm.Scalars.upload(ss)
    specs, so = [], 0
    for k, n in enumerate(sizes):
        specs.append((n, (k * 997) % (pool_n - n + 1), so))
        so += n
    eager = gm.msm_batch(b, s, specs, scalar_bits=299)
    bt = gm.Batch(b, s, specs, scalar_bits=299, use_graph=True)
    nog = gm.Batch(b, s, specs, scalar_bits=299, use_graph=False)
    dev = gm.Batch(b, s, specs, scalar_bits=299, use_graph=True, device_combine=True)
    for rep in range(4):
        assert same(bt.run(), eager), rep
        assert same(nog.run(), eager), rep
        assert same(dev.run(), eager), rep
        gm.msm_resident(b, s, n=100, scalar_bits=299)  # eager work between replays
    # launch/wait split and a second batch overlapping on the device
    bt.launch()
    nog.launch()
    assert same(bt.wait(), eager)
    assert same(nog.wait(), eager)
    bt.free()
    nog.free()
    dev.free()


def test_heavy_buckets_repeated_scalars_and_tiny_top_window():
    """All scalars equal (every window's entries in one bucket) and a window
    width whose top window has few bits: the chunk partials of those buckets
    go through the parallel heavy-bucket fix-up. Closed form: s * sum(k_i) * G."""
    n = 1 << 14
    ks = o.gen_u64(97, n)
    b = gm.Bases.from_k(gm.CURVE_G1, ks)
    s_val = m.gen_scalars_modr(98, 1)[0]
    sc = np.tile(o.int_to_limbs(s_val, 5), (n, 1))
    s = gm.Scalars.upload(sc)
    t = s_val * sum(int(k) for k in ks) % m.R_ORDER
    want = m.ec_mul(m.FP_OPS, m.G1_STD, t)
    for c in (0, 14, 16):
        r = gm.msm_resident(b, s, scalar_bits=299, window=c)
        assert as_point("g1", r[0]) == want, c
    # one bucket of 2^15 entries per window: its block of 128 buckets holds
    # more entries than the placement's LDS stage and is placed directly
    n2 = 1 << 15
    ks2 = o.gen_u64(96, n2)
    b2 = gm.Bases.from_k(gm.CURVE_G1, ks2)
    sr2 = gm.Scalars.upload(np.tile(o.int_to_limbs(s_val, 5), (n2, 1)))
    t2 = s_val * sum(int(k) for k in ks2) % m.R_ORDER
    r2 = gm.msm_resident(b2, sr2, scalar_bits=299, window=12)
    assert as_point("g1", r2[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, t2)
    # tiny top window with random scalars, against the oracle on a subset
    ss = o.gen_scalars_modr(99, n)
    s2 = gm.Scalars.upload(ss)
    full = gm.msm_resident(b, s2, scalar_bits=299, window=14)
    assert same(full, gm.msm_resident(b, s2, scalar_bits=299, window=11))


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_equal_and_cancelling_chunk_partials(curve):
    """Every base the same point and every scalar the same: each window's
    entries fill one bucket, every chunk partial is the same multiple of P, and
    the fix-up's sums double (heavy-bucket tree and serial fix-up alike).
    Then P and -P alternate under equal scalars: every even chunk's partial is
    O, and the MSM is O."""
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    ops, gen = (m.FP_OPS, m.G1_STD) if curve == "g1" else (m.FP2_OPS, m.G2_STD)
    n, k = 1 << 13, 12345
    b = gm.Bases.from_k(cv, np.full(n, k, dtype=np.uint64))
    s_val = m.gen_scalars_modr(77, 1)[0]
    s = gm.Scalars.upload(np.tile(o.int_to_limbs(s_val, 5), (n, 1)))
    want = m.ec_mul(ops, gen, n * k * s_val % m.R_ORDER)
    for c in (0, 9, 13):
        r = gm.msm_resident(b, s, scalar_bits=299, window=c)
        assert as_point(curve, r[0]) == want, c
    g = o.g1_generator() if curve == "g1" else o.g2_generator()
    p = o.scalar_mul(curve, g, k)
    pts = np.concatenate([p, o.scalar_mul(curve, p, m.R_ORDER - 1)] * (n // 2))
    proj = gm.G1Projective if curve == "g1" else gm.G2Projective
    r = proj.msm_multi(pts, np.full((n, 1), 0x9E3779B97F4A7C15, dtype=np.uint64), montgomery=True).rec
    assert not r.view(np.uint64).any()


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_global_sort_path_wide_windows(curve):
    """Windows wider than 16 bits place entries with one global atomic per
    entry (the sort's fallback path) instead of LDS tile counters: windows of
    17 and 20 bits on scalars below r, 24 bits on u64 scalars (3 windows of
    2^23 buckets), alone and in a batch that mixes c = 12 with c = 18 (the
    whole batch then takes the global path), against the oracle."""
    n = 1500 if curve == "g1" else 500
    ks, b = _pool(curve, n, 140 if curve == "g1" else 141)
    bases = o.bases_from_k(curve, ks)
    ss = o.gen_scalars_modr(142, n)
    s = gm.Scalars.upload(ss)
    want = o.msm(curve, bases, ss)
    for c in (17, 20):
        print(f"{curve} n={n} window={c}", flush=True)
        assert same(gm.msm_resident(b, s, scalar_bits=299, window=c), want), c
    su = o.gen_u64(143, n)
    s64 = gm.Scalars.upload(su.reshape(-1, 1))
    print(f"{curve} n={n} window=24 u64 scalars", flush=True)
    assert same(gm.msm_resident(b, s64, scalar_bits=64, window=24), o.msm(curve, bases, su.reshape(-1, 1)))
    specs = [(n // 2, 0, 0, 12), (n - n // 2, n // 2, n // 2, 18), (n // 3, n // 5, 0, 12)]
    print(f"{curve} batch of windows 12 / 18 / 12", flush=True)
    got = gm.msm_batch(b, s, specs, scalar_bits=299)
    for k, (m_, bo, so, _) in enumerate(specs):
        w = o.msm(curve, bases[bo:bo + m_], ss[so:so + m_])
        assert same(got[k:k + 1], w), k


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_montgomery_upload_stores_canonical_coordinates(curve):
    """Bases uploaded in Montgomery form are stored canonical, below p (the
    accumulation negates a base's y as p - y, which needs it): coordinates
    given as x + k p (k = 1..3, kept below 2^448) download as x and give the
    same MSM, bit for bit, as the canonical upload."""
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    ks = o.gen_u64(77, 256)
    ref = gm.Bases.from_k(cv, ks)
    recs = ref.download(0, 256)
    big = recs.copy()
    rng = np.random.default_rng(3)
    lifted = 0
    for i in range(len(big)):
        for f in ("x", "y"):
            w = big[f][i]
            for e in range(len(w) // 7):
                v = sum(int(w[7 * e + j]) << (64 * j) for j in range(7))
                v2 = v + int(rng.integers(1, 4)) * m.P
                if v2 < 1 << 448:
                    w[7 * e:7 * e + 7] = [(v2 >> (64 * j)) & ((1 << 64) - 1) for j in range(7)]
                    lifted += 1
    assert lifted > 256
    up = gm.Bases.upload(cv, big, montgomery=True)
    assert same(up.download(0, 256), recs)
    s = gm.Scalars.upload(o.gen_scalars_modr(78, 256))
    assert same(gm.msm_resident(up, s, scalar_bits=299), gm.msm_resident(ref, s, scalar_bits=299))
    up.free()
    ref.free()
