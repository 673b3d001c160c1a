"""Pins the CPU oracle (oracle/lda_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something it does not share code with:

  philox            Random123 known-answer vectors (SC'11 paper's reference implementation) and
                    64 blocks from rocRAND's open-source engine (tests/golden/philox_rocrand.json)
  orc_log/sqrt/     libm (math.log / math.sqrt / math.lgamma) within a few ulp
  log_factorial
  scan64 / sum64    exact integer inputs (any association gives the exact sum) and the reference's
                    compute_product_cumsum known answer (src/cu/test/test_topics.cu:191-237)
  alias_build       exact PMF reconstruction from (prob, alias) on random, sparse, one-hot,
                    uniform and the reference's test_build_alias rows (test_spalias.cu:14-53)
  poisson tables    scipy.stats.poisson.pmf in double; table PMF reconstruction
  poisson (table,   chi-square goodness of fit against scipy's PMF; mean/variance bands of
  PTRS)             test_polyaurn.cu:16-74
  map               Robin Hood invariants, counts vs a dict, history independence, and the
                    reference's count_topics / hashmap known answers (test_topics.cu:134-189,
                    test_hashmap.cu:224-266)
  sweep             exact two-token joint conditional (enumerated) by chi-square; a one-topic
                    phi row forces every z to that topic (restates test_topics.cu:276-334)
  count             numpy bincount
  loglik            closed form with scipy.special.gammaln
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import special, stats

HERE = os.path.dirname(os.path.abspath(__file__))

RANDOM123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_known_answers(orc):
    for c, k, r in RANDOM123_KAT:
        assert tuple(int(x) for x in orc.philox([c], [k])[0]) == r


def test_philox_matches_rocrand_vectors(orc):
    vec = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    ctr = np.array([v[0] for v in vec], dtype=np.uint32)
    key = np.array([v[1] for v in vec], dtype=np.uint32)
    exp = np.array([v[2] for v in vec], dtype=np.uint32)
    assert np.array_equal(orc.philox(ctr, key), exp)


def test_log_sqrt_lfact_against_libm(orc):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 2000)), rng.uniform(0.5, 2.0, 2000), [1.0, 2.0, 5e-324]])
    for x in xs:
        a, b = orc.lib().orc_log(float(x)), math.log(float(x))
        assert abs(a - b) <= 4 * math.ulp(b) + 1e-300, (x, a, b)
    for x in np.concatenate([rng.uniform(1, 1e9, 2000), [1.0, 100.01, 4e9]]):
        a, b = orc.lib().orc_sqrt(float(x)), math.sqrt(float(x))
        assert abs(a - b) <= 2 * math.ulp(b), (x, a, b)
    for k in list(range(0, 40)) + [100, 1000, 12345, 10**6, 4 * 10**9]:
        a, b = orc.lib().orc_log_factorial(k), math.lgamma(k + 1)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), (k, a, b)


def test_scan64_exact_on_integers(orc):
    rng = np.random.default_rng(1)
    x = rng.integers(0, 1000, size=(50, 64)).astype(np.float32)
    x[3] = 0
    x[7, ::3] = 0
    out = orc.scan64(x)
    assert np.array_equal(out, np.cumsum(x.astype(np.float64), axis=1).astype(np.float32))
    assert orc.sum64(x.reshape(-1)) == np.float32(x.sum())


def test_scan64_close_on_reals(orc):
    rng = np.random.default_rng(2)
    x = rng.random((20, 64)).astype(np.float32) * np.float32(1e-3)
    out = orc.scan64(x)
    ref = np.cumsum(x.astype(np.float64), axis=1)
    assert np.allclose(out, ref, rtol=1e-6, atol=0)
    assert np.all(np.diff(out, axis=1) >= 0)


def test_product_cumsum_known_answer(orc):
    # test_topics.cu:191-237: Phi = 6i, count = i over 100 slots; exclusive prefix at i is
    # (i-1) i (2i-1) and the total 99*100*199.  Here as a 128-slot inclusive scan over lines.
    v = np.zeros(128, dtype=np.float32)
    i = np.arange(100)
    v[:100] = (6.0 * i) * i
    lines = orc.scan64(v)
    pre = np.concatenate([lines[0], lines[0][-1] + lines[1]])
    excl = np.concatenate([[0.0], pre[:-1]])
    expect = np.array([(k - 1) * k * (2 * k - 1) if k > 0 else 0 for k in range(100)], dtype=np.float64)
    assert np.array_equal(excl[:100], expect)
    assert pre[99] == 99 * 100 * 199


def reconstruct(prob, al):
    m = len(prob)
    p = prob.astype(np.float64).copy()
    for i in range(m):
        if al[i] != i:
            p[al[i]] += 1.0 - float(prob[i])
    return p / m


def _alias_ok(orc, w):
    w = np.asarray(w, dtype=np.float32)
    prob, al = orc.alias_build(w)
    m = len(w)
    assert np.all((al >= 0) & (al < m))
    assert np.all((prob >= 0) & (prob <= 1))
    total = float(orc.sum64(w))
    if total <= 0:
        assert np.all(prob == 1) and np.array_equal(al, np.arange(m))
        return
    target = w.astype(np.float64) / total
    rec = reconstruct(prob, al)
    # f32 rounding of each stored probability contributes at most 2^-24 per slot it touches
    touch = np.bincount(al, minlength=m) + 1
    assert np.all(np.abs(rec - target) <= touch * 2.0**-23 / m + 1e-12), np.abs(rec - target).max()
    # zero-weight outcomes are never produced
    reach = (prob > 0) | (np.arange(m) != al)
    produced = np.zeros(m, bool)
    produced[np.arange(m)[prob > 0]] = True
    produced[al[prob < 1]] = True
    assert not np.any(produced & (w == 0)) or np.all(w == 0)
    assert reach.shape == (m,)


@pytest.mark.parametrize("m", [1, 2, 5, 20, 64, 65, 200, 1024, 3000])
def test_alias_reconstruction_random(orc, m):
    rng = np.random.default_rng(m)
    _alias_ok(orc, rng.random(m))
    w = rng.random(m) * (rng.random(m) < 0.1)
    _alias_ok(orc, w)
    _alias_ok(orc, np.ones(m))
    one = np.zeros(m)
    one[m // 2] = 3.0
    _alias_ok(orc, one)
    _alias_ok(orc, np.exp(rng.normal(0, 5, m)))


def test_alias_reference_case_and_zero_row(orc):
    # test_spalias.cu:26-46 input (the reference's unscaled convention is not reproduced; the
    # PMF it encodes is)
    _alias_ok(orc, [0.9, 0.02] + [0.01] * 8)
    prob, al = orc.alias_build(np.zeros(7, np.float32))
    assert np.all(prob == 1) and np.array_equal(al, np.arange(7))


@pytest.mark.parametrize("beta", [0.01, 0.1])
def test_poisson_tables_pmf(orc, beta):
    prob, al, _ = orc.poisson_tables(beta)
    for n in [0, 1, 2, 10, 50, 99]:
        lam = np.float64(n) + np.float64(np.float32(beta))
        pmf = stats.poisson.pmf(np.arange(200), lam)
        rec = reconstruct(prob[n], al[n])
        assert np.abs(rec - pmf).max() < 5e-7, (n, np.abs(rec - pmf).max())


def _chisq_counts(draws, pmf_fn, lo, hi):
    vals = np.arange(lo, hi + 1)
    expected = pmf_fn(vals) * len(draws)
    obs = np.array([(draws == v).sum() for v in vals], dtype=np.float64)
    # merge tails into the bins below / above
    lo_tail = (draws < lo).sum()
    hi_tail = (draws > hi).sum()
    cdf_lo = 1.0 - expected.sum() / len(draws)
    keep = expected >= 5
    e = np.concatenate([expected[keep], [max(len(draws) - expected[keep].sum(), 1e-9)]])
    o = np.concatenate([obs[keep], [len(draws) - obs[keep].sum()]])
    del lo_tail, hi_tail, cdf_lo
    return stats.chisquare(o, e).pvalue


@pytest.mark.parametrize("n", [0, 1, 10, 99, 100, 1000, 10000])
def test_poisson_draws_chisquare(orc, n):
    beta = 0.01
    tables = orc.poisson_tables(beta)[2]
    N = 20000
    lib = orc.lib()
    import ctypes as C
    draws = np.array([lib.orc_poisson(n, C.c_float(beta), tables, i, 7, 1, 3, 99) for i in range(N)])
    lam = n + float(np.float32(beta))
    sd = math.sqrt(lam)
    lo, hi = max(0, int(lam - 5 * sd) - 1), int(lam + 5 * sd) + 2
    p = _chisq_counts(draws, lambda v: stats.poisson.pmf(v, lam), lo, hi)
    assert p > 1e-4, p
    # mean / variance bands of test_polyaurn.cu:16-74 (there with 10^4 draws at +-10%)
    if n >= 10:
        assert abs(draws.mean() - lam) < 0.05 * lam
        assert abs(draws.var() - lam) < 0.1 * lam


def _check_map(T, cap, K, counts: dict, orc):
    hm = (2**32 - 1) // K  # the home multiplier: floor((2^32 - 1) / K)
    assert orc.map_home_mult(K) == hm
    home = lambda k: ((k * hm) * cap) >> 32  # noqa: E731
    assert all(k * hm < 2**32 for k in range(K))  # no wrap: homes are non-decreasing in the key
    # every empty slot holds count 0 at key 0xFFFF, past any row of K < 65536 topics
    empty = lambda e: e & 0xFFFF == 0  # noqa: E731
    seen = {}
    for s in range(cap):
        e = int(T[s])
        if empty(e):
            assert e == 0xFFFF0000
            continue
        k, c = e >> 16, e & 0xFFFF
        assert k not in seen
        seen[k] = c
        h = home(k)
        d = (s - h) % cap
        for j in range(d):  # every slot between home and position is occupied
            assert not empty(int(T[(h + j) % cap]))
        if d > 0:  # Robin Hood order against the previous slot
            p = int(T[(s - 1) % cap])
            pk = p >> 16
            pd = (s - 1 - home(pk)) % cap
            assert d <= pd + 1
            if d == pd + 1:
                assert pk < k
    assert seen == {k: c for k, c in counts.items() if c > 0}


def test_map_closed_form_layout(orc):
    """The ordered layout the sweep kernels build in two scans over the per-topic counts: keys
    ascending, key i at slot (i + max(w, max_{j<=i}(h_j - j))) mod cap, h the home slot and w the
    number of slots the run passing the table's end occupies at its start.  It must equal the
    oracle's table built by sequential Robin Hood inserts, for key sets of every load (caps that a
    regrow or a shrink uses included)."""
    rng = np.random.default_rng(11)
    for trial in range(1500):
        K = int(rng.choice([7, 20, 64, 300, 1024, 4096]))
        n = int(rng.integers(1, K + 1))
        cap = orc.map_capacity(n, K)
        if rng.random() < 0.3:
            cap = orc.map_capacity(int(rng.integers(n, K + 1)), K)
        keys = np.sort(rng.choice(K, n, replace=False)).astype(np.int64)
        hm = orc.map_home_mult(K)
        h = ((keys * hm) * cap) >> 32
        i = np.arange(n)
        M = np.maximum.accumulate(h - i)
        w = max(0, int(n - 1 + M[-1]) + 1 - cap)
        pos = (i + np.maximum(w, M)) % cap
        assert len(set(pos.tolist())) == n
        want = np.full(cap, 0xFFFF0000, np.uint32)
        want[pos] = (keys.astype(np.uint32) << 16) | 1
        got = orc.map_apply(rng.permutation(keys).astype(np.uint32), cap, K)
        assert np.array_equal(got, want), (K, n, cap)


@pytest.mark.parametrize("K,L", [(20, 50), (1024, 600), (300, 64), (4096, 5000)])
def test_map_random_ops(orc, K, L):
    rng = np.random.default_rng(K + L)
    cap = orc.map_capacity(L, K)
    counts = {}
    ops = []
    total = 0
    for _ in range(6 * L):
        if total > 0 and (total >= L or rng.random() < 0.45):
            keys = [k for k, c in counts.items() if c > 0]
            k = int(rng.choice(keys))
            ops.append(k | (1 << 31))
            counts[k] -= 1
            total -= 1
        else:
            k = int(rng.integers(0, K))
            ops.append(k)
            counts[k] = counts.get(k, 0) + 1
            total += 1
    T = orc.map_apply(ops, cap, K)
    _check_map(T, cap, K, counts, orc)
    # history independence: the final multiset inserted in two other orders gives the same table
    final = [k for k, c in counts.items() for _ in range(c)]
    for seed in (1, 2):
        order = list(np.random.default_rng(seed).permutation(final)) if final else []
        assert np.array_equal(orc.map_apply(order, cap, K), T)
    # order-preserving homes and Robin Hood order make the table sorted by key, read from the
    # slot after the last wrapped run (at most one descent in the cyclic sequence of keys)
    keys = [int(e) >> 16 for e in T if int(e) & 0xFFFF]
    descents = sum(1 for i in range(len(keys)) if keys[i] > keys[(i + 1) % len(keys)])
    assert len(keys) < 2 or descents == 1


def test_map_reference_known_answers(orc):
    # count_topics (test_topics.cu:134-189): 100 tokens, z = i % 25 -> each count 4
    cap = orc.map_capacity(100, 25)
    T = orc.map_apply([i % 25 for i in range(100)], cap, 25)
    got = {int(e) >> 16: int(e) & 0xFFFF for e in T if int(e) & 0xFFFF}
    assert got == {k: 4 for k in range(25)}
    # test_hashmap.cu:224-266: 90 inserts over 9 keys -> every count is 10
    cap = orc.map_capacity(90, 9)
    T = orc.map_apply([i % 9 for i in range(90)], cap, 9)
    got = {int(e) >> 16: int(e) & 0xFFFF for e in T if int(e) & 0xFFFF}
    assert got == {k: 10 for k in range(9)}
    assert orc.map_capacity(100, 1024) == 128 and orc.map_capacity(50, 20) == 64
    # load factor <= 32/33, rounded up to whole lines
    assert orc.map_capacity(5000, 4096) == 64 * math.ceil((4096 + 128 + 1) / 64)
    assert orc.map_capacity(600, 1024) == 640 and orc.map_capacity(4096, 1024) == 1088


def _two_token_joint(phi, alpha, w, z):
    """Exact distribution of (z0', z1') after one sweep of a two-token document."""
    K = phi.shape[1]
    out = np.zeros((K, K))
    p0 = phi[w[0]] * (alpha + (np.arange(K) == z[1]))
    p0 = p0 / p0.sum()
    for a in range(K):
        p1 = phi[w[1]] * (alpha + (np.arange(K) == a))
        out[a] = p0[a] * p1 / p1.sum()
    return out


def _tiny_model(orc, K=4, V=3, seed=5):
    rng = np.random.default_rng(seed)
    phi_t = rng.integers(0, 5, size=(V, K)).astype(np.float32)
    phi_t[:, 0] += 1
    ts = phi_t.sum(axis=0).astype(np.uint64)
    alpha = 0.3
    phi, sig, prob, al = orc.phi_finish(phi_t, ts, alpha)
    return phi, sig, prob, al, np.float32(alpha)


def test_sweep_exact_conditional(orc):
    phi, sig, prob, al, alpha = _tiny_model(orc)
    K = phi.shape[1]
    w = np.array([1, 2], np.uint32)
    z0 = np.array([2, 3], np.uint32)
    n_docs = 40000
    W = np.tile(w, n_docs)
    Z = np.tile(z0, n_docs)
    off = np.arange(0, 2 * n_docs + 1, 2, dtype=np.uint64)
    znew, _ = orc.sweep(Z, W, off, 0, 1, 11, K, phi, sig, prob, al, threads=4)
    pairs = znew.reshape(-1, 2)
    obs = np.bincount(pairs[:, 0] * K + pairs[:, 1], minlength=K * K).astype(float)
    exp = _two_token_joint(phi.astype(np.float64), float(alpha), w, z0).reshape(-1) * n_docs
    keep = exp > 0
    assert obs[~keep].sum() == 0
    assert stats.chisquare(obs[keep], exp[keep]).pvalue > 1e-4


def test_sweep_single_topic_row_forces_topic(orc):
    # restates test_topics.cu:276-334: V=3, K=5, docs of length 3 and 2, all w = 0; phi_0 has
    # mass on topic 0 only, so every draw must return 0.
    V, K = 3, 5
    phi_t = np.zeros((V, K), np.float32)
    phi_t[0, 0] = 7
    phi_t[1:, 1:] = 1
    phi, sig, prob, al = orc.phi_finish(phi_t, phi_t.sum(0).astype(np.uint64), 0.1)
    z = np.array([4, 1, 0, 4, 0], np.uint32)
    w = np.zeros(5, np.uint32)
    off = np.array([0, 3, 5], np.uint64)
    znew, _ = orc.sweep(z, w, off, 0, 1, 0, K, phi, sig, prob, al)
    assert np.all(znew == 0)


def test_sweep_zero_row_keeps_topic(orc):
    V, K = 2, 4
    phi_t = np.zeros((V, K), np.float32)
    phi_t[1] = [1, 2, 3, 4]
    phi, sig, prob, al = orc.phi_finish(phi_t, np.array([1, 2, 3, 4], np.uint64), 0.1)
    assert sig[0] == 0
    z = np.array([3, 2, 1], np.uint32)
    w = np.array([0, 0, 1], np.uint32)
    znew, _ = orc.sweep(z, w, np.array([0, 3], np.uint64), 0, 1, 0, K, phi, sig, prob, al)
    assert znew[0] == 3 and znew[1] == 2


def test_count_matches_bincount(orc):
    rng = np.random.default_rng(3)
    V, K, N = 50, 7, 5000
    w = rng.integers(0, V, N).astype(np.uint32)
    z = rng.integers(0, K, N).astype(np.uint32)
    n = orc.count(w, z, V, K)
    assert np.array_equal(n.reshape(-1), np.bincount(w.astype(np.int64) * K + z, minlength=V * K))


def test_loglik_closed_form(orc):
    rng = np.random.default_rng(4)
    V, K = 6, 3
    alpha, beta = 0.2, 0.05
    d_len = np.array([3, 5, 1])
    off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.uint64)
    N = int(off[-1])
    w = rng.integers(0, V, N).astype(np.uint32)
    z = rng.integers(0, K, N).astype(np.uint32)
    n = orc.count(w, z, V, K)
    ll = special.gammaln(V * beta) * K - special.gammaln(V * beta + n.sum(0)).sum()
    ll += (special.gammaln(beta + n) - special.gammaln(beta)).sum()
    for d in range(3):
        zd = z[off[d]:off[d + 1]]
        ndk = np.bincount(zd, minlength=K)
        ll += special.gammaln(K * alpha) - special.gammaln(K * alpha + len(zd))
        ll += (special.gammaln(alpha + ndk) - special.gammaln(alpha)).sum()
    assert abs(orc.loglik(n, z, off, alpha, beta) - ll) < 1e-9 * abs(ll)


def test_phi_step_properties(orc):
    rng = np.random.default_rng(6)
    V, K = 40, 6
    n = rng.integers(0, 150, size=(V, K)).astype(np.uint32)
    n[::7] = 0
    pt, ts = orc.phi_draw(n, 0.01, 3, 5)
    assert np.all(pt == np.floor(pt)) and np.array_equal(ts, pt.astype(np.uint64).sum(0))
    # same counters -> same draws; different iteration -> different draws
    pt2, _ = orc.phi_draw(n, 0.01, 3, 5)
    pt3, _ = orc.phi_draw(n, 0.01, 4, 5)
    assert np.array_equal(pt, pt2) and not np.array_equal(pt, pt3)
    phi, sig, prob, al = orc.phi_finish(pt, ts, 0.5)
    colsum = phi.astype(np.float64).sum(0)
    assert np.allclose(colsum[ts > 0], 1.0, atol=1e-5)
    assert np.allclose(sig, 0.5 * phi.astype(np.float64).sum(1), rtol=1e-5)


def test_sweep_growth_known_answer(orc):
    # Every token starts at topic 0 (one distinct topic: a one-line map); word i has all its phi
    # mass on topic 3i+1, so after one sweep token i must hold topic 3i+1 -- 300 distinct
    # topics, reached only through repeated map growth.
    V, K, L = 300, 1024, 300
    phi_t = np.zeros((V, K), np.float32)
    for v in range(V):
        phi_t[v, (3 * v + 1) % K] = 5
    phi, sig, prob, al = orc.phi_finish(phi_t, np.maximum(phi_t.sum(0), 1).astype(np.uint64), 0.05)
    w = np.arange(L, dtype=np.uint32)
    z = np.zeros(L, np.uint32)
    znew, visited = orc.sweep(z, w, np.array([0, L], np.uint64), 0, 1, 3, K, phi, sig, prob, al)
    assert np.array_equal(znew, (3 * w + 1) % K)
    assert visited > 0


def test_foldin_exact_conditional(orc):
    """One fold-in sweep of two-token requests: z0, z1 start independently from p(k | w) (the
    alias tables), then z0' | z1 and z1' | z0' follow the full conditionals.  The joint law of
    (z0', z1') is enumerated exactly and compared by chi-square."""
    phi, sig, prob, al, alpha = _tiny_model(orc)
    K = phi.shape[1]
    ph = phi.astype(np.float64)
    w = np.array([1, 2], np.uint32)
    n_req = 40000
    W = np.tile(w, n_req)
    off = np.arange(0, 2 * n_req + 1, 2, dtype=np.uint64)
    z, _ = orc.foldin(W, off, np.arange(n_req, dtype=np.uint32), 1, 7, K, phi, sig, prob, al, threads=4)
    pairs = z.reshape(-1, 2)
    obs = np.bincount(pairs[:, 0] * K + pairs[:, 1], minlength=K * K).astype(float)
    init1 = ph[w[1]] / ph[w[1]].sum()
    ex = np.zeros((K, K))
    for z1 in range(K):
        p0 = ph[w[0]] * (float(alpha) + (np.arange(K) == z1))
        p0 /= p0.sum()
        for a in range(K):
            p1 = ph[w[1]] * (float(alpha) + (np.arange(K) == a))
            ex[a] += init1[z1] * p0[a] * p1 / p1.sum()
    ex = ex.reshape(-1) * n_req
    keep = ex > 0
    assert obs[~keep].sum() == 0
    assert stats.chisquare(obs[keep], ex[keep]).pvalue > 1e-4


def test_foldin_recovers_topic(orc):
    """Requests drawn from one topic's vocabulary fold in onto that topic."""
    V, K = 40, 4
    phi_t = np.zeros((V, K), np.float32)
    for k in range(K):
        phi_t[k * 10:(k + 1) * 10, k] = 3
    phi_t[:, :] += (np.random.default_rng(0).random((V, K)) < 0.05) * 1
    phi, sig, prob, al = orc.phi_finish(phi_t, phi_t.sum(0).astype(np.uint64), 0.1)
    rng = np.random.default_rng(1)
    reqs = [rng.integers(20, 30, 200).astype(np.uint32) for _ in range(20)]  # topic 2's words
    off = np.concatenate([[0], np.cumsum([len(r) for r in reqs])]).astype(np.uint64)
    z, _ = orc.foldin(np.concatenate(reqs), off, np.arange(20, dtype=np.uint32), 20, 3, K, phi, sig, prob, al)
    assert (z == 2).mean() > 0.9


PPU_COUNTS_REF = np.array([[1, 10, 100], [1, 1, 1], [1000, 1000, 1000]], np.uint32)  # [topic][word]


def ppu_draws(orc, iters):
    """Normalised Polya-urn draws of test_polyaurn.cu:76-118's counts over `iters` iterations."""
    n = np.ascontiguousarray(PPU_COUNTS_REF.T)  # word-major here
    tables = orc.poisson_tables(0.01)[2]
    out = []
    for it in range(1, iters + 1):
        pt, ts = orc.phi_draw(n, 0.01, it, 0, tables)
        out.append(orc.phi_finish(pt, ts, 0.1)[0])
    return np.array(out)  # (iters, V, K)


def test_ppu_row_shapes_reference(orc):
    """test_polyaurn.cu:76-118 restated: topic 0 with counts (1, 10, 100) over three words comes out
    near (0.01, 0.09, 0.9), topic 1 with (1, 1, 1) as ratios of small integers, topic 2 with
    (1000, 1000, 1000) near 1/3 each.  The reference asserts single seed-dependent draws (its
    0/0.5/1 set for topic 1 rejects valid draws such as 1/3); here every draw is checked for exactness
    and the averages over 400 independent iterations against the Poisson means."""
    phis = ppu_draws(orc, 400).astype(np.float64)
    pt_all = []
    for it in (1, 2, 3):
        pt, ts = orc.phi_draw(np.ascontiguousarray(PPU_COUNTS_REF.T), 0.01, it, 0)
        assert np.all(pt == np.floor(pt))
        pt_all.append((pt, ts))
    colsum = phis.sum(axis=1)  # (iters, K): each topic column sums to 1 (or is all zero)
    assert np.all((np.abs(colsum - 1) < 1e-6) | (colsum == 0))
    lam = PPU_COUNTS_REF.T.astype(np.float64) + 0.01
    mean = phis.mean(axis=0)
    assert np.allclose(mean[:, 0], lam[:, 0] / lam[:, 0].sum(), atol=0.006)  # ~ (0.009, 0.090, 0.901)
    p_any = 1 - np.exp(-lam[:, 1].sum())
    assert np.allclose(mean[:, 1], p_any / 3, atol=0.03)
    assert np.all(np.abs(phis[:, :, 2] - 1 / 3) < 0.05) and np.allclose(mean[:, 2], 1 / 3, atol=0.004)
    # topic 1: every value is x / S for Poisson integers x
    for pt, ts in pt_all:
        phi = orc.phi_finish(pt, ts, 0.1)[0]
        if ts[1]:
            assert np.array_equal(phi[:, 1], (pt[:, 1] / np.float32(ts[1])).astype(np.float32))
