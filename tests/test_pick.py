"""The sparse pick of the z-sweep, pinned independently of both implementations.

The sweep draws a token's new topic from p(k) proportional to phi[w][k] (alpha + n_dk): with
probability sigma_b / (sigma_a + sigma_b) from the doc-topic map's masses count * phi[w][key], by
inverse CDF over the map's slots (the reference's compute_product_cumsum + draw_wary_search,
src/cu/topics.cuh:163-192, 88-137).  Here the slots are laid end to end column by column -- lane
0's slots in line order, then lane 1's, ... (slot j*64+l is line j, lane l) -- and the pick is the
slot whose interval holds the target t = u * sigma_b.  That enumeration is what both sides compute
with f32 scans; the tests below derive the expected slot from it in float64, without either
implementation:

  * closed-form edge targets on a two-line map of equal masses (t = 0, a boundary between the two
    lines of a column, a boundary between columns, the last slot, t = sigma_b);
  * the reference's own known answers of draw_wary_search (test_topics.cu:78-132): 96 entries of
    equal mass, u in {0, .02, .16, .4, .985, 1.0} -> entries {0, 1, 15, 38, 94, 95};
  * every slot of maps of 3 to 67 lines hit by a target at the midpoint of its interval (the
    column search with a carry across 64-line scans included), and a chi-square of uniform
    targets;
  * the exact full conditional of the first token of a document whose map has 5 lines (K=512) or
    67 lines (K=4096), over tens of thousands of replicas under different document ids, against
    p(k) proportional to phi[w0][k] (alpha + n_dk without token 0).

The reference's draw_alias and compute_d_idx known answers (test_topics.cu:16-76) are restated at
the end.  GPU tests call the library's stage entry points (the sweep's own device functions) and
also compare with the oracle bit for bit.
"""
import numpy as np
import pytest
from scipy import stats

NO = -1  # dense draw


def two_level_order(T):
    """Slots of map T (occupied ones) in the pick's enumeration order: lane-major, then line."""
    lines = len(T) // 64
    order = []
    for lane in range(64):
        for j in range(lines):
            s = j * 64 + lane
            if T[s] & 0xFFFF:
                order.append(s)
    return order


def masses(T, prow):
    cnt = (T & 0xFFFF).astype(np.float64)
    key = (T >> 16).astype(np.int64)
    return np.where(cnt > 0, cnt * prow.astype(np.float64)[np.minimum(key, len(prow) - 1)], 0.0)


def uniform_map(n_entries, lines, K):
    """n_entries entries of count 1 in slots 0.. of a map of `lines` lines; the entry at position p of
    the enumeration order holds key p."""
    T = np.zeros(lines * 64, np.uint32)
    T[:n_entries] = 1
    for p, s in enumerate(two_level_order(T)):
        T[s] = (p << 16) | 1
    assert n_entries <= K
    return T


def random_map(rng, K, n_keys, cap, max_count=6):
    """A map of cap slots holding n_keys distinct topics (counts 1..max_count) at random slots."""
    T = np.zeros(cap, np.uint32)
    slots = rng.choice(cap, n_keys, replace=False)
    keys = rng.choice(K, n_keys, replace=False)
    T[slots] = (keys.astype(np.uint32) << 16) | rng.integers(1, max_count + 1, n_keys).astype(np.uint32)
    return T


def midpoint_targets(T, prow):
    """For each occupied slot with mass, the target fraction at the middle of its interval."""
    m = masses(T, prow)
    order = [s for s in two_level_order(T) if m[s] > 0]
    cm = np.cumsum([m[s] for s in order])
    total = cm[-1]
    lo = np.concatenate([[0.0], cm[:-1]])
    mid = (lo + cm) / 2 / total
    return np.array(order), mid.astype(np.float32), m, total


EDGE_CASES = [  # (u_target as k/128 or a special value, expected position in the order)
    (0.0, 0),                 # t = 0: first slot of the first column
    (1 / 128, 1),             # t = 0.5: boundary between the two lines of column 0 -> line 1
    (2 / 128, 2),             # t = 1.0: boundary between columns 0 and 1 -> column 1, line 0
    (3 / 128, 3),             # column 1, line 1
    (0.5, 64),                # t = 32: column 32, line 0
    (127 / 128, 127),         # last slot
    (float(np.float32(1 - 2**-24)), 127),  # largest uniform: last slot
    (1.0, 127),               # t = sigma_b (no interval holds it): the last slot with mass
]


def _edge_inputs():
    K = 128
    T = uniform_map(128, 2, K)
    prow = np.full(K, 0.5, np.float32)  # every mass 0.5, sigma_b = 64, exact in f32
    order = two_level_order(T)
    return K, T, prow, order


def test_oracle_pick_edge_targets(orc):
    K, T, prow, order = _edge_inputs()
    for u, p in EDGE_CASES:
        assert orc.sparse_pick(T, prow, 1e-30, 0.9, u)[0] == order[p], (u, p)
    # positions follow slot (p // 2) + 64 * (p % 2): two lines per column
    assert order[:4] == [0, 64, 1, 65] and order[127] == 127
    # sparse iff ux * (sigma_a + sigma_b) > sigma_a: with sigma_a = sigma_b = 64, ux = 1/2 is dense
    assert orc.sparse_pick(T, prow, 64.0, 0.5, 0.3)[0] == NO
    assert orc.sparse_pick(T, prow, 64.0, float(np.float32(0.5 + 2**-24)), 0.3)[0] != NO


def test_oracle_pick_reference_wary_search_vectors(orc):
    # test_topics.cu:78-132: 96 entries of equal mass, the entry at position p has key p
    K = 96
    T = uniform_map(96, 2, K)
    prow = np.full(K, 0.5, np.float32)
    for u, want in [(0.0, 0), (0.02, 1), (0.16, 15), (0.4, 38), (0.985, 94), (1.0, 95)]:
        s = orc.sparse_pick(T, prow, 1e-30, 0.9, np.float32(u))[0]
        assert T[s] >> 16 == want, (u, int(T[s] >> 16), want)


@pytest.mark.parametrize("K,n_keys,lines", [(512, 150, 3), (1024, 480, 8), (2048, 1200, 20), (4096, 4096, 67)])
def test_oracle_pick_every_slot_at_midpoint(orc, K, n_keys, lines):
    rng = np.random.default_rng(K + lines)
    T = random_map(rng, K, n_keys, lines * 64)
    prow = rng.gamma(0.7, 1.0, K).astype(np.float32) / np.float32(K)
    prow[rng.random(K) < 0.1] = 0  # zero-mass keys are skipped
    order, mid, m, total = midpoint_targets(T, prow)
    # keep slots whose interval is wide against f32 rounding of the prefix sums
    ok = np.array([m[s] > 2e-5 * total for s in order])
    got = orc.sparse_pick(T, prow, 1e-30, 0.99, mid[ok])
    assert np.array_equal(got, order[ok])
    assert lines < 64 or any((s // 64) >= 64 for s in order[ok])  # the second 64-line scan is reached


def test_oracle_pick_chisquare_uniform_targets(orc):
    rng = np.random.default_rng(5)
    K, lines = 4096, 67
    T = random_map(rng, K, 600, lines * 64, max_count=3)
    prow = rng.gamma(0.7, 1.0, K).astype(np.float32)
    n = 40000
    uw = (rng.integers(0, 2**24, n) * 2.0**-24).astype(np.float32)
    got = orc.sparse_pick(T, prow, 1e-30, 0.99, uw)
    m = masses(T, prow)
    occ = np.nonzero(m > 0)[0]
    obs = np.array([(got == s).sum() for s in occ], float)
    assert obs.sum() == n
    ex = m[occ] / m[occ].sum() * n
    assert chisq_pooled(obs, ex) > 1e-4


def first_token_law(phi_row, alpha, z_rest, K):
    """p(k) of token 0's new topic: phi[w0][k] (alpha + n_dk without token 0), normalised."""
    n = np.bincount(z_rest, minlength=K).astype(np.float64)
    p = phi_row.astype(np.float64) * (alpha + n)
    return p / p.sum()


def chisq_pooled(obs, ex, min_ex=5.0):
    big = ex >= min_ex
    o = np.concatenate([obs[big], [obs[~big].sum()]])
    e = np.concatenate([ex[big], [ex[~big].sum()]])
    if e[-1] == 0:
        o, e = o[:-1], e[:-1]
    return stats.chisquare(o, e).pvalue


def _multiline_doc(orc, K, V, L, distinct, seed, alpha=0.1):
    """One document of L tokens with `distinct` topics; phi from random counts; returns the pieces
    and the number of map lines of its first sweep."""
    rng = np.random.default_rng(seed)
    pt = rng.integers(1, 6, size=(V, K)).astype(np.float32) * (rng.random((V, K)) < 0.8)
    pt[:, 0] += 1
    phi, sig, prob, al = orc.phi_finish(pt, pt.sum(0).astype(np.uint64), alpha)
    topics = rng.choice(K, distinct, replace=False)
    z = np.concatenate([topics, rng.choice(topics, L - distinct)]).astype(np.uint32)
    rng.shuffle(z)
    w = rng.integers(0, V, L).astype(np.uint32)
    lines = orc.map_capacity(len(np.unique(z)), K) // 64
    return phi, sig, prob, al, z, w, lines


def test_oracle_sweep_first_token_conditional_five_lines(orc):
    K, V, L = 512, 50, 260
    phi, sig, prob, al, z, w, lines = _multiline_doc(orc, K, V, L, 255, seed=3)
    assert lines == 5
    reps = 20000
    off = np.arange(0, L * reps + 1, L, dtype=np.uint64)
    znew, _ = orc.sweep(np.tile(z, reps), np.tile(w, reps), off, 1000, 2, 77, K, phi, sig, prob, al, threads=8)
    first = znew.reshape(reps, L)[:, 0]
    p = first_token_law(phi[w[0]], float(np.float32(0.1)), z[1:], K)
    obs = np.bincount(first, minlength=K).astype(float)
    assert obs[p == 0].sum() == 0
    assert chisq_pooled(obs, p * reps) > 1e-4


def test_oracle_alias_draw_reference_known_answer(orc):
    # test_topics.cu:49-76: prob = 0.5, alias = 1 over 5 slots.  The reference draws slot floor(u n)
    # and keeps it when (u mod 1/n) n < prob: u = 0.6 -> slot 3, kept; u = 0.75 -> slot 3, u' = 0.75,
    # its alias 1.  Here slot and keep test take two words: r_slot = u 2^32, r_keep = u' 2^32.
    prob = np.full(5, 0.5, np.float32)
    al = np.ones(5, np.uint32)
    r_slot = [int(np.float32(0.6) * 2.0**32), int(np.float32(0.75) * 2.0**32)]
    r_keep = [0, int(0.75 * 2**32)]
    assert list(orc.alias_draw(prob, al, r_slot, r_keep)) == [3, 1]


# ---------------------------------------------------------------------------------------------
# GPU: the sweep's device functions through the stage entry points
# ---------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def S():
    from gpupolyaurnlda_amd import stages
    return stages


def say(msg):
    print(f"[gpu-pick] {msg}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("use_regs", [True, False])
def test_gpu_pick_edge_targets_and_reference_vectors(S, orc, use_regs):
    K, T, prow, order = _edge_inputs()
    u = np.array([c[0] for c in EDGE_CASES], np.float32)
    got = S.pick(T, K, prow, 1e-30, np.full(len(u), 0.9, np.float32), u, use_regs=use_regs)
    assert list(got) == [order[p] for _, p in EDGE_CASES]
    assert S.pick(T, K, prow, 64.0, [0.5], [0.3], use_regs)[0] == NO
    assert S.pick(T, K, prow, 64.0, [np.float32(0.5 + 2**-24)], [0.3], use_regs)[0] != NO
    K = 96
    T = uniform_map(96, 2, K)
    prow = np.full(K, 0.5, np.float32)
    uu = np.array([0.0, 0.02, 0.16, 0.4, 0.985, 1.0], np.float32)
    got = S.pick(T, K, prow, 1e-30, np.full(6, 0.9, np.float32), uu, use_regs=use_regs)
    assert [int(T[s] >> 16) for s in got] == [0, 1, 15, 38, 94, 95]


@pytest.mark.gpu
@pytest.mark.parametrize("K,n_keys,lines", [(512, 150, 3), (1024, 480, 8), (2048, 1200, 20), (4096, 4096, 67)])
def test_gpu_pick_every_slot_and_bit_exact(S, orc, K, n_keys, lines):
    rng = np.random.default_rng(K + lines)
    T = random_map(rng, K, n_keys, lines * 64)
    prow = rng.gamma(0.7, 1.0, K).astype(np.float32) / np.float32(K)
    prow[rng.random(K) < 0.1] = 0
    order, mid, m, total = midpoint_targets(T, prow)
    ok = np.array([m[s] > 2e-5 * total for s in order])
    forms = [True, False] if lines <= 8 else [False]
    for regs in forms:
        got = S.pick(T, K, prow, 1e-30, np.full(ok.sum(), 0.99, np.float32), mid[ok], use_regs=regs)
        assert np.array_equal(got, order[ok]), regs
    # random decisions and targets, against the oracle bit for bit (dense draws included)
    ux = rng.random(3000).astype(np.float32)
    uw = rng.random(3000).astype(np.float32)
    sa = np.float32(0.3 * total)
    want = orc.sparse_pick(T, prow, sa, ux, uw)
    for regs in forms:
        assert np.array_equal(S.pick(T, K, prow, sa, ux, uw, use_regs=regs), want), regs
    say(f"pick K={K} lines={lines}: {ok.sum()} midpoints, 3000 random draws bit-exact")


@pytest.mark.gpu
def test_gpu_sweep_first_token_conditional_five_lines(S, orc):
    K, V, L = 512, 50, 260
    phi, sig, prob, al, z, w, lines = _multiline_doc(orc, K, V, L, 255, seed=3)
    assert lines == 5
    reps = 40000
    off = np.arange(0, L * reps + 1, L, dtype=np.uint64)
    znew, _ = S.sweep(np.tile(z, reps), np.tile(w, reps), off, 1000, 2, 77, K, phi, sig, prob, al)
    first = znew.reshape(reps, L)[:, 0]
    p = first_token_law(phi[w[0]], float(np.float32(0.1)), z[1:], K)
    obs = np.bincount(first, minlength=K).astype(float)
    assert obs[p == 0].sum() == 0
    pv = chisq_pooled(obs, p * reps)
    say(f"first-token law, 5-line maps, {reps} replicas: p = {pv:.3g}")
    assert pv > 1e-4
    zc, _ = orc.sweep(np.tile(z, 2000), np.tile(w, 2000), off[:2001], 1000, 2, 77, K, phi, sig, prob, al, threads=16)
    assert np.array_equal(znew[:L * 2000], zc)


@pytest.mark.gpu
def test_gpu_sweep_first_token_conditional_67_lines(S, orc):
    """K=4096 and every topic in the document: a 67-line map, so the in-column search runs two
    64-line scans with a carry.  Token 0's word has phi mass on 32 low topics (line 0) and 32
    high ones (lines 65-66)."""
    K, V, L = 4096, 8, 4200
    rng = np.random.default_rng(17)
    pt = rng.integers(1, 4, size=(V, K)).astype(np.float32)
    low = np.arange(0, 64, 2)
    high = np.arange(K - 64, K, 2)
    pt[0] = 0
    pt[0, low] = rng.integers(1, 9, 32)
    pt[0, high] = rng.integers(1, 9, 32)
    phi, sig, prob, al = orc.phi_finish(pt, pt.sum(0).astype(np.uint64), 0.1)
    # every topic once, plus L - K more tokens on the marked topics (their counts vary)
    z = np.concatenate([np.arange(K), rng.choice(np.concatenate([low, high]), L - K)]).astype(np.uint32)
    rng.shuffle(z)
    w = rng.integers(1, V, L).astype(np.uint32)
    w[0] = 0
    assert orc.map_capacity(len(np.unique(z)), K) // 64 == 67
    reps = 10000
    off = np.arange(0, L * reps + 1, L, dtype=np.uint64)
    znew, _ = S.sweep(np.tile(z, reps), np.tile(w, reps), off, 50000, 3, 91, K, phi, sig, prob, al)
    first = znew.reshape(reps, L)[:, 0]
    p = first_token_law(phi[0], float(np.float32(0.1)), z[1:], K)
    obs = np.bincount(first, minlength=K).astype(float)
    assert obs[p == 0].sum() == 0
    pv = chisq_pooled(obs, p * reps)
    say(f"first-token law, 67-line maps, {reps} replicas: p = {pv:.3g}; high topics drawn {obs[high].sum():.0f}")
    assert pv > 1e-4 and obs[high].sum() > 0.2 * reps
    zc, _ = orc.sweep(z, w, off[:2], 50000, 3, 91, K, phi, sig, prob, al)
    assert np.array_equal(znew[:L], zc)


@pytest.mark.gpu
def test_gpu_alias_draw_reference_known_answer(S):
    prob = np.full(5, 0.5, np.float32)
    al = np.ones(5, np.uint32)
    r_slot = np.array([int(np.float32(0.6) * 2.0**32), int(np.float32(0.75) * 2.0**32)], np.uint32)
    r_keep = np.array([0, int(0.75 * 2**32)], np.uint32)
    assert list(S.alias_draw(prob, al, r_slot, r_keep)) == [3, 1]


@pytest.mark.gpu
def test_gpu_alias_draw_matches_oracle(S, orc):
    rng = np.random.default_rng(9)
    w = rng.gamma(0.5, 1.0, 1000).astype(np.float32)
    prob, al = orc.alias_build(w)
    rs = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)
    rk = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(S.alias_draw(prob, al, rs, rk), orc.alias_draw(prob, al, rs, rk))


@pytest.mark.gpu
def test_gpu_doc_offsets_reference_known_answer(S):
    # test_topics.cu:16-47: 527 documents of lengths 1..527; offset i = 1 + 2 + ... + i
    d = np.arange(1, 528, dtype=np.uint32)
    off = S.doc_offsets(d)
    i = np.arange(528, dtype=np.uint64)
    assert np.array_equal(off, i * (i + 1) // 2)
    assert np.array_equal(S.doc_offsets(np.array([0, 3, 0, 0, 2], np.uint32)), np.array([0, 0, 3, 3, 3, 5]))
