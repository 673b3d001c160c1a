"""GPU parity: every stage of the HIP library against the CPU oracle, bit for bit, plus the
five-call ABI end to end (tiny corpus, 100 iterations, reference train.rs buffer loop).

Integer results (z, n_wk, topic sums, alias indices, Poisson draws, map layouts) must be equal;
floating-point results (phi, sigma_a, alias probabilities, scans) must be bit-identical because
both sides evaluate the same operations in the same order.  The log-likelihood is computed in
double with different lgamma implementations: tolerance 1e-9 relative.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def S():
    from gpupolyaurnlda_amd import stages
    return stages


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)


def test_device_is_gfx950():
    from gpupolyaurnlda_amd._lib import lib
    name = C.create_string_buffer(256)
    cu = C.c_int(0)
    hbm = C.c_uint64(0)
    lds = C.c_int(0)
    assert lib().gpulda_device_info(name, 256, C.byref(cu), C.byref(hbm), C.byref(lds)) == 0
    say(f"device {name.value.decode()} CUs={cu.value} HBM={hbm.value / 2**30:.0f} GiB LDS/block={lds.value}")
    assert b"gfx950" in name.value


def test_philox(S, orc):
    vec = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    ctr = np.array([v[0] for v in vec], dtype=np.uint32)
    key = np.array([v[1] for v in vec], dtype=np.uint32)
    exp = np.array([v[2] for v in vec], dtype=np.uint32)
    assert np.array_equal(S.philox(ctr, key), exp)
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2**32, size=(5000, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2**32, size=(5000, 2), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(S.philox(ctr, key), orc.philox(ctr, key))


def test_scan64(S, orc):
    rng = np.random.default_rng(1)
    x = (rng.random((300, 64)) * rng.choice([0, 1, 1e-3, 1e3], size=(300, 64))).astype(np.float32)
    assert np.array_equal(S.scan64(x).view(np.uint32), orc.scan64(x).view(np.uint32))


@pytest.mark.parametrize("K,L", [(20, 50), (1024, 600), (300, 64), (4096, 5000), (5, 3)])
def test_map_layout(S, orc, K, L):
    rng = np.random.default_rng(K * 7 + L)
    cap = orc.map_capacity(L, K)
    counts = {}
    ops = []
    total = 0
    for _ in range(4 * L):
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
    assert np.array_equal(S.map_apply(ops, cap, K), orc.map_apply(ops, cap, K))


@pytest.mark.parametrize("m", [1, 5, 20, 64, 200, 256, 257, 1024, 1025, 3000, 4096, 4097, 6000])
def test_alias_rows(S, orc, m):
    """Rows of up to 4,096 slots take the four-wavefront kernel (one to sixteen slots per thread),
    longer rows the one-wavefront kernel; both against the oracle's split, bit for bit."""
    rng = np.random.default_rng(m)
    rows = np.stack([rng.random(m), rng.random(m) * (rng.random(m) < 0.1), np.ones(m), np.zeros(m),
                     np.exp(rng.normal(0, 6, m))]).astype(np.float32)
    ro, sums, prob, al = S.alias(rows, alpha=0.25)
    for r in range(len(rows)):
        t = orc.sum64(rows[r])
        assert sums[r].view(np.uint32) == np.float32(0.25 * t).view(np.uint32) or sums[r] == np.float32(0.25) * t
        p2, a2 = orc.alias_build(rows[r], t)
        assert np.array_equal(prob[r].view(np.uint32), p2.view(np.uint32)), r
        assert np.array_equal(al[r], a2), r


def test_phi_step_bit_exact(S, orc):
    rng = np.random.default_rng(2)
    V, K = 300, 64
    n = rng.integers(0, 30, size=(V, K)).astype(np.uint32)
    n[rng.random((V, K)) < 0.7] = 0
    n[5, 3] = 100
    n[7, 9] = 5000
    n[11, 0] = 2_000_000
    pt, ts = S.phi_draw(n, 0.01, 4, seed=9)
    pt2, ts2 = orc.phi_draw(n, 0.01, 4, 9)
    assert np.array_equal(pt, pt2) and np.array_equal(ts, ts2)
    rows_out, sums, prob, al = S.alias(pt, divisor=ts, alpha=0.1)
    phi, sig, p2, a2 = orc.phi_finish(pt2, ts2, 0.1)
    assert np.array_equal(rows_out.view(np.uint32), phi.view(np.uint32))
    assert np.array_equal(sums.view(np.uint32), sig.view(np.uint32))
    assert np.array_equal(prob.view(np.uint32), p2.view(np.uint32)) and np.array_equal(al, a2)


@pytest.mark.parametrize("n", [0, 1, 10, 99, 100, 1000, 100000])
def test_poisson_bit_exact(S, orc, n):
    vals = np.full(4000, n, np.uint32)
    g = S.poisson(vals, 0.01, c1=3, c2=2, tag=3, seed=77)
    c = orc.poisson(vals, 0.01, c1=3, c2=2, tag=3, seed=77)
    assert np.array_equal(g, c)


def _tiny_phi(orc, V, K, seed=5, alpha=0.3):
    rng = np.random.default_rng(seed)
    pt = rng.integers(0, 5, size=(V, K)).astype(np.float32) * (rng.random((V, K)) < 0.6)
    pt[:, 0] += 1
    ts = pt.sum(0).astype(np.uint64)
    return orc.phi_finish(pt, ts, alpha)


@pytest.mark.parametrize("K,V,D,meanL", [(5, 3, 200, 6), (20, 1000, 500, 50), (256, 2000, 300, 300),
                                         (1024, 3000, 40, 1500), (4096, 2000, 12, 6000)])
def test_sweep_bit_exact(S, orc, K, V, D, meanL):
    rng = np.random.default_rng(K + D)
    phi, sig, prob, al = _tiny_phi(orc, V, K, seed=K)
    L = np.maximum(1, rng.poisson(meanL, D)).astype(np.uint32)
    L[0] = 1
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.uint64)
    N = int(off[-1])
    w = rng.integers(0, V, N).astype(np.uint32)
    z = rng.integers(0, K, N).astype(np.uint32)
    say(f"sweep K={K} V={V} D={D} N={N}")
    zg, vg = S.sweep(z, w, off, 17, 3, 5, K, phi, sig, prob, al)
    zc, vc = orc.sweep(z, w, off, 17, 3, 5, K, phi, sig, prob, al, threads=8)
    assert vg == vc
    assert np.array_equal(zg, zc)
    assert not np.array_equal(zg, z)


def test_sweep_length_limits(S, orc):
    """Empty documents beside one of the longest allowed length (65,535 tokens, 16-bit counts)."""
    rng = np.random.default_rng(77)
    V, K = 500, 64
    phi, sig, prob, al = _tiny_phi(orc, V, K, seed=9)
    L = np.array([0, 65535, 0, 3, 0], np.uint32)
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.uint64)
    N = int(off[-1])
    w = rng.integers(0, V, N).astype(np.uint32)
    z = np.zeros(N, np.uint32)  # one topic holds every token of the long document
    zg, vg = S.sweep(z, w, off, 5, 1, 3, K, phi, sig, prob, al)
    zc, vc = orc.sweep(z, w, off, 5, 1, 3, K, phi, sig, prob, al, threads=8)
    assert vg == vc and np.array_equal(zg, zc)
    assert len(np.unique(zg[1:65536])) > 1


def test_sweep_exact_conditional_gpu(S, orc):
    """Two-token documents: the joint law of (z0', z1') after one sweep, enumerated exactly."""
    from scipy import stats
    phi, sig, prob, al = _tiny_phi(orc, 3, 4, seed=5)
    K = 4
    w = np.array([1, 2], np.uint32)
    z0 = np.array([2, 3], np.uint32)
    n_docs = 100000
    off = np.arange(0, 2 * n_docs + 1, 2, dtype=np.uint64)
    zg, _ = S.sweep(np.tile(z0, n_docs), np.tile(w, n_docs), off, 0, 1, 123, K, phi, sig, prob, al)
    pairs = zg.reshape(-1, 2)
    obs = np.bincount(pairs[:, 0] * K + pairs[:, 1], minlength=K * K).astype(float)
    ph = phi.astype(np.float64)
    alpha = 0.3
    ex = np.zeros((K, K))
    p0 = ph[w[0]] * (alpha + (np.arange(K) == z0[1]))
    p0 /= p0.sum()
    for a in range(K):
        p1 = ph[w[1]] * (alpha + (np.arange(K) == a))
        ex[a] = p0[a] * p1 / p1.sum()
    ex = ex.reshape(-1) * n_docs
    keep = ex > 0
    assert obs[~keep].sum() == 0
    assert stats.chisquare(obs[keep], ex[keep]).pvalue > 1e-4


@pytest.mark.parametrize("V,K", [(2000, 300), (500, 1023), (40, 20000), (40, 50000)])
def test_count(S, orc, V, K):
    """n_wk through the word-order reduction. K = 1023 leaves a partial group of four bins; K =
    20000 fits two histograms a block in LDS; K = 50000 fits none and adds counts directly."""
    rng = np.random.default_rng(8)
    N = 300000
    w = (rng.zipf(1.3, N) % V).astype(np.uint32)
    z = rng.integers(0, K, N).astype(np.uint32)
    assert np.array_equal(S.count(w, z, V, K), orc.count(w, z, V, K))


def _tiny_corpus():
    from gpupolyaurnlda_amd import corpus
    return corpus.generate(corpus.TINY)


@pytest.mark.parametrize("f32_rows", [None, 0])
def test_abi_tiny_100_iterations(orc, f32_rows):
    """train.rs buffer loop through the five ABI calls vs the oracle Model (counts from C).  With
    f32_rows = 0 every word's sweep reads its u16 count row (by default the tiny model's rows are
    all kept f32: 1 MiB of them is more than its whole Φ)."""
    from gpupolyaurnlda_amd._lib import check, lib
    from gpupolyaurnlda_amd.corpus import TINY
    from gpupolyaurnlda_amd.train import Sampler, train
    c = _tiny_corpus()
    w, d, z0 = c["w"], c["d_len"], c["z"]
    k_d = np.array([len(np.unique(z0[c["doc_off"][i]:c["doc_off"][i + 1]])) for i in range(len(d))], np.uint32)
    buffer_size, max_D = 4096, 64  # several buffers per iteration, as train.rs streams the corpus
    model = orc.Model(w, c["doc_off"], z0, TINY.V, TINY.K, TINY.alpha, TINY.beta, seed=0, counts="C", threads=8)
    z = z0.copy()
    checkpoints = {1, 10, 100}
    with Sampler(TINY.alpha, TINY.beta, TINY.K, c["C"], buffer_size, max_D, seed=0) as smp:
        if f32_rows is not None:
            check(lib().gpulda_set_f32_rows(f32_rows), "set_f32_rows")

        def on_it(it):
            # the oracle replays the same iteration: phi step, sweep of every buffer, counts
            model.phi_step()
            doc = 0
            tok = 0
            from gpupolyaurnlda_amd.train import fill_buffer
            while True:
                nd, nt = fill_buffer(d, doc, buffer_size, max_D)
                if nd == 0:
                    break
                model.z_step(doc_base=doc, lo=doc, hi=doc + nd)
                doc += nd
                tok += nt
            model.n = orc.count(model.w, model.z, TINY.V, TINY.K)
            if it + 1 in checkpoints:
                say(f"ABI iteration {it + 1}")
                assert np.array_equal(z, model.z), f"z differs after iteration {it + 1}"
                assert np.array_equal(smp.counts(), model.n), f"n_wk differs after iteration {it + 1}"
                assert np.array_equal(smp.phi().view(np.uint32), model.phi.view(np.uint32))
                p, a = smp.alias()
                assert np.array_equal(p.view(np.uint32), model.prob.view(np.uint32)) and np.array_equal(a, model.al)
        train(smp, z, w, d, k_d, n_mc=100, buffer_size=buffer_size, max_D=max_D, on_iteration=on_it)
        assert smp.iteration == 100
        if f32_rows == 0:
            on, bad, fit = C.c_int(0), np.zeros(1, np.uint64), np.zeros(1, np.uint64)
            check(lib().gpulda_check_count_rows(C.byref(on), bad.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                fit.ctypes.data_as(C.POINTER(C.c_uint64))), "check_count_rows")
            say(f"ABI count rows: on={on.value} mismatches={int(bad[0])} words read as counts={int(fit[0])}")
            assert on.value == 1 and bad[0] == 0 and fit[0] > 0


@pytest.mark.parametrize("buffer_size,max_D", [(1 << 20, 1 << 16), (4096, 64)])
def test_abi_pipelined_parts(orc, buffer_size, max_D):
    """sample_z_async split into parts whose copies overlap other parts' sweeps (small parts so
    that a tiny corpus has up to eight of them): the same chain as the oracle."""
    from gpupolyaurnlda_amd._lib import check, lib
    from gpupolyaurnlda_amd.corpus import TINY
    from gpupolyaurnlda_amd.train import Sampler, fill_buffer, train
    c = _tiny_corpus()
    w, d, z0 = c["w"], c["d_len"], c["z"]
    k_d = np.zeros(len(d), np.uint32)
    model = orc.Model(w, c["doc_off"], z0, TINY.V, TINY.K, TINY.alpha, TINY.beta, seed=2, counts="C", threads=8)
    z = z0.copy()
    check(lib().gpulda_set_part_tokens(257), "set_part_tokens")
    try:
        with Sampler(TINY.alpha, TINY.beta, TINY.K, c["C"], buffer_size, max_D, seed=2) as smp:
            train(smp, z, w, d, k_d, n_mc=5, buffer_size=buffer_size, max_D=max_D)
            n_gpu = smp.counts()
    finally:
        check(lib().gpulda_set_part_tokens(4 << 20), "set_part_tokens")
    for _ in range(5):
        model.phi_step()
        doc = 0
        while True:
            nd, _ = fill_buffer(d, doc, buffer_size, max_D)
            if nd == 0:
                break
            model.z_step(doc_base=doc, lo=doc, hi=doc + nd)
            doc += nd
        model.n = orc.count(model.w, model.z, TINY.V, TINY.K)
    assert np.array_equal(z, model.z)
    assert np.array_equal(n_gpu, model.n)


def test_abi_reports_topic_out_of_range():
    """A topic indicator >= K in the caller's z ends the run at sync_buffer with a message, as
    every ABI error does (src/cu/error.cuh:13-18); the sweep itself runs normally."""
    import subprocess
    import sys
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "from gpupolyaurnlda_amd.train import Sampler, train\n"
        "rng = np.random.default_rng(0)\n"
        "V, K, D = 50, 8, 20\n"
        "d = np.full(D, 10, np.uint32); w = rng.integers(0, V, 200).astype(np.uint32)\n"
        "z = rng.integers(0, K, 200).astype(np.uint32); z[37] = K\n"
        "C = np.bincount(w, minlength=V).astype(np.uint32)\n"
        "with Sampler(0.1, 0.01, K, C, 1000, 100) as s:\n"
        "    train(s, z, w, d, np.zeros(D, np.uint32), 1, 1000, 100)\n"
        "print('not reached')\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode != 0
    assert "topic indicators were >= K" in out.stdout + out.stderr
    assert "not reached" not in out.stdout


def test_resident_matches_oracle_and_loglik(orc):
    from gpupolyaurnlda_amd.corpus import TINY
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    c = _tiny_corpus()
    model = orc.Model(c["w"], c["doc_off"], c["z"], TINY.V, TINY.K, TINY.alpha, TINY.beta, seed=3, counts="z",
                      threads=8)
    with ResidentCorpus(c["w"], c["d_len"], TINY.K, TINY.V, TINY.alpha, TINY.beta, seed=3, z=c["z"]) as rc:
        assert np.array_equal(rc.counts(), model.n)
        for it in range(1, 21):
            rc.iterate_untimed(1)
            model.iterate()
            if it in (1, 5, 20):
                assert np.array_equal(rc.z(), model.z), it
                assert np.array_equal(rc.counts(), model.n), it
                ll_g = rc.loglik()
                ll_c = orc.loglik(model.n, model.z, c["doc_off"], float(np.float32(TINY.alpha)),
                                  float(np.float32(TINY.beta)))
                say(f"resident it={it} LL gpu={ll_g:.6f} cpu={ll_c:.6f}")
                assert abs(ll_g - ll_c) <= 1e-9 * abs(ll_c)
        assert rc.visited() == model.visited


def test_resident_device_init_equals_generator_init():
    from gpupolyaurnlda_amd.corpus import TINY
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    c = _tiny_corpus()
    with ResidentCorpus(c["w"], c["d_len"], TINY.K, TINY.V, TINY.alpha, TINY.beta, seed=0, z=None) as rc:
        assert np.array_equal(rc.z(), c["z"])


def test_sweep_growth_bit_exact(S, orc):
    """Documents whose distinct-topic count grows far beyond the starting map (regrow path)."""
    rng = np.random.default_rng(11)
    V, K = 400, 1024
    pt = (rng.random((V, K)) < 0.02).astype(np.float32) * rng.integers(1, 4, (V, K)).astype(np.float32)
    pt[:, 0] += 1
    phi, sig, prob, al = orc.phi_finish(pt, pt.sum(0).astype(np.uint64), 0.5)
    L = np.array([1, 64, 300, 900, 2000], np.uint32)
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.uint64)
    N = int(off[-1])
    w = rng.integers(0, V, N).astype(np.uint32)
    z = np.zeros(N, np.uint32)  # one distinct topic per document at the start
    zg, vg = S.sweep(z, w, off, 5, 2, 9, K, phi, sig, prob, al)
    zc, vc = orc.sweep(z, w, off, 5, 2, 9, K, phi, sig, prob, al)
    assert vg == vc and np.array_equal(zg, zc)
    assert len(np.unique(zg[int(off[3]):int(off[4])])) > 100
    # known answer: one-hot rows force topic 3v+1
    pt1 = np.zeros((300, K), np.float32)
    pt1[np.arange(300), (3 * np.arange(300) + 1) % K] = 5
    phi1, sig1, prob1, al1 = orc.phi_finish(pt1, np.maximum(pt1.sum(0), 1).astype(np.uint64), 0.05)
    w1 = np.arange(300, dtype=np.uint32)
    z1, _ = S.sweep(np.zeros(300, np.uint32), w1, np.array([0, 300], np.uint64), 0, 1, 3, K, phi1, sig1, prob1, al1)
    assert np.array_equal(z1, (3 * w1 + 1) % K)


def _schedules_corpus():
    from gpupolyaurnlda_amd import corpus
    spec = corpus.BATCHED.scaled(3000)
    return spec, corpus.generate(spec)


def test_schedules_and_graph_bit_identical():
    """Per-bucket launches, the persistent queue, persistent + block-per-document, the split
    persistent launches and a replayed hipGraph give the same chain (same Philox streams) on a
    power-law corpus."""
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec, c = _schedules_corpus()
    say(f"schedules corpus N={len(c['w'])} max L={int(c['d_len'].max())}")
    results = {}
    for name, sched, graph in [("buckets", 0, False), ("persistent", 1, False), ("block", 2, False),
                               ("split", 3, False), ("graph", 0, True), ("graph-block", 2, True),
                               ("graph-split", 3, True)]:
        with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=4, z=c["z"]) as rc:
            rc.schedule(sched, 8, {2: 1000, 3: 64}.get(sched, 0))
            rc.use_graph(graph)
            zs = []
            for _ in range(3):  # replays with eager reads in between
                rc.iterate_untimed(1)
                zs.append(rc.z())
            results[name] = (zs, rc.counts(), rc.visited())
    ref = results["buckets"]
    for name, (zs, n, vis) in results.items():
        for it in range(3):
            assert np.array_equal(zs[it], ref[0][it]), (name, it)
        assert np.array_equal(n, ref[1]) and vis == ref[2], name


def test_autotune_keeps_chain():
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec, c = _schedules_corpus()
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=4, z=c["z"]) as rc:
        rc.iterate_untimed(1)
        z1 = rc.z()
        rows, chosen = rc.autotune()
        say(f"autotune: {rows} chosen={chosen}")
        assert len(rows) >= 5 and 0 <= chosen < len(rows)
        assert np.array_equal(rc.z(), z1)
        rc.iterate_untimed(1)
        z_tuned = rc.z()
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=4, z=c["z"]) as rc:
        rc.iterate_untimed(2)
        assert np.array_equal(rc.z(), z_tuned)


def test_count_rows_exact_and_unfit_rows(orc):
    """The u16 count rows the persistent sweep reads where a word's counts fit 16 bits: every phi
    recovered from them equals the f32 row bit for bit, and the chain equals the f32 rows' chain and
    the oracle's.  Word 0 takes half of the 600k tokens over four topics, so its counts pass 16 bits
    and its row stays f32 (the kernel mixes both kinds of row in one sweep); the two next most
    frequent words are kept on f32 rows by gpulda_set_f32_rows."""
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    rng = np.random.default_rng(21)
    V, K, D, L = 60, 4, 2000, 300
    N = D * L
    w = np.where(rng.random(N) < 0.5, 0, rng.integers(1, V, N)).astype(np.uint32)
    d_len = np.full(D, L, np.uint32)
    z = rng.integers(0, K, N).astype(np.uint32)
    doc_off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.uint64)
    alpha, beta = 0.1, 0.01
    runs = {}
    for on in (True, False):
        with ResidentCorpus(w, d_len, K, V, alpha, beta, seed=7, z=z) as rc:
            rc.schedule(1, 32)
            rc.count_rows(on)
            rc.f32_rows(3)  # word 0 and the two next most frequent
            zs = []
            for it in range(3):
                rc.iterate_untimed(1)
                zs.append(rc.z())
                if on:
                    cr_on, bad, fit = rc.check_count_rows()
                    say(f"count rows it={it + 1}: mismatches={bad} words held as counts={fit} of {V}")
                    assert cr_on and bad == 0 and fit == V - 3
            runs[on] = (zs, rc.counts())
    model = orc.Model(w, doc_off, z, V, K, alpha, beta, seed=7, counts="z", threads=8)
    for it in range(3):
        model.iterate()
        assert np.array_equal(runs[True][0][it], model.z), it
        assert np.array_equal(runs[False][0][it], model.z), it
    assert np.array_equal(runs[True][1], model.n) and np.array_equal(runs[False][1], model.n)


@pytest.mark.parametrize("V,K,Lmax,n_req,zero_rows", [(500, 256, 800, 60, 0), (3000, 1024, 2500, 24, 40)])
def test_foldin_bit_exact(S, orc, V, K, Lmax, n_req, zero_rows):
    """Fold-in against the oracle.  The second case has maps beyond eight lines (the line-loop
    draw) and words whose phi row is all zero (no draw), both of which end the next-token
    gathers."""
    rng = np.random.default_rng(21 + K)
    if zero_rows:
        pt = rng.integers(0, 5, size=(V, K)).astype(np.float32) * (rng.random((V, K)) < 0.6)
        pt[:, 0] += 1
        pt[rng.choice(V, zero_rows, replace=False)] = 0
        phi, sig, prob, al = orc.phi_finish(pt, pt.sum(0).astype(np.uint64), 0.1)
        assert (sig == 0).sum() == zero_rows
    else:
        phi, sig, prob, al = _tiny_phi(orc, V, K, seed=3, alpha=0.1)
    L = np.concatenate([[1, 5], rng.integers(10, Lmax, n_req)]).astype(np.uint32)
    off = np.concatenate([[0], np.cumsum(L)]).astype(np.uint64)
    w = rng.integers(0, V, int(off[-1])).astype(np.uint32)
    ids = (np.arange(len(L)) * 7 + 100).astype(np.uint32)
    zg, vg = S.foldin(w, off, ids, 12, 5, K, phi, sig, prob, al)
    zc, vc = orc.foldin(w, off, ids, 12, 5, K, phi, sig, prob, al, threads=8)
    assert vg == vc and np.array_equal(zg, zc)


def test_server_alone_equals_mixed(orc):
    """Continuous batching: each request's result is bit-identical whether it runs alone or in
    a mixed stream, and equals the stage fold-in of the same request id."""
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    from gpupolyaurnlda_amd.serve import FoldInServer, run_serial, run_stream
    from gpupolyaurnlda_amd import stages
    spec = corpus.MEDIUM.scaled(2000)
    c = corpus.generate(spec)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=1, z=c["z"]) as rc:
        rc.iterate_untimed(3)
        phi = rc.phi()
        sig = rc.sigma_a()
        prob, al = rc.alias()
        d_len, arr, w = corpus.requests(phi, 60, median=120, sigma=1.0, lo=10, hi=1500, rate=3000.0, seed=4)
        with FoldInServer(hbm_budget_bytes=64 << 20, sweeps=20, waves_per_cu=1, max_slots=256) as srv:
            mixed = run_stream(srv, d_len, arr, w)
            alone = run_serial(srv, d_len[:12], w[:int(d_len[:12].sum())])
        say(f"server p50={np.median(mixed['latency_s']) * 1e3:.2f} ms serial mean={alone['service_s'].mean() * 1e3:.2f} ms")
    off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.uint64)
    ids = np.arange(len(d_len), dtype=np.uint32)
    zs, _ = stages.foldin(w, off, ids, 20, 1, spec.K, phi, sig, prob, al)
    zc, _ = orc.foldin(w, off, ids, 20, 1, spec.K, phi, sig, prob, al, threads=8)
    assert np.array_equal(zs, zc)
    want = [zs[off[i]:off[i + 1]] for i in range(len(d_len))]
    bad_mixed = [i for i in range(len(d_len)) if not np.array_equal(mixed["results"][i], want[i])]
    bad_alone = [i for i in range(12) if not np.array_equal(alone["results"][i], want[i])]
    assert not bad_mixed and not bad_alone, f"mixed wrong {bad_mixed}, alone wrong {bad_alone}"


def test_server_deadline_reports_lost_and_recovers(orc):
    """A request with no result within the deadline is reported by poll and collected as failed
    (here a deadline of 1 us, so the request is reported while its worker may still run); its
    slot is held until the pool kernels have ended, then serves the next request, whose result
    equals the stage fold-in."""
    import time
    from gpupolyaurnlda_amd import corpus, stages
    from gpupolyaurnlda_amd._lib import check, lib
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    from gpupolyaurnlda_amd.serve import FoldInServer
    spec = corpus.MEDIUM.scaled(2000)
    c = corpus.generate(spec)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=1, z=c["z"]) as rc:
        rc.iterate_untimed(3)
        phi, sig = rc.phi(), rc.sigma_a()
        prob, al = rc.alias()
        d_len, _, w = corpus.requests(phi, 2, median=120, sigma=1.0, lo=10, hi=800, rate=1000.0, seed=5)
        w0, w1 = w[:int(d_len[0])], w[int(d_len[0]):]
        with FoldInServer(hbm_budget_bytes=64 << 20, sweeps=20, waves_per_cu=1, max_slots=1) as srv:
            try:
                check(lib().gpulda_serve_set_deadline(1e-6), "serve_set_deadline")
                slot = srv.submit(w0, 0)
                assert slot == 0
                srv.flush()
                got, t0 = [], time.time()
                while not got and time.time() - t0 < 10:
                    got = srv.poll()
                assert got == [0]
                with pytest.raises(RuntimeError, match="no result within"):
                    srv.result(0, len(w0))
            finally:
                check(lib().gpulda_serve_set_deadline(30.0), "serve_set_deadline")
            say("lost request reported; waiting for the pool to end")
            slot, t0 = None, time.time()
            while slot is None and time.time() - t0 < 20:  # the pool stops claiming after 2 s
                srv.poll()
                slot = srv.submit(w1, 1)
                if slot is None:
                    time.sleep(0.05)
            assert slot == 0, "quarantined slot was not released"
            srv.flush()
            got, t0 = [], time.time()
            while not got and time.time() - t0 < 20:
                got = srv.poll()
            assert got == [0]
            z1, rid = srv.result(0, len(w1))
    assert rid == 1
    off = np.array([0, len(w1)], dtype=np.uint64)
    zs, _ = stages.foldin(w1, off, np.array([1], dtype=np.uint32), 20, 1, spec.K, phi, sig, prob, al)
    assert np.array_equal(z1, zs)


def test_medium_config_parity_and_ll_trace(orc):
    """Medium workload shape (V=50k Zipf, K=256, lognormal lengths) on a 20k-document slice:
    three iterations bit-exact against the oracle, then a 60-iteration log-likelihood trace on
    the GPU that must rise from the random start and settle (last 10 within 0.5%)."""
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec = corpus.MEDIUM.scaled(20000)
    c = corpus.generate(spec)
    say(f"medium slice N={len(c['w'])}")
    model = orc.Model(c["w"], c["doc_off"], c["z"], spec.V, spec.K, spec.alpha, spec.beta, seed=1, counts="z",
                      threads=16)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=1, z=c["z"]) as rc:
        lls = [rc.loglik()]
        for it in range(3):
            rc.iterate_untimed(1)
            model.iterate()
            assert np.array_equal(rc.z(), model.z), it
            assert np.array_equal(rc.counts(), model.n), it
            lls.append(rc.loglik())
        ll_c = orc.loglik(model.n, model.z, c["doc_off"], float(np.float32(spec.alpha)), float(np.float32(spec.beta)))
        assert abs(lls[-1] - ll_c) <= 1e-9 * abs(ll_c)
        for _ in range(57):
            rc.iterate_untimed(1)
            lls.append(rc.loglik())
    say(f"LL trace: start {lls[0]:.4e} it3 {lls[3]:.4e} it60 {lls[-1]:.4e}")
    assert lls[-1] > lls[0]
    tail = np.array(lls[-10:])
    assert np.abs(tail - tail.mean()).max() <= 0.005 * abs(tail.mean())


def test_ppu_row_shapes_match_oracle(S, orc):
    """The reference's Polya-urn row-shape inputs (test_polyaurn.cu:76-118) through the device
    Poisson draw and normalisation, bit for bit against the oracle over 50 iterations."""
    n = np.ascontiguousarray(np.array([[1, 10, 100], [1, 1, 1], [1000, 1000, 1000]], np.uint32).T)
    for it in range(1, 51):
        pt, ts = S.phi_draw(n, 0.01, it, seed=0)
        pt2, ts2 = orc.phi_draw(n, 0.01, it, 0)
        assert np.array_equal(pt, pt2) and np.array_equal(ts, ts2), it
        rows, _, _, _ = S.alias(pt, divisor=ts, alpha=0.1)
        assert np.array_equal(rows.view(np.uint32), orc.phi_finish(pt2, ts2, 0.1)[0].view(np.uint32)), it


def _cache_stats():
    from gpupolyaurnlda_amd._lib import lib, ptr
    seg = np.zeros(1, np.uint32)
    by = np.zeros(1, np.uint64)
    builds = np.zeros(1, np.uint64)
    assert lib().gpulda_slice_cache_stats(ptr(seg, C.c_uint32), ptr(by, C.c_uint64), ptr(builds, C.c_uint64)) == 0
    return int(seg[0]), int(by[0]), int(builds[0])


@pytest.mark.parametrize("mode", [1, 0])
def test_slice_cache_refilled_buffer(orc, mode):
    """One host buffer refilled in place with a slice that differs from the previous one in a single
    word: the default cache (whole-content hash) and no cache both upload it again, and z equals the
    oracle's.  (The reference uploads every call, train.cu:160-163.)"""
    from gpupolyaurnlda_amd._lib import check, lib
    from gpupolyaurnlda_amd.train import Sampler
    rng = np.random.default_rng(31)
    V, K, D, L = 3000, 64, 60, 300
    n = D * L
    wA = rng.integers(0, V, n).astype(np.uint32)
    wB = wA.copy()
    wB[1] = (wB[1] + 1) % V  # off the sampled fingerprint's grid
    zA = rng.integers(0, K, n).astype(np.uint32)
    zB = rng.integers(0, K, n).astype(np.uint32)
    d = np.full(D, L, np.uint32)
    w_all = np.concatenate([wA, wB])
    C_counts = np.bincount(w_all, minlength=V).astype(np.uint32)
    off = np.arange(0, 2 * n + 1, L, dtype=np.uint64)
    model = orc.Model(w_all, off, np.concatenate([zA, zB]), V, K, 0.1, 0.01, seed=6, counts="C", threads=8)
    model.phi_step()
    model.z_step(doc_base=0, lo=0, hi=D)
    model.z_step(doc_base=D, lo=D, hi=2 * D)
    wbuf, zbuf = wA.copy(), zA.copy()
    kd = np.zeros(D, np.uint32)
    check(lib().gpulda_set_slice_cache(mode), "set_slice_cache")
    try:
        with Sampler(0.1, 0.01, K, C_counts, n, D, seed=6) as smp:
            smp.sample_phi()
            smp.set_buffer(0, zbuf, wbuf, d, kd)
            smp.sample_z_async(0)
            smp.sync_buffer(0)
            out_a = zbuf.copy()
            b0 = _cache_stats()[2]
            wbuf[:] = wB
            zbuf[:] = zB
            smp.sample_z_async(0)
            smp.sync_buffer(0)
            assert _cache_stats()[2] == b0 + 1  # w uploaded again
    finally:
        check(lib().gpulda_set_slice_cache(1), "set_slice_cache")
    assert np.array_equal(out_a, model.z[:n])
    assert np.array_equal(zbuf, model.z[n:])


def test_slice_cache_tiny_budget_two_buffers_in_flight(orc):
    """A one-byte cache budget with two buffers in flight: each new slice evicts every cached slice
    except the one the other buffer is still sweeping, so slices are uploaded again and again, and
    z equals the oracle replaying the one-buffer loop of train.rs."""
    from gpupolyaurnlda_amd._lib import check, lib
    from gpupolyaurnlda_amd.corpus import TINY
    from gpupolyaurnlda_amd.train import Sampler, fill_buffer
    c = _tiny_corpus()
    w, d, z = c["w"], c["d_len"], c["z"].copy()
    buffer_size, max_D, n_mc = 3000, 40, 3
    k_d = np.zeros(len(d), np.uint32)
    model = orc.Model(w, c["doc_off"], c["z"], TINY.V, TINY.K, TINY.alpha, TINY.beta, seed=8, counts="C", threads=8)
    check(lib().gpulda_set_slice_cache_budget(1), "budget")
    try:
        with Sampler(TINY.alpha, TINY.beta, TINY.K, c["C"], buffer_size, max_D, n_buffers=2, seed=8) as smp:
            b0 = _cache_stats()[2]
            n_calls = 0
            for _ in range(n_mc):
                smp.sample_phi()
                doc = tok = 0
                busy = [None, None]
                slot = 0
                while True:
                    nd, nt = fill_buffer(d, doc, buffer_size, max_D)
                    if nd == 0:
                        break
                    if busy[slot] is not None:
                        smp.sync_buffer(slot)
                    smp.set_buffer(slot, z[tok:tok + nt], w[tok:tok + nt], d[doc:doc + nd], k_d[doc:doc + nd])
                    smp.sample_z_async(slot)
                    busy[slot] = True
                    n_calls += 1
                    doc += nd
                    tok += nt
                    slot ^= 1
                for s in (0, 1):
                    if busy[s] is not None:
                        smp.sync_buffer(s)
                segs, _, builds = _cache_stats()
                assert segs <= 2
            assert builds - b0 == n_calls  # nothing was kept: every call uploaded its slice
    finally:
        check(lib().gpulda_set_slice_cache_budget(0), "budget")
    for _ in range(n_mc):
        model.phi_step()
        doc = 0
        while True:
            nd, _ = fill_buffer(d, doc, buffer_size, max_D)
            if nd == 0:
                break
            model.z_step(doc_base=doc, lo=doc, hi=doc + nd)
            doc += nd
        model.n = orc.count(model.w, model.z, TINY.V, TINY.K)
    assert np.array_equal(z, model.z)
