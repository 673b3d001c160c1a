"""Every workload shape the sampler is measured on, checked against the CPU oracle on the GPU.

  medium  V=50k Zipf(1.07), K=256, D=200k lognormal(100, 0.8) <= 2000 -- the whole corpus: two
          iterations bit-exact (z, n_wk, phi), then a 200-iteration log-likelihood trace
  large   V=100k Zipf(1.07), K=1024, lognormal(200, 1.0) <= 4096 -- a 20,000-document slice under
          the persistent schedule the benchmark runs: two iterations, z, n_wk and the full V x K
          phi bit-exact; and the whole corpus (328M tokens): exact n_wk, Phi / sigma_a / alias
          tables of the third iteration equal the oracle's, and the first 2,000 documents' z
          (the whole batched corpus likewise, under the split schedule as a replayed hipGraph)
  batched V=100k, K=1024, power-law lengths on [4, 20000] -- a 20,000-document slice under the
          persistent + block-per-document and the persistent split schedules, eagerly and as a
          replayed hipGraph, against the oracle at every iteration
  serving fold-in against a K=1024, V=100k model: 50 sweeps per request, lognormal(300, 1.2)
          lengths clipped to [10, 5000]; the server's results, mixed and alone, equal the oracle's
          fold-in of the same request ids

(The tiny shape runs 100 iterations through the five-call ABI in test_gpu_parity.py.)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-configs] {msg}", flush=True)


def _model(orc, spec, c, seed):
    return orc.Model(c["w"], c["doc_off"], c["z"], spec.V, spec.K, spec.alpha, spec.beta, seed=seed, counts="z",
                     threads=16)


def _same_state(rc, model, it, phi=True):
    assert np.array_equal(rc.z(), model.z), f"z differs at iteration {it}"
    assert np.array_equal(rc.counts(), model.n), f"n_wk differs at iteration {it}"
    if phi:
        assert np.array_equal(rc.phi().view(np.uint32), model.phi.view(np.uint32)), f"phi differs at iteration {it}"


def test_medium_full_corpus_bit_exact_and_ll_trace(orc):
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec = corpus.MEDIUM
    c = corpus.generate(spec)
    N = len(c["w"])
    say(f"medium full corpus: D={spec.D} N={N}")
    assert spec.D == 200000 and N > 20_000_000
    model = _model(orc, spec, c, seed=1)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=1, z=c["z"]) as rc:
        rc.schedule(1, 32)
        lls = [rc.loglik()]
        for it in (1, 2):
            rc.iterate_untimed(1)
            model.iterate()
            _same_state(rc, model, it)
            lls.append(rc.loglik())
            say(f"medium iteration {it} bit-exact")
        ll_c = orc.loglik(model.n, model.z, c["doc_off"], float(np.float32(spec.alpha)), float(np.float32(spec.beta)))
        assert abs(lls[-1] - ll_c) <= 1e-9 * abs(ll_c)
        for it in range(3, 201):
            rc.iterate_untimed(1)
            lls.append(rc.loglik())
    tail = np.array(lls[-10:])
    say(f"medium LL: start {lls[0]:.6e} it2 {lls[2]:.6e} it200 {lls[-1]:.6e}; last-10 spread "
        f"{np.abs(tail - tail.mean()).max() / abs(tail.mean()):.2e}")
    assert lls[-1] > lls[0]
    assert np.abs(tail - tail.mean()).max() <= 0.005 * abs(tail.mean())


def test_large_slice_bit_exact_full_phi(orc):
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec = corpus.LARGE.scaled(20000)
    c = corpus.generate(spec)
    say(f"large slice: D={spec.D} N={len(c['w'])} V={spec.V} K={spec.K}")
    model = _model(orc, spec, c, seed=2)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=2, z=c["z"]) as rc:
        rc.schedule(1, 32)  # the benchmark's schedule: one persistent launch, 32 wavefronts per CU
        for it in (1, 2):
            rc.iterate_untimed(1)
            model.iterate()
            _same_state(rc, model, it)
            p, a = rc.alias()
            assert np.array_equal(p.view(np.uint32), model.prob.view(np.uint32)) and np.array_equal(a, model.al)
            assert np.array_equal(rc.sigma_a().view(np.uint32), model.sigma_a.view(np.uint32))
            say(f"large slice iteration {it}: z, n_wk, phi, sigma_a, alias bit-exact")
        assert rc.visited() == model.visited


@pytest.mark.parametrize("name,sched,min_len,graph", [("large", 1, 0, False), ("batched", 3, 2048, True)])
def test_full_corpus_properties(orc, name, sched, min_len, graph):
    """A whole corpus -- large (D = 1M documents, about 328M tokens) under the benchmark's
    persistent schedule, batched (D = 2M, about 142M tokens) under the split schedule replayed as
    a hipGraph -- through properties that do not depend on its size: after two iterations n_wk equals the exact
    count of the (w, z) pairs, every topic lies below K, each Phi column sums to 1 within f32
    rounding, the alias tables are well formed; and the next iteration's Phi, sigma_a (alpha times
    the canonical row sums), alias tables and the z of the first 2,000 documents equal the
    oracle's, run from the same state with the full Phi."""
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec = corpus.SPECS[name]
    c = corpus.generate(spec)
    N = len(c["w"])
    say(f"{name} full corpus: D={spec.D} N={N} schedule={sched} graph={graph}")
    assert N > 100_000_000
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=spec.seed, z=c["z"]) as rc:
        rc.schedule(sched, 32, min_len)
        rc.use_graph(graph)
        rc.iterate_untimed(2)
        z = rc.z()
        n = rc.counts()
        assert int(z.max()) < spec.K
        want = np.bincount(c["w"].astype(np.int64) * spec.K + z, minlength=spec.V * spec.K)
        assert np.array_equal(n.reshape(-1), want.astype(np.uint32)), "n_wk differs from the count of (w, z)"
        del want
        say("n_wk equals the exact count of the (w, z) pairs")
        it = rc.iteration
        rc.iterate_untimed(1)
        rc.use_graph(False)
        phi = rc.phi()
        col = phi.astype(np.float64).sum(axis=0)
        used = col > 0
        assert np.all(np.abs(col[used] - 1.0) < 1e-3), "a Phi column does not sum to 1"
        sig = rc.sigma_a()
        prob, al = rc.alias()
        assert np.all((prob >= 0) & (prob <= 1)) and int(al.max()) < spec.K
        z_gpu = rc.z()
    tables = orc.poisson_tables(spec.beta)[2]
    pt, ts = orc.phi_draw(n, spec.beta, it + 1, spec.seed, tables)
    phi_c, sig_c, prob_c, al_c = orc.phi_finish(pt, ts, spec.alpha)
    assert np.array_equal(phi_c.view(np.uint32), phi.view(np.uint32)), "Phi differs from the oracle's"
    assert np.array_equal(sig_c.view(np.uint32), sig.view(np.uint32))
    assert np.array_equal(prob_c.view(np.uint32), prob.view(np.uint32)) and np.array_equal(al_c, al)
    say("Phi, sigma_a and alias tables of iteration 3 equal the oracle's")
    D_s = 2000
    off = c["doc_off"]
    n_s = int(off[D_s])
    z_s, _ = orc.sweep(z[:n_s], c["w"][:n_s], off[:D_s + 1], 0, it + 1, spec.seed, spec.K, phi_c, sig_c, prob_c, al_c,
                       threads=16)
    assert np.array_equal(z_s, z_gpu[:n_s]), "z of the first documents differs from the oracle's"
    say(f"z of the first {D_s} documents ({n_s} tokens) equal the oracle's")


@pytest.mark.parametrize("sched,graph", [(2, False), (2, True), (3, False), (3, True)])
def test_batched_slice_vs_oracle(orc, sched, graph):
    """sched 2: persistent + block-per-document (documents of at least 1,000 tokens, four
    wavefronts each); sched 3: persistent split (documents over 1,000 tokens in the long-document
    persistent launch, the rest in the register-variant one)."""
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    spec = corpus.BATCHED.scaled(20000)
    c = corpus.generate(spec)
    longest = int(c["d_len"].max())
    say(f"batched slice: D={spec.D} N={len(c['w'])} longest={longest} schedule={sched} graph={graph}")
    min_len = 1000
    assert (c["d_len"] > min_len).sum() > 10  # the long-document kernel gets work
    model = _model(orc, spec, c, seed=3)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=3, z=c["z"]) as rc:
        rc.schedule(sched, 16, min_len)
        rc.use_graph(graph)
        for it in (1, 2, 3):  # graph replays with eager reads in between
            rc.iterate_untimed(1)
            model.iterate()
            _same_state(rc, model, it, phi=(it == 3))
            say(f"batched slice iteration {it} bit-exact (schedule={sched}, graph={graph})")


def test_serving_large_model_vs_oracle(orc):
    from gpupolyaurnlda_amd import corpus, stages
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    from gpupolyaurnlda_amd.serve import FoldInServer, run_serial, run_stream
    spec = corpus.LARGE.scaled(20000)
    c = corpus.generate(spec)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=2, z=c["z"]) as rc:
        rc.iterate_untimed(3)
        phi = rc.phi()
        sig = rc.sigma_a()
        prob, al = rc.alias()
        d_len, arr, w = corpus.requests(phi, 30, median=300.0, sigma=1.2, lo=10, hi=5000, rate=2000.0, seed=4)
        say(f"serving: V={spec.V} K={spec.K}, 30 requests, lengths {int(d_len.min())}..{int(d_len.max())}, 50 sweeps")
        with FoldInServer(hbm_budget_bytes=16 << 30, sweeps=50, waves_per_cu=2, max_slots=1024) as srv:
            mixed = run_stream(srv, d_len, arr, w)
            alone = run_serial(srv, d_len, w)
    off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.uint64)
    ids = np.arange(len(d_len), dtype=np.uint32)
    zc, vc = orc.foldin(w, off, ids, 50, 2, spec.K, phi, sig, prob, al, threads=16)
    zs, vs = stages.foldin(w, off, ids, 50, 2, spec.K, phi, sig, prob, al)
    assert vs == vc and np.array_equal(zs, zc)
    want = [zc[off[i]:off[i + 1]] for i in range(len(d_len))]
    bad_mixed = [i for i in range(len(d_len)) if not np.array_equal(mixed["results"][i], want[i])]
    bad_alone = [i for i in range(len(d_len)) if not np.array_equal(alone["results"][i], want[i])]
    assert not bad_mixed and not bad_alone, f"mixed wrong {bad_mixed}, alone wrong {bad_alone}"
