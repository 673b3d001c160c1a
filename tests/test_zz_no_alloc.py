"""The allocation-free fold-in (gpulda_foldin_noalloc) against the fold-in it is a variant of
(gpulda_stage_foldin), bit for bit, on two calls in a row -- the second smaller, so that anything
the first call left in the shared working memory would show -- and its size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_foldin_noalloc_matches_stage_foldin():
    from gpupolyaurnlda_amd import corpus, stages
    from gpupolyaurnlda_amd.resident import ResidentCorpus
    from gpupolyaurnlda_amd.serve import foldin_noalloc
    spec = corpus.MEDIUM.scaled(2000)
    c = corpus.generate(spec)
    with ResidentCorpus(c["w"], c["d_len"], spec.K, spec.V, spec.alpha, spec.beta, seed=3, z=c["z"]) as rc:
        rc.iterate_untimed(2)
        phi, sig = rc.phi(), rc.sigma_a()
        prob, al = rc.alias()
        for n_req, lo, hi in ((48, 10, 3000), (5, 10, 400)):
            d_len, _, w = corpus.requests(phi, n_req, median=200.0, sigma=1.0, lo=lo, hi=hi, seed=n_req)
            off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.uint64)
            ids = (np.arange(n_req) * 13 + 7).astype(np.uint32)
            z_na, v_na = foldin_noalloc(w, off, ids, 20)
            z_st, v_st = stages.foldin(w, off, ids, 20, 3, spec.K, phi, sig, prob, al)
            print(f"[no-alloc] {n_req} requests, {len(w)} tokens: identical={np.array_equal(z_na, z_st)}", flush=True)
            assert v_na == v_st and np.array_equal(z_na, z_st)
        # beyond the working memory (18 MB for 3M tokens): rejected with a message, and the next
        # call still works
        big = np.zeros(3_000_000, np.uint32)
        offs = np.arange(0, 3_000_001, 60000, dtype=np.uint64)
        with pytest.raises(RuntimeError, match="working memory"):
            foldin_noalloc(big, offs, np.arange(len(offs) - 1, dtype=np.uint32), 1)
        z_again, _ = foldin_noalloc(w, off, ids, 20)
        assert np.array_equal(z_again, z_st)
