"""The C++ train.rs-equivalent driver (gpupolyaurnlda_amd/gpulda_train) on the reference's on-disk
formats (c.txt, z.bin, w.bin, d.bin, k_d.bin), checked against the oracle replaying the same
buffer sequence."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write_formats(tmp, c, V):
    with open(os.path.join(tmp, "c.txt"), "w") as f:
        for v in range(V):
            f.write(f"tok{v}\t{v}\t{int(c['C'][v])}\n")
    c["z"].astype("<u4").tofile(os.path.join(tmp, "z.bin"))
    c["w"].astype("<u4").tofile(os.path.join(tmp, "w.bin"))
    c["d_len"].astype("<u4").tofile(os.path.join(tmp, "d.bin"))
    k_d = np.array([len(np.unique(c["z"][c["doc_off"][i]:c["doc_off"][i + 1]])) for i in range(len(c["d_len"]))],
                   np.uint32)
    k_d.astype("<u4").tofile(os.path.join(tmp, "k_d.bin"))


def test_driver_matches_oracle(tmp_path, orc):
    from gpupolyaurnlda_amd import corpus
    from gpupolyaurnlda_amd.train import fill_buffer
    spec = corpus.TINY.scaled(300)
    c = corpus.generate(spec)
    tmp = str(tmp_path)
    _write_formats(tmp, c, spec.V)
    n_mc, buffer_size, max_D, seed = 3, 2048, 32, 7
    exe = os.path.join(ROOT, "gpupolyaurnlda_amd", "gpulda_train")
    out = subprocess.run([exe, "--stages", "train", "--alpha", str(spec.alpha), "--beta", str(spec.beta), "--K", str(spec.K),
                          "--buffer_size", str(buffer_size), "--max_D", str(max_D), "--n_mc", str(n_mc),
                          "--seed", str(seed),
                          "--z_temp_file", f"{tmp}/z.bin", "--w_temp_file", f"{tmp}/w.bin",
                          "--d_temp_file", f"{tmp}/d.bin", "--K_d_temp_file", f"{tmp}/k_d.bin",
                          "--c_temp_file", f"{tmp}/c.txt"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "finished iteration: 2" in out.stdout
    z_gpu = np.fromfile(f"{tmp}/z.bin", dtype="<u4")

    model = orc.Model(c["w"], c["doc_off"], c["z"], spec.V, spec.K, spec.alpha, spec.beta, seed=seed, counts="C")
    d = c["d_len"]
    for _ in range(n_mc):
        model.phi_step()
        doc = 0
        while True:
            nd, _ = fill_buffer(d, doc, buffer_size, max_D)
            if nd == 0:
                break
            model.z_step(doc_base=doc, lo=doc, hi=doc + nd)
            doc += nd
        model.n = orc.count(model.w, model.z, spec.V, spec.K)
    assert np.array_equal(z_gpu, model.z)


def _replay(orc, w, d_len, z0, V, K, alpha, beta, seed, n_mc, buffer_size, max_D):
    """The oracle replaying train.rs's buffer sequence from the given start."""
    from gpupolyaurnlda_amd.train import fill_buffer
    doc_off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.uint64)
    model = orc.Model(w, doc_off, z0, V, K, alpha, beta, seed=seed, counts="C")
    for _ in range(n_mc):
        model.phi_step()
        doc = 0
        while True:
            nd, _ = fill_buffer(d_len, doc, buffer_size, max_D)
            if nd == 0:
                break
            model.z_step(doc_base=doc, lo=doc, hi=doc + nd)
            doc += nd
        model.n = orc.count(model.w, model.z, V, K)
    return model


def _mallet_file(path, c):
    """A Mallet-style input whose tokens are the synthetic word ids."""
    with open(path, "w") as f:
        for i in range(len(c["d_len"])):
            toks = c["w"][c["doc_off"][i]:c["doc_off"][i + 1]]
            f.write(f"doc{i}\tlabel\t" + " ".join(f"t{v}" for v in toks) + "\n")


def _driver(tmp, *args):
    exe = os.path.join(ROOT, "gpupolyaurnlda_amd", "gpulda_train")
    cmd = [exe, "--z_temp_file", f"{tmp}/z.bin", "--w_temp_file", f"{tmp}/w.bin", "--d_temp_file", f"{tmp}/d.bin",
           "--K_d_temp_file", f"{tmp}/k_d.bin", "--c_temp_file", f"{tmp}/c.txt", *map(str, args)]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def test_driver_full_pipeline(tmp_path, orc):
    """preprocess -> train -> output as src/main.rs runs them, plus the Phi and theta exports:
    the trained z equals the oracle's replay from the preprocessed files, Phi is the oracle's last
    Phi bit for bit, and the output and theta files restate that z."""
    from gpupolyaurnlda_amd import corpus
    spec = corpus.TINY.scaled(200)
    c = corpus.generate(spec)
    tmp = str(tmp_path)
    _mallet_file(f"{tmp}/in.txt", c)
    K, seed, n_mc, buffer_size, max_D = spec.K, 11, 3, 4096, 64
    common = ("--K", K, "--alpha", spec.alpha, "--beta", spec.beta, "--seed", seed)
    _driver(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", *common)
    z0 = np.fromfile(f"{tmp}/z.bin", "<u4")
    w = np.fromfile(f"{tmp}/w.bin", "<u4")
    d_len = np.fromfile(f"{tmp}/d.bin", "<u4")
    names = [ln.split("\t")[0] for ln in open(f"{tmp}/c.txt").read().splitlines()]
    V = len(names)
    out = _driver(tmp, "--stages", "train,output", "--output", f"{tmp}/out.txt", "--phi_file", f"{tmp}/phi.bin",
                  "--theta_file", f"{tmp}/theta.txt", "--n_mc", n_mc, "--buffer_size", buffer_size,
                  "--max_D", max_D, *common)
    assert "finished iteration: 2" in out and "writing output" in out
    z = np.fromfile(f"{tmp}/z.bin", "<u4")
    model = _replay(orc, w, d_len, z0, V, K, spec.alpha, spec.beta, seed, n_mc, buffer_size, max_D)
    assert np.array_equal(z, model.z)
    phi = np.fromfile(f"{tmp}/phi.bin", "<f4").reshape(V, K)
    assert np.array_equal(phi.view(np.uint32), np.asarray(model.phi, np.float32).view(np.uint32))
    off = np.concatenate([[0], np.cumsum(d_len)]).astype(np.int64)
    want_out, want_theta = [], []
    for i in range(len(d_len)):
        zz, ww = z[off[i]:off[i + 1]], w[off[i]:off[i + 1]]
        want_out.append("".join(f"{names[v]}:{k}\t" for v, k in zip(ww, zz)))
        ks, ns = np.unique(zz, return_counts=True)
        want_theta.append(" ".join(f"{k}:{n}" for k, n in zip(ks, ns)))
    assert open(f"{tmp}/out.txt").read() == "".join(s + "\n" for s in want_out)
    assert open(f"{tmp}/theta.txt").read() == "".join(s + "\n" for s in want_theta)


def _buffers_per_iteration(d_len, buffer_size, max_D):
    from gpupolyaurnlda_amd.train import fill_buffer
    n, doc = 0, 0
    while True:
        nd, _ = fill_buffer(d_len, doc, buffer_size, max_D)
        if nd == 0:
            return n
        n += 1
        doc += nd


def test_driver_checkpoint_resume(tmp_path, orc):
    """Resume from a checkpoint, after a stop at an iteration boundary and after a stop in the
    middle of an iteration (z.bin then holds part of the next iteration's topics; the resumed run
    restores it from the checkpoint's snapshot).  Both resumed runs, with another buffer size, end
    with the z of the oracle replaying four uninterrupted iterations."""
    import shutil
    from gpupolyaurnlda_amd import corpus
    spec = corpus.TINY.scaled(300)
    c = corpus.generate(spec)
    base, b, m = (str(tmp_path / n) for n in ("base", "b", "m"))
    os.makedirs(base)
    _mallet_file(f"{base}/in.txt", c)
    seed, n_mc = 5, 4
    common = ("--K", spec.K, "--alpha", spec.alpha, "--beta", spec.beta, "--seed", seed)
    _driver(base, "--stages", "preprocess", "--input", f"{base}/in.txt", *common)
    shutil.copytree(base, b)
    shutil.copytree(base, m)
    z0 = np.fromfile(f"{base}/z.bin", "<u4")
    w = np.fromfile(f"{base}/w.bin", "<u4")
    d_len = np.fromfile(f"{base}/d.bin", "<u4")
    V = len(open(f"{base}/c.txt").read().splitlines())
    want = _replay(orc, w, d_len, z0, V, spec.K, spec.alpha, spec.beta, seed, n_mc, 3000, 40).z
    assert not np.array_equal(want, z0)
    # stop at the boundary after iteration 2, resume to 4
    ck = f"{b}/ck.txt"
    _driver(b, "--stages", "train", "--n_mc", 2, "--buffer_size", 3000, "--max_D", 40, "--checkpoint_file", ck,
            *common)
    assert "iterations 2" in open(ck).read() and os.path.exists(f"{ck}.z2")
    out = _driver(b, "--stages", "train", "--n_mc", n_mc, "--buffer_size", 1000, "--max_D", 17,
                  "--checkpoint_file", ck, "--resume", *common)
    assert "resuming after iteration 2" in out and "finished iteration: 3" in out
    assert "finished iteration: 1" not in out and "restored" not in out
    assert "iterations 4" in open(ck).read() and os.path.exists(f"{ck}.z4") and not os.path.exists(f"{ck}.z2")
    assert np.array_equal(np.fromfile(f"{b}/z.bin", "<u4"), want)
    # stop in the middle of the third iteration (checkpoints every 2 iterations), resume to 4
    per_it = _buffers_per_iteration(d_len, 3000, 40)
    assert per_it >= 4
    ck = f"{m}/ck.txt"
    exe = os.path.join(ROOT, "gpupolyaurnlda_amd", "gpulda_train")
    cmd = [exe, "--z_temp_file", f"{m}/z.bin", "--w_temp_file", f"{m}/w.bin", "--d_temp_file", f"{m}/d.bin",
           "--K_d_temp_file", f"{m}/k_d.bin", "--c_temp_file", f"{m}/c.txt", "--stages", "train", "--n_mc", n_mc,
           "--buffer_size", 3000, "--max_D", 40, "--checkpoint_file", ck, "--checkpoint_every", 2,
           "--stop_after_buffers", 2 * per_it + per_it // 2, *common]
    stop = subprocess.run([str(x) for x in cmd], capture_output=True, text=True, timeout=120)
    assert stop.returncode == 3 and "stopped after" in stop.stdout, stop.stdout + stop.stderr
    assert "iterations 2" in open(ck).read()
    z_stopped = np.fromfile(f"{m}/z.bin", "<u4")
    z_ck = np.fromfile(f"{ck}.z2", "<u4")
    assert not np.array_equal(z_stopped, z_ck)  # z.bin moved on past the checkpoint
    out = _driver(m, "--stages", "train", "--n_mc", n_mc, "--buffer_size", 1000, "--max_D", 17,
                  "--checkpoint_file", ck, "--resume", *common)
    assert "restored z.bin from ck.txt.z2" in out and "resuming after iteration 2" in out
    assert np.array_equal(np.fromfile(f"{m}/z.bin", "<u4"), want)


def test_driver_buffers_in_flight(tmp_path, orc):
    """Three buffers in flight (each synchronised before reuse) give the z of the oracle replaying
    train.rs's one-buffer loop."""
    import shutil
    from gpupolyaurnlda_amd import corpus
    spec = corpus.TINY.scaled(400)
    c = corpus.generate(spec)
    base, b = (str(tmp_path / n) for n in ("base", "b"))
    os.makedirs(base)
    _mallet_file(f"{base}/in.txt", c)
    seed, n_mc, buffer_size, max_D = 9, 3, 1500, 30
    common = ("--K", spec.K, "--alpha", spec.alpha, "--beta", spec.beta, "--seed", seed, "--n_mc", n_mc,
              "--buffer_size", buffer_size, "--max_D", max_D)
    _driver(base, "--stages", "preprocess", "--input", f"{base}/in.txt", *common)
    shutil.copytree(base, b)
    _driver(b, "--stages", "train", "--n_buffers", 3, *common)
    z0 = np.fromfile(f"{base}/z.bin", "<u4")
    w = np.fromfile(f"{base}/w.bin", "<u4")
    d_len = np.fromfile(f"{base}/d.bin", "<u4")
    V = len(open(f"{base}/c.txt").read().splitlines())
    assert _buffers_per_iteration(d_len, buffer_size, max_D) > 3
    want = _replay(orc, w, d_len, z0, V, spec.K, spec.alpha, spec.beta, seed, n_mc, buffer_size, max_D).z
    assert not np.array_equal(want, z0)
    assert np.array_equal(np.fromfile(f"{b}/z.bin", "<u4"), want)
