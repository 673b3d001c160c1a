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
    out = subprocess.run([exe, "--alpha", str(spec.alpha), "--beta", str(spec.beta), "--K", str(spec.K),
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
