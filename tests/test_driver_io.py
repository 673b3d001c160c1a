"""The C++ driver's host stages on the reference's on-disk formats, without a GPU: preprocessing of
Mallet-style input (src/preprocess.rs:12-72), the output writer (src/output.rs:27-62), the theta
export and the checkpoint checks that run before the device is touched.

Expected values are restated from the reference's Rust: its get_tokens known answer
(preprocess.rs:83-93), the vocabulary ordering rule, the stream layouts and the output line
format.  Initial topics are pinned to the Philox draws the oracle computes."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gpupolyaurnlda_amd", "gpulda_train")


def _paths(tmp):
    return {"z": f"{tmp}/z.bin", "w": f"{tmp}/w.bin", "d": f"{tmp}/d.bin", "k_d": f"{tmp}/k_d.bin",
            "c": f"{tmp}/c.txt"}


def _run(tmp, *args, check=True):
    p = _paths(tmp)
    cmd = [EXE, "--z_temp_file", p["z"], "--w_temp_file", p["w"], "--d_temp_file", p["d"],
           "--K_d_temp_file", p["k_d"], "--c_temp_file", p["c"], *map(str, args)]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    if check:
        assert out.returncode == 0, out.stdout + out.stderr
    return out


def _tokens(line):
    # get_tokens (preprocess.rs:12-14): third tab field, split_whitespace
    return line.split("\t")[2].split()


def _vocab(lines):
    # count_tokens (preprocess.rs:16-32): decreasing count; ties here by first appearance
    counts, order = {}, []
    for ln in lines:
        for t in _tokens(ln):
            if t not in counts:
                counts[t] = 0
                order.append(t)
            counts[t] += 1
    ranked = sorted(order, key=lambda t: -counts[t])  # stable
    return ranked, counts


def _init_topics(orc, L, doc, K, seed):
    ctr = np.zeros((L, 4), np.uint32)
    ctr[:, 0] = np.arange(L)
    ctr[:, 1] = doc
    ctr[:, 3] = 1 << 24  # INIT_Z tag, sub 0
    key = np.tile(np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32), (L, 1))
    r = orc.philox(ctr, key)
    return ((r[:, 0].astype(np.uint64) * K) >> 32).astype(np.uint32)


def test_get_tokens_known_answer(tmp_path):
    # the reference's own test: "docno:1\tX\thi alex and Kunal\n" has 4 tokens, "hi" .. "Kunal"
    tmp = str(tmp_path)
    with open(f"{tmp}/in.txt", "w") as f:
        f.write("docno:1\tX\thi alex and Kunal\n")
    _run(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", "--K", 5)
    names = [ln.split("\t")[0] for ln in open(f"{tmp}/c.txt").read().splitlines()]
    w = np.fromfile(f"{tmp}/w.bin", "<u4")
    assert np.fromfile(f"{tmp}/d.bin", "<u4").tolist() == [4]
    assert [names[i] for i in w] == ["hi", "alex", "and", "Kunal"]


def test_preprocess_restates_reference(tmp_path, orc):
    tmp = str(tmp_path)
    rng = np.random.default_rng(3)
    words = [f"w{i}" for i in range(40)]
    lines = []
    for d in range(60):
        L = int(rng.integers(0, 30))
        toks = [words[int(min(39, rng.zipf(1.5) - 1))] for _ in range(L)]
        lines.append(f"doc{d}\tlabel\t" + " ".join(toks))
    # separators the reference's split_whitespace accepts, and a CRLF line ending
    lines.append("docA\tX\t  w1 w2 w1\tignored field")
    lines.append("docB\tX\tw3 w3\r")
    with open(f"{tmp}/in.txt", "w", encoding="utf-8", newline="") as f:
        f.write("\n".join(lines) + "\n")
    K, seed = 13, 0x1234_0000_0007
    out = _run(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", "--K", K, "--seed", seed)
    assert "preprocessing" in out.stdout

    clean = [ln.rstrip("\r") for ln in lines]
    ranked, counts = _vocab(clean)
    c_lines = open(f"{tmp}/c.txt", encoding="utf-8").read().splitlines()
    assert c_lines == [f"{t}\t{i}\t{counts[t]}" for i, t in enumerate(ranked)]
    ids = {t: i for i, t in enumerate(ranked)}

    z = np.fromfile(f"{tmp}/z.bin", "<u4")
    w = np.fromfile(f"{tmp}/w.bin", "<u4")
    d = np.fromfile(f"{tmp}/d.bin", "<u4")
    k_d = np.fromfile(f"{tmp}/k_d.bin", "<u4")
    assert len(d) == len(k_d) == len(clean)
    pos = 0
    for doc, ln in enumerate(clean):
        toks = _tokens(ln)
        assert d[doc] == len(toks)
        assert w[pos:pos + len(toks)].tolist() == [ids[t] for t in toks]
        z0 = _init_topics(orc, len(toks), doc, K, seed)
        assert np.array_equal(z[pos:pos + len(toks)], z0)
        assert k_d[doc] == len(np.unique(z0))
        pos += len(toks)
    assert pos == len(w) == len(z)
    assert _tokens(clean[-2]) == ["w1", "w2", "w1"]


def test_preprocess_rejects_line_without_text_field(tmp_path):
    tmp = str(tmp_path)
    with open(f"{tmp}/in.txt", "w") as f:
        f.write("a\tb\tx y\nonly\ttwo\n")
    out = _run(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", check=False)
    assert out.returncode != 0 and "line 2" in out.stderr


def test_output_and_theta_formats(tmp_path):
    tmp = str(tmp_path)
    with open(f"{tmp}/in.txt", "w") as f:
        f.write("a\tx\tred blue red\nb\tx\t\nc\tx\tgreen red\n")
    _run(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", "--K", 4)
    # topics as a trained run would have left them
    np.array([3, 0, 3, 1, 2], "<u4").tofile(f"{tmp}/z.bin")
    _run(tmp, "--stages", "output", "--output", f"{tmp}/out.txt", "--theta_file", f"{tmp}/theta.txt", "--K", 4)
    # output.rs:51 writes "token:topic\t" per token and a newline per document
    assert open(f"{tmp}/out.txt").read() == "red:3\tblue:0\tred:3\t\n\ngreen:1\tred:2\t\n"
    assert open(f"{tmp}/theta.txt").read() == "0:1 3:2\n\n1:1 2:1\n"


def test_output_rejects_unknown_word(tmp_path):
    tmp = str(tmp_path)
    with open(f"{tmp}/in.txt", "w") as f:
        f.write("a\tx\tred blue\n")
    _run(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", "--K", 4)
    np.array([0, 7], "<u4").tofile(f"{tmp}/w.bin")
    out = _run(tmp, "--stages", "output", "--output", f"{tmp}/out.txt", check=False)
    assert out.returncode != 0 and "word id 7" in out.stderr


def test_resume_checks_run_before_the_device(tmp_path):
    tmp = str(tmp_path)
    with open(f"{tmp}/in.txt", "w") as f:
        f.write("a\tx\tred blue red\nb\tx\tgreen\n")
    _run(tmp, "--stages", "preprocess", "--input", f"{tmp}/in.txt", "--K", 4)
    ck = f"{tmp}/ck.txt"
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "no checkpoint" in out.stderr
    with open(ck, "w") as f:
        f.write("something else\n")
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "not a checkpoint" in out.stderr
    with open(ck, "w") as f:
        f.write("gpulda-checkpoint 1\niterations 2\nseed 0\nK 4\nV 3\ntokens 4\ndocuments 2\n"
                "alpha_bits 3dcccccd\nbeta_bits 3dcccccd\nbuffer_size 100\nmax_D 10\nz_hash 0000000000000001\n")
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "z.bin changed" in out.stderr
    out = _run(tmp, "--stages", "train", "--K", 5, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "other parameters" in out.stderr
    # a record naming a z snapshot: the snapshot must exist, hold every token and match the hash
    rec = open(ck).read().replace("gpulda-checkpoint 1\n", "gpulda-checkpoint 1\nz_snapshot ck.txt.z2\n")
    with open(ck, "w") as f:
        f.write(rec)
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "snapshot" in out.stderr and "missing" in out.stderr
    np.array([1, 2, 3], "<u4").tofile(f"{tmp}/ck.txt.z2")
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "does not hold 4 topics" in out.stderr
    np.array([1, 2, 3, 0], "<u4").tofile(f"{tmp}/ck.txt.z2")
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "does not match its record" in out.stderr
    with open(ck, "w") as f:
        f.write(rec.replace("z_snapshot ck.txt.z2", "z_snapshot ../ck.txt.z2"))
    out = _run(tmp, "--stages", "train", "--K", 4, "--resume", "--checkpoint_file", ck, check=False)
    assert out.returncode != 0 and "bad checkpoint snapshot name" in out.stderr


def test_unknown_stage_and_flag_rejected(tmp_path):
    out = _run(str(tmp_path), "--stages", "train,bogus", check=False)
    assert out.returncode != 0 and "unknown stage" in out.stderr
    out = _run(str(tmp_path), "--no_such_flag", 1, check=False)
    assert out.returncode != 0 and "unknown flag" in out.stderr
    out = _run(str(tmp_path), "--n_buffers", 0, check=False)
    assert out.returncode != 0 and "n_buffers" in out.stderr
