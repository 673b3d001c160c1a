// train_driver.cpp -- C++ host that runs the reference's program (src/main.rs:17-21: preprocess,
// train, output) against libGPUPolyaUrnLDA, driving the library exactly as the reference's Rust
// host does (src/train.rs:60-105), over the reference's on-disk formats:
//   c.txt      token \t id \t count per line, ids 0..V-1 in order   (preprocess.rs:27)
//   z.bin      u32 little-endian topic indicators, one per token     (read-write, mmap)
//   w.bin      u32 word ids, one per token                           (read-only, mmap)
//   d.bin      u32 document lengths, one per document                (read-only, mmap)
//   k_d.bin    u32 distinct-topic hints, one per document            (read-only, mmap)
//
// Flags restate src/args.rs:17-59 (names and defaults).  The reference ignores --seed
// (args.rs:36-37); here it is passed to the library with gpulda_set_seed.  Added flags:
//   --stages LIST          which of preprocess,train,output run (default: all three, in order)
//   --phi_file PATH        after training, Phi of the last iteration (V x K f32, word-major)
//   --theta_file PATH      after training, per-document topic counts (`k:n` pairs per line)
//   --checkpoint_file PATH after every --checkpoint_every iterations, a snapshot of z (PATH.z<it>,
//                          written and flushed first) and a record of the completed iterations,
//                          the parameters and the snapshot's hash (written, flushed, renamed)
//   --resume               continue from --checkpoint_file: z.bin is restored from the snapshot
//                          (the iterations after the checkpoint rewrite z.bin in place, so a run
//                          stopped anywhere leaves it part-way), n_wk is rebuilt from it and the
//                          iteration counter restored, so the run continues the uninterrupted
//                          chain bit for bit (any buffer_size / max_D)
//   --stop_after_buffers N stop the run once N buffers have been synchronised (all device work
//                          finished, z.bin flushed, no checkpoint written): a stop in the middle
//                          of an iteration, for tests of --resume
//   --n_buffers N          buffers in flight (default 1, as train.rs): buffer i + 1 is packed and
//                          submitted while buffer i is swept, each synchronised before reuse;
//                          results do not depend on N
//
// Buffer packing restates fill_buffer (train.rs:19-45), including its panic on a document larger
// than buffer_size.  End of input is the end of d.bin; a zero-length document does not end the
// iteration (the reference reads a 0 length as end-of-file, train.rs:24,30).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "corpus_io.h"
#include "gpulda_abi.h"
#include "gpulda_ext.h"

namespace {

using gpulda_io::die;

enum : unsigned { kPreprocess = 1, kTrain = 2, kOutput = 4 };

struct Options {
  float alpha = 0.1f;
  float beta = 0.1f;
  uint32_t K = 100;
  uint32_t buffer_size = 33554432;
  uint32_t max_D = 16384;
  uint32_t n_mc = 1;
  uint64_t seed = 0;
  std::string input = "data/small.txt";
  std::string output = "output/small.txt";
  gpulda_io::TempPaths tmp{"temp/z.bin", "temp/w.bin", "temp/d.bin", "temp/k_d.bin", "temp/c.txt"};
  unsigned stages = kPreprocess | kTrain | kOutput;
  std::string phi_file, theta_file, checkpoint_file;
  uint32_t checkpoint_every = 1;
  bool resume = false;
  uint32_t n_buffers = 1;
  uint64_t stop_after_buffers = 0;
};

unsigned parse_stages(const std::string& v) {
  unsigned st = 0;
  size_t i = 0;
  while (i <= v.size()) {
    size_t j = v.find(',', i);
    if (j == std::string::npos) j = v.size();
    const std::string name = v.substr(i, j - i);
    if (name == "preprocess") st |= kPreprocess;
    else if (name == "train") st |= kTrain;
    else if (name == "output") st |= kOutput;
    else die("unknown stage '" + name + "' (preprocess, train, output)");
    i = j + 1;
  }
  return st;
}

uint32_t parse_u32(const std::string& flag, const std::string& v) {
  char* end = nullptr;
  errno = 0;
  unsigned long long x = strtoull(v.c_str(), &end, 10);
  if (errno || end == v.c_str() || *end || x > 0xFFFFFFFFull) die("bad value for " + flag + ": " + v);
  return (uint32_t)x;
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value for " + a);
      return argv[++i];
    };
    if (a == "--alpha") o.alpha = std::stof(val());
    else if (a == "--beta") o.beta = std::stof(val());
    else if (a == "--K") o.K = parse_u32(a, val());
    else if (a == "--buffer_size") o.buffer_size = parse_u32(a, val());
    else if (a == "--max_D") o.max_D = parse_u32(a, val());
    else if (a == "--n_mc") o.n_mc = parse_u32(a, val());
    else if (a == "--seed") o.seed = std::stoull(val());
    else if (a == "--input") o.input = val();
    else if (a == "--output") o.output = val();
    else if (a == "--z_temp_file") o.tmp.z = val();
    else if (a == "--w_temp_file") o.tmp.w = val();
    else if (a == "--d_temp_file") o.tmp.d = val();
    else if (a == "--K_d_temp_file") o.tmp.k_d = val();
    else if (a == "--c_temp_file") o.tmp.c = val();
    else if (a == "--stages") o.stages = parse_stages(val());
    else if (a == "--phi_file") o.phi_file = val();
    else if (a == "--theta_file") o.theta_file = val();
    else if (a == "--checkpoint_file") o.checkpoint_file = val();
    else if (a == "--checkpoint_every") o.checkpoint_every = parse_u32(a, val());
    else if (a == "--resume") o.resume = true;
    else if (a == "--n_buffers") o.n_buffers = parse_u32(a, val());
    else if (a == "--stop_after_buffers") o.stop_after_buffers = parse_u32(a, val());
    else die("unknown flag " + a);
  }
  if (o.resume && o.checkpoint_file.empty()) die("--resume needs --checkpoint_file");
  if (o.checkpoint_every == 0) die("--checkpoint_every must be positive");
  if (o.n_buffers == 0 || o.n_buffers > 64) die("--n_buffers must be in 1..64");
  return o;
}

struct Map {
  uint32_t* p = nullptr;
  size_t n = 0;  // u32 elements
  size_t bytes = 0;
};

Map map_file(const std::string& path, bool writable) {
  int fd = open(path.c_str(), writable ? O_RDWR : O_RDONLY);
  if (fd < 0) die("cannot open " + path);
  struct stat st;
  if (fstat(fd, &st) != 0) die("cannot stat " + path);
  Map m;
  m.bytes = (size_t)st.st_size;
  m.n = m.bytes / sizeof(uint32_t);
  if (m.bytes) {
    void* p = mmap(nullptr, m.bytes, writable ? PROT_READ | PROT_WRITE : PROT_READ, MAP_SHARED, fd, 0);
    if (p == MAP_FAILED) die("cannot mmap " + path);
    m.p = (uint32_t*)p;
  }
  close(fd);
  return m;
}

uint32_t f32_bits(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  return b;
}

void ext_check(int rc, const char* what) {
  if (rc != 0) die(std::string(what) + ": " + gpulda_last_error());
}

gpulda_io::Checkpoint checkpoint_of(const Options& o, uint32_t V, const Map& z, const Map& d, uint32_t iterations) {
  gpulda_io::Checkpoint c;
  c.iterations = iterations;
  c.seed = o.seed;
  c.K = o.K;
  c.V = V;
  c.tokens = z.n;
  c.docs = d.n;
  c.alpha_bits = f32_bits(o.alpha);
  c.beta_bits = f32_bits(o.beta);
  c.buffer_size = o.buffer_size;
  c.max_D = o.max_D;
  return c;
}

// Snapshot of z after `iterations` iterations, then the record that names it; the previous
// snapshot is removed once the new record is in place.
void write_checkpoint(const Options& o, uint32_t V, const Map& z, const Map& d, uint32_t iterations,
                      std::string* prev_snapshot) {
  gpulda_io::Checkpoint c = checkpoint_of(o, V, z, d, iterations);
  c.z_hash = gpulda_io::bytes_hash(z.p, z.bytes);
  c.z_snapshot = gpulda_io::snapshot_name(o.checkpoint_file, iterations);
  gpulda_io::write_snapshot(gpulda_io::snapshot_file(o.checkpoint_file, c.z_snapshot), z.p, z.bytes);
  gpulda_io::write_checkpoint(o.checkpoint_file, c);
  if (!prev_snapshot->empty() && *prev_snapshot != c.z_snapshot)
    (void)unlink(gpulda_io::snapshot_file(o.checkpoint_file, *prev_snapshot).c_str());
  *prev_snapshot = c.z_snapshot;
}

// train.rs:60-105, plus resume, checkpoints and exports.
void train(const Options& o) {
  printf("training\n");
  std::vector<uint32_t> C = gpulda_io::token_counts(o.tmp.c);
  Args args;
  args.alpha = o.alpha;
  args.beta = o.beta;
  args.K = o.K;
  args.V = (uint32_t)C.size();
  args.C = C.data();
  args.buffer_size = o.buffer_size;
  args.max_D = o.max_D;
  Map z = map_file(o.tmp.z, true);
  Map w = map_file(o.tmp.w, false);
  Map d = map_file(o.tmp.d, false);
  Map k_d = map_file(o.tmp.k_d, false);
  if (z.n != w.n) die("z.bin and w.bin hold different token counts");
  if (d.n != k_d.n) die("d.bin and k_d.bin hold different document counts");

  // the checkpoint is checked before the device is touched
  uint32_t start = 0;
  std::vector<uint32_t> n_resume;
  std::string snapshot;  // snapshot named by the newest record of this run
  if (o.resume) {
    gpulda_io::Checkpoint ck;
    if (!gpulda_io::read_checkpoint(o.checkpoint_file, &ck)) die("no checkpoint at " + o.checkpoint_file);
    const gpulda_io::Checkpoint now = checkpoint_of(o, args.V, z, d, ck.iterations);
    if (ck.seed != now.seed || ck.K != now.K || ck.V != now.V || ck.tokens != now.tokens || ck.docs != now.docs ||
        ck.alpha_bits != now.alpha_bits || ck.beta_bits != now.beta_bits)
      die("checkpoint " + o.checkpoint_file + " was written for other parameters or another corpus");
    if (ck.z_snapshot.empty()) {
      // a record without a snapshot: z.bin itself must still hold the checkpointed topics
      if (ck.z_hash != gpulda_io::file_hash(o.tmp.z))
        die("z.bin changed after checkpoint " + o.checkpoint_file + " was written");
    } else {
      const std::vector<unsigned char> snap = gpulda_io::read_snapshot(ck, o.checkpoint_file);
      if (z.bytes && memcmp(z.p, snap.data(), z.bytes) != 0) {
        memcpy(z.p, snap.data(), z.bytes);
        if (msync(z.p, z.bytes, MS_SYNC) != 0) die("cannot flush " + o.tmp.z);
        printf("restored z.bin from %s\n", ck.z_snapshot.c_str());
      }
      snapshot = ck.z_snapshot;
    }
    n_resume = gpulda_io::counts_from_files(o.tmp, args.V, o.K);
    start = ck.iterations;
  }

  std::vector<Buffer> buffers(o.n_buffers);
  memset(buffers.data(), 0, buffers.size() * sizeof(Buffer));
  std::vector<char> in_flight(o.n_buffers, 0);
  ext_check(gpulda_set_seed(o.seed), "set_seed");
  initialize(&args, buffers.data(), o.n_buffers);
  if (o.resume) {
    ext_check(gpulda_set_counts(n_resume.data()), "set_counts");
    ext_check(gpulda_set_iteration(start), "set_iteration");
    std::vector<uint32_t>().swap(n_resume);
    printf("resuming after iteration %u\n", start);
  }

  uint64_t synced = 0;  // buffers synchronised in this run (--stop_after_buffers)
  auto synchronise = [&](Buffer* b) {
    sync_buffer(b);
    if (o.stop_after_buffers && ++synced == o.stop_after_buffers) {
      for (uint32_t i = 0; i < o.n_buffers; ++i)
        if (in_flight[i] && &buffers[i] != b) sync_buffer(&buffers[i]);
      if (z.bytes) msync(z.p, z.bytes, MS_SYNC);
      printf("stopped after %llu buffers\n", (unsigned long long)synced);
      fflush(stdout);
      exit(3);
    }
  };
  for (uint32_t it = start; it < o.n_mc; ++it) {
    sample_phi();
    size_t doc = 0, tok = 0;
    for (uint32_t slot = 0;; slot = (slot + 1) % o.n_buffers) {
      // fill_buffer (train.rs:19-45)
      uint32_t n_docs = 0, n_tokens = 0;
      while (doc + n_docs < d.n) {
        uint32_t size = d.p[doc + n_docs];
        if (size > o.buffer_size) die("document size exceeded buffer size: " + std::to_string(size));
        if (n_docs + 1 > o.max_D || (uint64_t)n_tokens + size > o.buffer_size) break;
        ++n_docs;
        n_tokens += size;
      }
      if (n_docs == 0) break;
      if (tok + n_tokens > z.n) die("d.bin lengths exceed the tokens in w.bin");
      Buffer& b = buffers[slot];
      if (in_flight[slot]) {
        in_flight[slot] = 0;
        synchronise(&b);
      }
      b.z = z.p + tok;
      b.w = w.p + tok;
      b.d = d.p + doc;
      b.K_d = k_d.p + doc;
      b.n_docs = n_docs;
      b.n_tokens = n_tokens;
      sample_z_async(&b);
      in_flight[slot] = 1;
      if (o.n_buffers == 1) {
        in_flight[slot] = 0;
        synchronise(&b);
      }
      // empty_buffer (train.rs:47-58)
      doc += n_docs;
      tok += n_tokens;
    }
    for (uint32_t i = 0; i < o.n_buffers; ++i)
      if (in_flight[i]) {
        in_flight[i] = 0;
        synchronise(&buffers[i]);
      }
    if (z.bytes) msync(z.p, z.bytes, MS_SYNC);  // z.get_mut().flush() (train.rs:99)
    printf("finished iteration: %u\n", it);
    if (!o.checkpoint_file.empty() && ((it + 1) % o.checkpoint_every == 0 || it + 1 == o.n_mc))
      write_checkpoint(o, args.V, z, d, it + 1, &snapshot);
  }
  if (!o.phi_file.empty()) {
    if (gpulda_iteration() == 0) die("--phi_file: no iteration has run");
    std::vector<float> phi((size_t)args.V * o.K);
    ext_check(gpulda_copy_phi(phi.data()), "copy_phi");
    gpulda_io::write_phi(o.phi_file, phi.data(), args.V, o.K);
  }
  cleanup(buffers.data(), o.n_buffers);
  for (Map* m : {&z, &w, &d, &k_d})
    if (m->p) munmap(m->p, m->bytes);
}

}  // namespace

int main(int argc, char** argv) {
  Options o = parse(argc, argv);
  if (o.stages & kPreprocess) {
    printf("preprocessing\n");
    gpulda_io::CorpusShape sh = gpulda_io::preprocess(o.input, o.tmp, o.K, o.seed);
    printf("documents %llu tokens %llu vocabulary %u\n", (unsigned long long)sh.docs, (unsigned long long)sh.tokens,
           sh.vocab);
  }
  if (o.stages & kTrain) train(o);
  if (!o.theta_file.empty()) gpulda_io::write_theta(o.theta_file, o.tmp, o.K);
  if (o.stages & kOutput) {
    printf("writing output\n");
    gpulda_io::write_output(o.output, o.tmp);
  }
  return 0;
}
