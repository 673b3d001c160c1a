// train_driver.cpp -- C++ host that drives libGPUPolyaUrnLDA exactly as the reference's Rust
// host does (src/train.rs:60-105), over the reference's on-disk formats:
//   c.txt      token \t id \t count per line, ids 0..V-1 in order   (preprocess.rs:27)
//   z.bin      u32 little-endian topic indicators, one per token     (read-write, mmap)
//   w.bin      u32 word ids, one per token                           (read-only, mmap)
//   d.bin      u32 document lengths, one per document                (read-only, mmap)
//   k_d.bin    u32 distinct-topic hints, one per document            (read-only, mmap)
//
// Flags restate src/args.rs:17-59 (names and defaults).  The reference ignores --seed
// (args.rs:36-37); here it is passed to the library with gpulda_set_seed.
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
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gpulda_abi.h"
#include "gpulda_ext.h"

namespace {

struct Options {
  float alpha = 0.1f;
  float beta = 0.1f;
  uint32_t K = 100;
  uint32_t buffer_size = 33554432;
  uint32_t max_D = 16384;
  uint32_t n_mc = 1;
  uint64_t seed = 0;
  std::string z_file = "temp/z.bin";
  std::string w_file = "temp/w.bin";
  std::string d_file = "temp/d.bin";
  std::string k_d_file = "temp/k_d.bin";
  std::string c_file = "temp/c.txt";
};

[[noreturn]] void die(const std::string& msg) {
  fprintf(stderr, "gpulda_train: %s\n", msg.c_str());
  exit(1);
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
    else if (a == "--K") o.K = (uint32_t)std::stoul(val());
    else if (a == "--buffer_size") o.buffer_size = (uint32_t)std::stoul(val());
    else if (a == "--max_D") o.max_D = (uint32_t)std::stoul(val());
    else if (a == "--n_mc") o.n_mc = (uint32_t)std::stoul(val());
    else if (a == "--seed") o.seed = std::stoull(val());
    else if (a == "--z_temp_file") o.z_file = val();
    else if (a == "--w_temp_file") o.w_file = val();
    else if (a == "--d_temp_file") o.d_file = val();
    else if (a == "--K_d_temp_file") o.k_d_file = val();
    else if (a == "--c_temp_file") o.c_file = val();
    else die("unknown flag " + a);
  }
  return o;
}

// get_token_counts (train.rs:14-17): third tab-separated column of c.txt.
std::vector<uint32_t> token_counts(const std::string& path) {
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  std::vector<uint32_t> c;
  std::string line;
  while (std::getline(f, line)) {
    std::stringstream ss(line);
    std::string tok, id, cnt;
    if (!std::getline(ss, tok, '\t') || !std::getline(ss, id, '\t') || !std::getline(ss, cnt, '\t'))
      die("malformed line in " + path + ": " + line);
    c.push_back((uint32_t)std::stoul(cnt));
  }
  return c;
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

}  // namespace

int main(int argc, char** argv) {
  Options o = parse(argc, argv);
  printf("training\n");
  std::vector<uint32_t> C = token_counts(o.c_file);
  Args args;
  args.alpha = o.alpha;
  args.beta = o.beta;
  args.K = o.K;
  args.V = (uint32_t)C.size();
  args.C = C.data();
  args.buffer_size = o.buffer_size;
  args.max_D = o.max_D;
  Buffer buffers[1];
  memset(buffers, 0, sizeof(buffers));
  gpulda_set_seed(o.seed);
  initialize(&args, buffers, 1);

  Map z = map_file(o.z_file, true);
  Map w = map_file(o.w_file, false);
  Map d = map_file(o.d_file, false);
  Map k_d = map_file(o.k_d_file, false);
  if (z.n != w.n) die("z.bin and w.bin hold different token counts");
  if (d.n != k_d.n) die("d.bin and k_d.bin hold different document counts");

  for (uint32_t it = 0; it < o.n_mc; ++it) {
    sample_phi();
    size_t doc = 0, tok = 0;
    for (;;) {
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
      Buffer& b = buffers[0];
      b.z = z.p + tok;
      b.w = w.p + tok;
      b.d = d.p + doc;
      b.K_d = k_d.p + doc;
      b.n_docs = n_docs;
      b.n_tokens = n_tokens;
      sample_z_async(&b);
      sync_buffer(&b);
      // empty_buffer (train.rs:47-58)
      doc += n_docs;
      tok += n_tokens;
    }
    if (z.bytes) msync(z.p, z.bytes, MS_SYNC);  // z.get_mut().flush() (train.rs:99)
    printf("finished iteration: %u\n", it);
  }
  cleanup(buffers, 1);
  for (Map* m : {&z, &w, &d, &k_d})
    if (m->p) munmap(m->p, m->bytes);
  return 0;
}
