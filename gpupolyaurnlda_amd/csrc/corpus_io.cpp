// corpus_io.cpp -- on-disk corpus formats of the reference host (preprocess.rs, output.rs) and the
// driver's exports and checkpoints.  Host code only.
#include "corpus_io.h"

#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <fstream>
#include <numeric>
#include <sstream>
#include <thread>
#include <unistd.h>
#include <unordered_map>

#include "lda_core.h"

namespace gpulda_io {

void die(const std::string& msg) {
  fprintf(stderr, "gpulda_train: %s\n", msg.c_str());
  exit(1);
}

namespace {

// Buffered binary writer of u32 streams (little-endian: the host byte order of x86-64).
class U32Writer {
 public:
  explicit U32Writer(const std::string& path) : path_(path) {
    f_ = fopen(path.c_str(), "wb");
    if (!f_) die("cannot create " + path + ": " + strerror(errno));
    setvbuf(f_, nullptr, _IOFBF, 1 << 22);
  }
  ~U32Writer() { close(); }
  void put(const uint32_t* p, size_t n) {
    if (n && fwrite(p, sizeof(uint32_t), n, f_) != n) die("write failed: " + path_);
  }
  void put(uint32_t v) { put(&v, 1); }
  void close() {
    if (f_ && fclose(f_) != 0) die("write failed: " + path_);
    f_ = nullptr;
  }

 private:
  std::string path_;
  FILE* f_ = nullptr;
};

// Whole file as u32 values (a trailing partial word is ignored, as output.rs's as_u32 does).
std::vector<uint32_t> read_u32_file(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) die("cannot open " + path + ": " + strerror(errno));
  std::vector<uint32_t> v;
  uint32_t buf[1 << 14];
  size_t got;
  while ((got = fread(buf, sizeof(uint32_t), sizeof(buf) / sizeof(buf[0]), f)) > 0) v.insert(v.end(), buf, buf + got);
  fclose(f);
  return v;
}

// Unicode White_Space code points above ASCII (Rust's char::is_whitespace, which split_whitespace
// uses in preprocess.rs:13).
bool is_unicode_space(uint32_t cp) {
  return cp == 0x85 || cp == 0xA0 || cp == 0x1680 || (cp >= 0x2000 && cp <= 0x200A) || cp == 0x2028 ||
         cp == 0x2029 || cp == 0x202F || cp == 0x205F || cp == 0x3000;
}

// Length of a whitespace character at s[i] (0 if none): ASCII \t \n \v \f \r and space, or one
// of the multi-byte White_Space code points in UTF-8.
size_t space_len(const std::string& s, size_t i) {
  const unsigned char c = (unsigned char)s[i];
  if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
  if (c < 0xC2) return 0;
  uint32_t cp;
  size_t n;
  if (c < 0xE0) {
    cp = c & 0x1F;
    n = 2;
  } else if (c < 0xF0) {
    cp = c & 0x0F;
    n = 3;
  } else {
    return 0;  // no White_Space code point needs four bytes
  }
  if (i + n > s.size()) return 0;
  for (size_t k = 1; k < n; ++k) {
    const unsigned char cc = (unsigned char)s[i + k];
    if ((cc & 0xC0) != 0x80) return 0;
    cp = (cp << 6) | (cc & 0x3F);
  }
  return is_unicode_space(cp) ? n : 0;
}

// Tokens of one input line: the third tab-separated field split on whitespace (get_tokens,
// preprocess.rs:12-14).  A line with fewer than three fields ends the run, as the reference's
// unwrap does.
void line_tokens(const std::string& line, uint64_t line_no, std::vector<std::string>* out) {
  out->clear();
  size_t a = line.find('\t');
  size_t b = a == std::string::npos ? a : line.find('\t', a + 1);
  if (b == std::string::npos) die("input line " + std::to_string(line_no) + " has fewer than three tab-separated fields");
  size_t e = line.find('\t', b + 1);
  if (e == std::string::npos) e = line.size();
  size_t i = b + 1;
  while (i < e) {
    size_t sl;
    while (i < e && (sl = space_len(line, i)) > 0) i += sl;
    if (i >= e) break;
    size_t j = i;
    while (j < e && space_len(line, j) == 0) ++j;
    out->emplace_back(line, i, j - i);
    i = j;
  }
}

// Lines as BufRead::lines yields them: split at '\n', one trailing '\r' removed.
bool next_line(std::istream& in, std::string* line) {
  if (!std::getline(in, *line)) return false;
  if (!line->empty() && line->back() == '\r') line->pop_back();
  return true;
}

std::ifstream open_text(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) die("cannot open " + path + ": " + strerror(errno));
  return in;
}

FILE* create_text(const std::string& path) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) die("cannot create " + path + ": " + strerror(errno));
  setvbuf(f, nullptr, _IOFBF, 1 << 22);
  return f;
}

void close_text(FILE* f, const std::string& path) {
  if (fclose(f) != 0) die("write failed: " + path);
}

// Token names by id from c.txt (get_token_names, output.rs:9-19).
std::vector<std::string> token_names(const std::string& c_path) {
  std::ifstream in = open_text(c_path);
  std::vector<std::string> names;
  std::vector<char> seen;
  std::string line;
  uint64_t no = 0;
  while (next_line(in, &line)) {
    ++no;
    size_t a = line.find('\t');
    if (a == std::string::npos) die("malformed line " + std::to_string(no) + " in " + c_path);
    size_t b = line.find('\t', a + 1);
    std::string id_s = line.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
    char* end = nullptr;
    errno = 0;
    unsigned long id = strtoul(id_s.c_str(), &end, 10);
    if (errno || end == id_s.c_str() || *end || id > 0xFFFFFFFFul)
      die("bad token id on line " + std::to_string(no) + " of " + c_path);
    if (id >= names.size()) {
      names.resize(id + 1);
      seen.resize(id + 1, 0);
    }
    names[id] = line.substr(0, a);
    seen[id] = 1;
  }
  for (size_t i = 0; i < seen.size(); ++i)
    if (!seen[i]) names[i].clear();
  return names;
}

}  // namespace

CorpusShape preprocess(const std::string& input, const TempPaths& t, uint32_t K, uint64_t seed) {
  if (K == 0) die("K must be positive");
  // pass 1: vocabulary with counts, in order of first appearance (count_tokens, preprocess.rs:16-32)
  std::unordered_map<std::string, uint32_t> first;
  std::vector<std::string> names;
  std::vector<uint64_t> counts;
  std::vector<std::string> toks;
  std::string line;
  uint64_t docs = 0;
  {
    std::ifstream in = open_text(input);
    while (next_line(in, &line)) {
      line_tokens(line, docs + 1, &toks);
      ++docs;
      for (std::string& s : toks) {
        auto it = first.find(s);
        if (it == first.end()) {
          first.emplace(s, (uint32_t)names.size());
          names.push_back(s);
          counts.push_back(1);
        } else {
          ++counts[it->second];
        }
      }
    }
  }
  if (names.size() > 0xFFFFFFFFull) die("vocabulary larger than 2^32 entries");
  // ids by decreasing count (preprocess.rs:23); ties in order of first appearance, where the
  // reference's order is that of a hash map
  std::vector<uint32_t> order(names.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return counts[a] > counts[b]; });
  std::vector<uint32_t> id(names.size());
  FILE* c = create_text(t.c);
  for (uint32_t r = 0; r < order.size(); ++r) {
    id[order[r]] = r;
    if (counts[order[r]] > 0xFFFFFFFFull) die("token count above 2^32: " + names[order[r]]);
    fprintf(c, "%s\t%u\t%" PRIu64 "\n", names[order[r]].c_str(), r, counts[order[r]]);
  }
  close_text(c, t.c);

  // pass 2: the token streams (preprocess.rs:54-70)
  U32Writer zf(t.z), wf(t.w), df(t.d), kf(t.k_d);
  std::vector<uint32_t> zs, ws;
  std::vector<uint32_t> stamp(K, 0xFFFFFFFFu);
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  uint64_t tokens = 0;
  std::ifstream in = open_text(input);
  for (uint64_t doc = 0; doc < docs; ++doc) {
    if (!next_line(in, &line)) die("input changed while it was read: " + input);
    line_tokens(line, doc + 1, &toks);
    if (toks.size() > 0xFFFFFFFFull) die("document with more than 2^32 tokens");
    zs.resize(toks.size());
    ws.resize(toks.size());
    uint32_t distinct = 0;
    for (uint32_t i = 0; i < toks.size(); ++i) {
      ws[i] = id[first.at(toks[i])];
      lda::u32x4 r = lda::philox_tagged(i, (uint32_t)doc, 0, lda::kTagInitZ, 0, k0, k1);
      zs[i] = lda::bounded(r.x, K);
      if (stamp[zs[i]] != (uint32_t)doc) {
        stamp[zs[i]] = (uint32_t)doc;
        ++distinct;
      }
    }
    zf.put(zs.data(), zs.size());
    wf.put(ws.data(), ws.size());
    df.put((uint32_t)toks.size());
    kf.put(distinct);
    tokens += toks.size();
  }
  zf.close();
  wf.close();
  df.close();
  kf.close();
  CorpusShape sh;
  sh.docs = docs;
  sh.tokens = tokens;
  sh.vocab = (uint32_t)names.size();
  return sh;
}

std::vector<uint32_t> token_counts(const std::string& c_path) {
  std::ifstream in = open_text(c_path);
  std::vector<uint32_t> c;
  std::string line;
  while (next_line(in, &line)) {
    std::stringstream ss(line);
    std::string tok, idv, cnt;
    if (!std::getline(ss, tok, '\t') || !std::getline(ss, idv, '\t') || !std::getline(ss, cnt, '\t'))
      die("malformed line in " + c_path + ": " + line);
    c.push_back((uint32_t)std::stoul(cnt));
  }
  return c;
}

uint64_t write_output(const std::string& output, const TempPaths& t) {
  const std::vector<std::string> names = token_names(t.c);
  const std::vector<uint32_t> d = read_u32_file(t.d);
  const std::vector<uint32_t> w = read_u32_file(t.w);
  const std::vector<uint32_t> z = read_u32_file(t.z);
  if (w.size() != z.size()) die("z.bin and w.bin hold different token counts");
  FILE* f = create_text(output);
  uint64_t pos = 0;
  for (size_t doc = 0; doc < d.size(); ++doc) {
    if (pos + d[doc] > w.size()) die("d.bin lengths exceed the tokens in w.bin");
    for (uint32_t i = 0; i < d[doc]; ++i, ++pos) {
      const uint32_t v = w[pos];
      if (v >= names.size() || names[v].empty()) die("word id " + std::to_string(v) + " is not in " + t.c);
      fputs(names[v].c_str(), f);
      fprintf(f, ":%u\t", z[pos]);
    }
    fputc('\n', f);
  }
  close_text(f, output);
  return d.size();
}

uint64_t write_theta(const std::string& path, const TempPaths& t, uint32_t K) {
  const std::vector<uint32_t> d = read_u32_file(t.d);
  const std::vector<uint32_t> z = read_u32_file(t.z);
  std::vector<uint32_t> n(K, 0);
  std::vector<uint32_t> keys;
  FILE* f = create_text(path);
  uint64_t pos = 0;
  for (size_t doc = 0; doc < d.size(); ++doc) {
    if (pos + d[doc] > z.size()) die("d.bin lengths exceed the tokens in z.bin");
    keys.clear();
    for (uint32_t i = 0; i < d[doc]; ++i, ++pos) {
      const uint32_t k = z[pos];
      if (k >= K) die("topic " + std::to_string(k) + " >= K in " + t.z);
      if (n[k]++ == 0) keys.push_back(k);
    }
    std::sort(keys.begin(), keys.end());
    for (size_t j = 0; j < keys.size(); ++j) {
      fprintf(f, j ? " %u:%u" : "%u:%u", keys[j], n[keys[j]]);
      n[keys[j]] = 0;
    }
    fputc('\n', f);
  }
  close_text(f, path);
  return d.size();
}

void write_phi(const std::string& path, const float* phi, uint32_t V, uint32_t K) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) die("cannot create " + path + ": " + strerror(errno));
  const size_t n = (size_t)V * K;
  if (n && fwrite(phi, sizeof(float), n, f) != n) die("write failed: " + path);
  if (fclose(f) != 0) die("write failed: " + path);
}

std::vector<uint32_t> counts_from_files(const TempPaths& t, uint32_t V, uint32_t K) {
  const std::vector<uint32_t> w = read_u32_file(t.w);
  const std::vector<uint32_t> z = read_u32_file(t.z);
  if (w.size() != z.size()) die("z.bin and w.bin hold different token counts");
  std::vector<uint32_t> n((size_t)V * K, 0);
  for (size_t i = 0; i < w.size(); ++i) {
    if (w[i] >= V || z[i] >= K) die("token " + std::to_string(i) + " has a word id >= V or a topic >= K");
    ++n[(size_t)w[i] * K + z[i]];
  }
  return n;
}

namespace {
// 64-bit hash of one chunk (position c in the sequence, len bytes).
uint64_t chunk_hash(const unsigned char* p, uint64_t len, uint64_t c) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ (c * 0xD6E8FEB86659FD93ull) ^ len;
  uint64_t i = 0;
  for (; i + 8 <= len; i += 8) {
    uint64_t x;
    memcpy(&x, p + i, 8);
    h = (h ^ x) * 0x100000001B3ull;
    h ^= h >> 29;
  }
  for (; i < len; ++i) h = (h ^ p[i]) * 0x100000001B3ull;
  return h;
}

constexpr uint64_t kHashChunk = 64ull << 20;

uint64_t combine_chunks(const std::vector<uint64_t>& part, uint64_t bytes) {
  uint64_t h = 0xCBF29CE484222325ull ^ bytes;
  for (uint64_t p : part) {
    h = (h ^ p) * 0x100000001B3ull;
    h ^= h >> 31;
  }
  return h;
}

unsigned hash_threads() { return std::max(1u, std::min(8u, std::thread::hardware_concurrency())); }
}  // namespace

uint64_t bytes_hash(const void* data, uint64_t bytes) {
  // fixed 64 MiB chunks hashed in parallel, then combined in chunk order: the value does not
  // depend on the number of threads
  const unsigned char* p = static_cast<const unsigned char*>(data);
  const uint64_t n_chunks = (bytes + kHashChunk - 1) / kHashChunk;
  std::vector<uint64_t> part(n_chunks, 0);
  const unsigned n_threads = hash_threads();
  auto work = [&](unsigned tid) {
    for (uint64_t c = tid; c < n_chunks; c += n_threads)
      part[c] = chunk_hash(p + c * kHashChunk, std::min(kHashChunk, bytes - c * kHashChunk), c);
  };
  std::vector<std::thread> th;
  for (unsigned i = 0; i < n_threads; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
  return combine_chunks(part, bytes);
}

uint64_t file_hash(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) die("cannot open " + path + ": " + strerror(errno));
  fseek(f, 0, SEEK_END);
  const uint64_t bytes = (uint64_t)ftell(f);
  fclose(f);
  const uint64_t n_chunks = (bytes + kHashChunk - 1) / kHashChunk;
  std::vector<uint64_t> part(n_chunks, 0);
  const unsigned n_threads = hash_threads();
  auto work = [&](unsigned tid) {
    FILE* g = fopen(path.c_str(), "rb");
    if (!g) return;
    std::vector<unsigned char> buf(kHashChunk);
    for (uint64_t c = tid; c < n_chunks; c += n_threads) {
      const uint64_t len = std::min(kHashChunk, bytes - c * kHashChunk);
      fseek(g, (long)(c * kHashChunk), SEEK_SET);
      part[c] = fread(buf.data(), 1, len, g) == len ? chunk_hash(buf.data(), len, c) : ~0ull;
    }
    fclose(g);
  };
  std::vector<std::thread> th;
  for (unsigned i = 0; i < n_threads; ++i) th.emplace_back(work, i);
  for (auto& x : th) x.join();
  return combine_chunks(part, bytes);
}

namespace {
// Flushes a written file to stable storage and closes it.
void close_durable(FILE* f, const std::string& path) {
  if (fflush(f) != 0 || fsync(fileno(f)) != 0) die("write failed: " + path + ": " + strerror(errno));
  close_text(f, path);
}

std::string dir_of(const std::string& path) {
  const size_t sl = path.rfind('/');
  return sl == std::string::npos ? std::string(".") : path.substr(0, sl);
}

std::string base_of(const std::string& path) {
  const size_t sl = path.rfind('/');
  return sl == std::string::npos ? path : path.substr(sl + 1);
}

void sync_dir(const std::string& dir) {
  int fd = open(dir.c_str(), O_RDONLY);
  if (fd >= 0) {
    (void)fsync(fd);
    close(fd);
  }
}
}  // namespace

std::string snapshot_name(const std::string& record_path, uint32_t iterations) {
  return base_of(record_path) + ".z" + std::to_string(iterations);
}

std::string snapshot_file(const std::string& record_path, const std::string& name) {
  return dir_of(record_path) + "/" + name;
}

void write_snapshot(const std::string& path, const void* z, uint64_t bytes) {
  const std::string tmp = path + ".tmp";
  FILE* f = create_text(tmp);
  if (bytes && fwrite(z, 1, bytes, f) != bytes) die("write failed: " + tmp);
  close_durable(f, tmp);
  if (rename(tmp.c_str(), path.c_str()) != 0) die("cannot rename " + tmp + ": " + strerror(errno));
  sync_dir(dir_of(path));
}

std::vector<unsigned char> read_snapshot(const Checkpoint& c, const std::string& record_path) {
  const std::string path = snapshot_file(record_path, c.z_snapshot);
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) die("checkpoint snapshot " + path + " is missing: " + strerror(errno));
  std::vector<unsigned char> buf(c.tokens * sizeof(uint32_t));
  const size_t got = buf.empty() ? 0 : fread(buf.data(), 1, buf.size(), f);
  const bool extra = fgetc(f) != EOF;
  fclose(f);
  if (got != buf.size() || extra) die("checkpoint snapshot " + path + " does not hold " + std::to_string(c.tokens) + " topics");
  if (bytes_hash(buf.data(), buf.size()) != c.z_hash) die("checkpoint snapshot " + path + " does not match its record");
  return buf;
}

void write_checkpoint(const std::string& path, const Checkpoint& c) {
  const std::string tmp = path + ".tmp";
  FILE* f = create_text(tmp);
  fprintf(f, "gpulda-checkpoint 1\n");
  if (!c.z_snapshot.empty()) fprintf(f, "z_snapshot %s\n", c.z_snapshot.c_str());
  fprintf(f, "iterations %u\nseed %" PRIu64 "\nK %u\nV %u\ntokens %" PRIu64 "\ndocuments %" PRIu64 "\n", c.iterations,
          c.seed, c.K, c.V, c.tokens, c.docs);
  fprintf(f, "alpha_bits %08x\nbeta_bits %08x\nbuffer_size %u\nmax_D %u\nz_hash %016" PRIx64 "\n", c.alpha_bits,
          c.beta_bits, c.buffer_size, c.max_D, c.z_hash);
  close_durable(f, tmp);
  if (rename(tmp.c_str(), path.c_str()) != 0) die("cannot rename " + tmp + ": " + strerror(errno));
  sync_dir(dir_of(path));
}

bool read_checkpoint(const std::string& path, Checkpoint* c) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  if (!next_line(in, &line) || line != "gpulda-checkpoint 1") die(path + " is not a checkpoint of this driver");
  std::unordered_map<std::string, std::string> kv;
  while (next_line(in, &line)) {
    size_t sp = line.find(' ');
    if (sp == std::string::npos) die("malformed checkpoint line: " + line);
    kv[line.substr(0, sp)] = line.substr(sp + 1);
  }
  auto get = [&](const char* k, int base) -> uint64_t {
    auto it = kv.find(k);
    if (it == kv.end()) die(std::string("checkpoint lacks ") + k);
    char* end = nullptr;
    errno = 0;
    unsigned long long v = strtoull(it->second.c_str(), &end, base);
    if (errno || end == it->second.c_str() || *end) die(std::string("bad checkpoint value for ") + k);
    return v;
  };
  c->iterations = (uint32_t)get("iterations", 10);
  c->seed = get("seed", 10);
  c->K = (uint32_t)get("K", 10);
  c->V = (uint32_t)get("V", 10);
  c->tokens = get("tokens", 10);
  c->docs = get("documents", 10);
  c->alpha_bits = (uint32_t)get("alpha_bits", 16);
  c->beta_bits = (uint32_t)get("beta_bits", 16);
  c->buffer_size = (uint32_t)get("buffer_size", 10);
  c->max_D = (uint32_t)get("max_D", 10);
  c->z_hash = get("z_hash", 16);
  auto snap = kv.find("z_snapshot");
  c->z_snapshot = snap == kv.end() ? std::string() : snap->second;
  if (c->z_snapshot.find('/') != std::string::npos) die("bad checkpoint snapshot name: " + c->z_snapshot);
  return true;
}

}  // namespace gpulda_io
