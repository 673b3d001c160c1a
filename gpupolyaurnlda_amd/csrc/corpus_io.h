// corpus_io.h -- the reference host's on-disk corpus formats, for the C++ driver.
//
//   input      one document per line, `name \t label \t text`, tokens = the text's
//              whitespace-separated words (preprocess.rs:12-14)
//   c.txt      `token \t id \t count` per vocabulary entry, ids in order of decreasing count
//              (preprocess.rs:16-32)
//   z.bin, w.bin, d.bin, k_d.bin
//              little-endian u32 streams: topic and word id per token, length and number of
//              distinct initial topics per document (preprocess.rs:42-72)
//   output     one line per document, `token:topic \t` per token (output.rs:27-62)
//
// Added here: Phi and theta exports, and a checkpoint record that lets a run resume.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gpulda_io {

struct TempPaths {
  std::string z, w, d, k_d, c;
};

struct CorpusShape {
  uint64_t docs = 0;
  uint64_t tokens = 0;
  uint32_t vocab = 0;
};

// Builds the vocabulary and the temporary files from a Mallet-style input (preprocess.rs:42-72).
// Initial topics are Philox draws with counters (token in document, document, 0, INIT_Z) and the
// seed, uniform over K -- the same draws the library makes for a corpus loaded without z.
CorpusShape preprocess(const std::string& input, const TempPaths& t, uint32_t K, uint64_t seed);

// Third column of c.txt, in file order (train.rs:14-17).
std::vector<uint32_t> token_counts(const std::string& c_path);

// Writes `token:topic\t` per token and a newline per document (output.rs:27-62).  Returns the
// number of documents written.
uint64_t write_output(const std::string& output, const TempPaths& t);

// Theta export: one line per document, `k:n_dk` for every topic present, ascending k, separated by
// spaces (an empty document gives an empty line).
uint64_t write_theta(const std::string& path, const TempPaths& t, uint32_t K);

// Phi export: V rows of K little-endian f32, row v for word id v of c.txt.
void write_phi(const std::string& path, const float* phi, uint32_t V, uint32_t K);

// n_wk of a z.bin / w.bin pair, V x K word-major.
std::vector<uint32_t> counts_from_files(const TempPaths& t, uint32_t V, uint32_t K);

// Hash of a file's bytes (order-dependent, 64-bit); bytes_hash of the same bytes in memory is equal.
uint64_t file_hash(const std::string& path);
uint64_t bytes_hash(const void* data, uint64_t bytes);

// State after a completed iteration: the run's parameters, the corpus shape, and a snapshot of z
// (a file beside the record, named in it) with its hash.  The next iteration rewrites z.bin in
// place, so a run stopped at any point resumes from the snapshot, not from z.bin.
struct Checkpoint {
  uint32_t iterations = 0;
  uint64_t seed = 0;
  uint32_t K = 0, V = 0;
  uint64_t tokens = 0, docs = 0;
  uint32_t alpha_bits = 0, beta_bits = 0;
  uint32_t buffer_size = 0, max_D = 0;
  uint64_t z_hash = 0;
  std::string z_snapshot;  // file name of the z snapshot, in the record's directory
};

// Written to path + ".tmp", flushed to disk and renamed, so a reader sees either the old or the new
// record.  The snapshot it names is written first (write_snapshot), so every record names a
// complete snapshot.
void write_checkpoint(const std::string& path, const Checkpoint& c);
// false if the file does not exist; exits with a message if it is malformed.
bool read_checkpoint(const std::string& path, Checkpoint* c);
// Name of the snapshot of the record at `record_path` after `iterations` iterations, and the file
// a snapshot name refers to (in the record's directory).
std::string snapshot_name(const std::string& record_path, uint32_t iterations);
std::string snapshot_file(const std::string& record_path, const std::string& name);
// z as little-endian u32, written to path + ".tmp", flushed to disk and renamed.
void write_snapshot(const std::string& path, const void* z, uint64_t bytes);
// The record's snapshot, checked against its length and hash (exits with a message otherwise).
std::vector<unsigned char> read_snapshot(const Checkpoint& c, const std::string& record_path);

[[noreturn]] void die(const std::string& msg);

}  // namespace gpulda_io
