// draws.hpp -- the replayed normal stream on the host's cores.
//
// ccr::NormalGenerator (include/ccr/core/random.hpp:13-26) draws std::normal_distribution<double>
// from one std::mt19937_64: the distribution's polar method consumes the engine in pairs (one
// attempt = two uniforms) and yields both normals of an accepted attempt, the second on the next
// call. Attempts are independent given their two engine outputs, so the stream is rebuilt in
// parallel: one thread runs the engine sequentially into batches of chunks of an even number of
// outputs (the next batch while the current one is transformed), worker threads run a fresh
// std::normal_distribution over each chunk (a chunk never splits an attempt, and every chunk ends
// with no normal held back), and the chunks' normals are concatenated in chunk order --
// bit-identical to the sequential stream, from the same library code.
#pragma once

#include <cstdint>
#include <random>
#include <vector>

namespace ccrhip {

class NormalStream {
 public:
  // threads <= 0: min(16, hardware threads). expected: the number of normals the caller will
  // draw (sizes the batches; 0 if unknown).
  explicit NormalStream(uint64_t seed, int threads = 0, int64_t expected = 0);
  // The next m normals of the stream.
  void fill(double* out, int64_t m);
  int threads() const { return threads_; }

 private:
  int chunksFor(int64_t normals) const;
  void generate(std::vector<uint64_t>& raw, int chunks);
  void refill();
  std::mt19937_64 engine_;
  int threads_;
  int64_t expected_, produced_ = 0;
  std::vector<uint64_t> raw_, next_;
  int nextChunks_ = 0;
  std::vector<std::vector<double>> parts_;
  size_t part_ = 0, pos_ = 0;
};

// The first n normals of the stream for `seed` (parallel, bit-identical to the sequential one).
void parallelNormals(uint64_t seed, int64_t n, double* out, int threads = 0);

}  // namespace ccrhip
