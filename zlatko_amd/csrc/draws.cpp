// draws.cpp -- the replayed normal stream on the host's cores (see draws.hpp).
#include "draws.hpp"

#include <algorithm>
#include <cstring>
#include <thread>

namespace ccrhip {

namespace {

constexpr int64_t kChunkOutputs = int64_t(1) << 17;  // engine outputs per chunk (even: whole attempts)
constexpr int kChunksPerThread = 4;                    // chunks per worker per batch
// Normals per engine output: pi / 4 of the attempts are accepted, two normals each per two outputs.
constexpr double kYield = 0.785;

// Replays a chunk of recorded engine outputs to std::normal_distribution; the chunk's end is
// signalled by an exception, which only a new attempt's first draw can raise (chunks hold whole
// attempts, and the second normal of an accepted attempt is returned without a draw).
struct Replay {
  using result_type = std::mt19937_64::result_type;
  struct End {};
  const uint64_t* p;
  const uint64_t* end;
  static constexpr result_type min() { return std::mt19937_64::min(); }
  static constexpr result_type max() { return std::mt19937_64::max(); }
  result_type operator()() {
    if (p == end) throw End{};
    return *p++;
  }
};

void normalsOfChunk(const uint64_t* raw, int64_t n, std::vector<double>& out) {
  out.clear();
  out.reserve(static_cast<size_t>(n));  // at most one normal per output
  std::normal_distribution<double> normal(0.0, 1.0);
  Replay r{raw, raw + n};
  try {
    for (;;) out.push_back(normal(r));
  } catch (const Replay::End&) {
  }
}

}  // namespace

NormalStream::NormalStream(uint64_t seed, int threads, int64_t expected) : engine_(seed), expected_(expected) {
  const unsigned hw = std::thread::hardware_concurrency();
  threads_ = threads > 0 ? threads : std::max(1, std::min(16, static_cast<int>(hw == 0 ? 1 : hw)));
}

// Chunks of a batch that covers `normals` more draws (one batch's worth when unknown).
int NormalStream::chunksFor(int64_t normals) const {
  const int full = threads_ * kChunksPerThread;
  if (expected_ <= 0) return full;
  const double need = static_cast<double>(normals) / (kYield * static_cast<double>(kChunkOutputs));
  return std::max(1, std::min(full, static_cast<int>(need * 1.02) + 1));
}

void NormalStream::generate(std::vector<uint64_t>& raw, int chunks) {
  raw.resize(static_cast<size_t>(chunks * kChunkOutputs));
  for (auto& v : raw) v = engine_();
}

void NormalStream::refill() {
  if (nextChunks_ == 0) {  // first batch: nothing generated ahead
    nextChunks_ = chunksFor(expected_ - produced_);
    generate(next_, nextChunks_);
  }
  raw_.swap(next_);
  const int chunks = nextChunks_;
  // Expected draws left after this batch: run the engine for the next batch meanwhile.
  const int64_t after = expected_ - produced_ - static_cast<int64_t>(kYield * chunks * kChunkOutputs);
  nextChunks_ = expected_ > 0 && after <= 0 ? 0 : chunksFor(after);
  std::thread ahead;
  if (nextChunks_ > 0) ahead = std::thread([this] { generate(next_, nextChunks_); });
  parts_.resize(static_cast<size_t>(chunks));
  const int workers = std::max(1, std::min(chunks, threads_ - (nextChunks_ > 0 ? 1 : 0)));
  auto work = [&](int t) {
    for (int c = t; c < chunks; c += workers)
      normalsOfChunk(raw_.data() + c * kChunkOutputs, kChunkOutputs, parts_[static_cast<size_t>(c)]);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < workers; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  if (ahead.joinable()) ahead.join();
  part_ = 0;
  pos_ = 0;
}

void NormalStream::fill(double* out, int64_t m) {
  while (m > 0) {
    while (part_ < parts_.size() && pos_ == parts_[part_].size()) {
      ++part_;
      pos_ = 0;
    }
    if (part_ >= parts_.size()) refill();
    const std::vector<double>& p = parts_[part_];
    const int64_t k = std::min<int64_t>(m, static_cast<int64_t>(p.size() - pos_));
    std::memcpy(out, p.data() + pos_, static_cast<size_t>(k) * sizeof(double));
    pos_ += static_cast<size_t>(k);
    produced_ += k;
    out += k;
    m -= k;
  }
}

void parallelNormals(uint64_t seed, int64_t n, double* out, int threads) {
  NormalStream s(seed, threads, n);
  s.fill(out, n);
}

}  // namespace ccrhip
