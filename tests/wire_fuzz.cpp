// Mutation fuzzing of the frame decoders (include/phc_wire.h), built by
// tests/test_wire_fuzz.py with AddressSanitizer. Every decode runs on a
// heap buffer of exactly the frame's length, so a read past the end is an
// ASan error; a decode that succeeds has every view it returned read in
// full. Mutations: random bytes, truncation, extension, and header or
// count fields overwritten, with the checksum recomputed half of the time
// so that the structural checks behind it are reached.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "phc_wire.h"

namespace {

uint64_t rng_state = 88172645463325252ull;
uint64_t next_u64() {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}
int below(int n) { return (int)(next_u64() % (uint64_t)n); }

uint64_t fnv1a(const unsigned char* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

void resign(std::vector<unsigned char>& f) {
  if (f.size() < 32) return;
  const uint64_t h = fnv1a(f.data() + 32, f.size() - 32);
  memcpy(f.data() + 24, &h, 8);
}

volatile double sink = 0.0;

void touch(const void* p, size_t bytes) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  double s = 0.0;
  for (size_t i = 0; i < bytes; ++i) s += b[i];
  sink = sink + s;
}

std::vector<unsigned char> request_frame(int n, int neq, int terms) {
  std::vector<int32_t> nterms(neq, terms);
  const int total = neq * terms;
  std::vector<double> coefs(2 * (size_t)total + 2);
  std::vector<int32_t> exps((size_t)total * n + 1);
  for (auto& c : coefs) c = (double)below(1000) / 7.0 - 50.0;
  for (auto& e : exps) e = below(4) - (below(50) == 0 ? 1 : 0);
  phc_user_system sys{n, neq, nterms.data(), coefs.data(), exps.data()};
  phc_solve_options o;
  memset(&o, 0, sizeof(o));
  o.seed = 3;
  const int64_t bytes = phc_wire_request_bytes(&sys);
  std::vector<unsigned char> f((size_t)bytes);
  int64_t w = 0;
  if (phc_wire_encode_request(7, &sys, &o, f.data(), bytes, &w) != PHC_OK ||
      w != bytes) {
    fprintf(stderr, "encode failed: %s\n", phc_wire_last_error());
    exit(2);
  }
  return f;
}

std::vector<unsigned char> endpoints_frame(int np, int N, int flags) {
  std::vector<int32_t> i32(7 * (size_t)np + 7);
  std::vector<double> f64((3 + 2 * (size_t)N) * np + 3);
  phc_endpoints e{};
  e.status = i32.data();
  e.winding = i32.data() + np;
  e.eg_status = i32.data() + 2 * np;
  e.naccepted = i32.data() + 3 * np;
  e.nrejected = i32.data() + 4 * np;
  e.nlin = i32.data() + 5 * np;
  e.nbe = i32.data() + 6 * np;
  e.t = f64.data();
  e.eg_error = f64.data() + np;
  e.eg_residual = f64.data() + 2 * np;
  e.z = f64.data() + 3 * np;
  const int64_t bytes = phc_wire_endpoints_bytes(np, N, flags);
  std::vector<unsigned char> f((size_t)bytes);
  int64_t w = 0;
  if (phc_wire_encode_endpoints(9, np, N, flags, &e, f.data(), bytes, &w) !=
      PHC_OK) {
    fprintf(stderr, "encode failed: %s\n", phc_wire_last_error());
    exit(2);
  }
  return f;
}

void mutate(std::vector<unsigned char>& f) {
  switch (below(6)) {
    case 0:  // random bytes anywhere
      for (int k = 1 + below(4); k > 0 && !f.empty(); --k)
        f[below((int)f.size())] = (unsigned char)below(256);
      break;
    case 1:  // truncate
      f.resize(f.empty() ? 0 : (size_t)below((int)f.size()));
      break;
    case 2:  // extend
      f.resize(f.size() + 1 + below(64), (unsigned char)below(256));
      break;
    case 3:  // a 32-bit field in the body's first 64 bytes
      if (f.size() >= 40) {
        const int at = 32 + 4 * below((int)std::min<size_t>(16, (f.size() - 32) / 4));
        const uint32_t vals[] = {0u, 1u, 0xffffffffu, 0x7fffffffu, 0x80000000u,
                                 (uint32_t)below(1 << 20)};
        const uint32_t v = vals[below(6)];
        memcpy(f.data() + at, &v, 4);
      }
      break;
    case 4:  // the length field
      if (f.size() >= 16) {
        const uint64_t v = below(2) ? f.size() + below(17) - 8
                                    : (uint64_t)next_u64() >> below(64);
        memcpy(f.data() + 8, &v, 8);
      }
      break;
    default:  // magic or version
      if (f.size() >= 8) f[below(8)] ^= (unsigned char)(1 + below(255));
  }
  if (below(2)) resign(f);
}

// Decodes `f` from a heap buffer of exactly f.size() bytes.
void decode_one(const std::vector<unsigned char>& f, long& ok, long& bad) {
  const size_t n = f.size();
  unsigned char* buf = static_cast<unsigned char*>(malloc(n ? n : 1));
  if (n) memcpy(buf, f.data(), n);
  int32_t kind = 0;
  int64_t fb = 0;
  if (n >= 16) (void)phc_wire_peek(buf, (int64_t)n, &kind, &fb);
  uint64_t id = 0;
  phc_user_system sys;
  phc_solve_options o;
  if (phc_wire_decode_request(buf, (int64_t)n, &id, &sys, &o) == PHC_OK) {
    ++ok;
    int64_t total = 0;
    for (int i = 0; i < sys.neq; ++i) total += sys.nterms[i];
    touch(sys.nterms, 4 * (size_t)sys.neq);
    touch(sys.coefs, 16 * (size_t)total);
    touch(sys.exps, 4 * (size_t)total * sys.n);
  } else {
    ++bad;
  }
  int64_t np = 0;
  int32_t N = 0, flags = 0, err = 0, mlen = 0;
  const char* msg = nullptr;
  phc_endpoints e;
  if (phc_wire_decode_endpoints(buf, (int64_t)n, &id, &np, &N, &flags, &err,
                                &msg, &mlen, &e) == PHC_OK) {
    ++ok;
    if (err) {
      touch(msg, (size_t)mlen);
    } else {
      touch(e.status, 4 * (size_t)np);
      touch(e.t, 8 * (size_t)np);
      touch(e.z, 16 * (size_t)np * N);
      touch(e.winding, 4 * (size_t)np);
      touch(e.eg_status, 4 * (size_t)np);
      touch(e.eg_error, 8 * (size_t)np);
      touch(e.naccepted, 4 * (size_t)np);
      touch(e.nrejected, 4 * (size_t)np);
      if (e.eg_residual) touch(e.eg_residual, 8 * (size_t)np);
      if (e.nlin) touch(e.nlin, 4 * (size_t)np);
      if (e.nbe) touch(e.nbe, 4 * (size_t)np);
    }
  } else {
    ++bad;
  }
  free(buf);
}

}  // namespace

int main(int argc, char** argv) {
  const long rounds = argc > 1 ? atol(argv[1]) : 20000;
  std::vector<std::vector<unsigned char>> seeds = {
      request_frame(1, 1, 1), request_frame(3, 3, 5), request_frame(8, 9, 7),
      request_frame(5, 0, 0), endpoints_frame(0, 0, 0),
      endpoints_frame(5, 4, 7), endpoints_frame(17, 9, 2)};
  {
    const int64_t bytes = phc_wire_error_bytes("rejected");
    std::vector<unsigned char> f((size_t)bytes);
    int64_t w = 0;
    if (phc_wire_encode_error(4, PHC_ERR_INVALID_INPUT, "rejected", f.data(),
                              bytes, &w) != PHC_OK)
      return 2;
    seeds.push_back(f);
  }
  long ok = 0, bad = 0;
  for (const auto& s : seeds) decode_one(s, ok, bad);
  if (ok != (long)seeds.size()) {
    fprintf(stderr, "an unmutated frame did not decode\n");
    return 3;
  }
  for (long r = 0; r < rounds; ++r) {
    std::vector<unsigned char> f = seeds[below((int)seeds.size())];
    for (int k = 1 + below(3); k > 0; --k) mutate(f);
    decode_one(f, ok, bad);
  }
  printf("decoded %ld, rejected %ld\n", ok, bad);
  return 0;
}
