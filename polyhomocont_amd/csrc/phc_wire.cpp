// Request and endpoint frames for an out-of-process serving front end.
// Host code only (built with g++ into libphc_wire.so); the layout is
// described in include/phc_wire.h.

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/phc_wire.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__,
              "the frames are little-endian and written with memcpy");

namespace {

thread_local char g_err[256] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

constexpr uint32_t MAGIC_REQUEST = 0x52434850u;    // "PHCR"
constexpr uint32_t MAGIC_ENDPOINTS = 0x45434850u;  // "PHCE"
constexpr int64_t HDR = PHC_WIRE_HEADER_BYTES;
constexpr int64_t REQ_FIXED = 16;                  // n, neq, total
constexpr int64_t OPT_BYTES = 18 * 8 + 8 * 4;
constexpr int64_t EP_FIXED = 24;  // npaths, N, flags, error, message bytes
constexpr int32_t MAX_MESSAGE = 4096;
constexpr int32_t MAX_N = 1 << 16;
constexpr int64_t MAX_COUNT = (int64_t)1 << 31;

int64_t pad8(int64_t b) { return (b + 7) & ~int64_t(7); }

// a * b, or -1 when either is negative or the product passes `limit`
int64_t mul_checked(int64_t a, int64_t b, int64_t limit) {
  if (a < 0 || b < 0) return -1;
  if (a != 0 && b > limit / a) return -1;
  return a * b;
}

uint64_t fnv1a(const unsigned char* p, int64_t n) {
  uint64_t h = 1469598103934665603ull;
  for (int64_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct Writer {
  unsigned char* p;
  int64_t at = 0;
  template <typename T>
  void put(T v) {
    memcpy(p + at, &v, sizeof(T));
    at += sizeof(T);
  }
  void bytes(const void* src, int64_t n) {
    if (n > 0) memcpy(p + at, src, (size_t)n);
    at += n;
  }
  void align() {
    while (at % 8) p[at++] = 0;
  }
};

struct Reader {
  const unsigned char* p;
  int64_t at = 0;
  template <typename T>
  T get() {
    T v;
    memcpy(&v, p + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  template <typename T>
  const T* view(int64_t count) {
    const T* v = reinterpret_cast<const T*>(p + at);
    at = pad8(at + count * (int64_t)sizeof(T));
    return v;
  }
};

void write_header(unsigned char* buf, uint32_t magic, int64_t frame,
                  uint64_t request_id) {
  Writer w{buf};
  w.put<uint32_t>(magic);
  w.put<uint16_t>(PHC_WIRE_VERSION);
  w.put<uint16_t>((uint16_t)HDR);
  w.put<uint64_t>((uint64_t)frame);
  w.put<uint64_t>(request_id);
  w.put<uint64_t>(fnv1a(buf + HDR, frame - HDR));
}

// Checks the header of a frame of exactly `len` bytes; *id the request id.
int check_header(const void* buf, int64_t len, uint32_t magic,
                 uint64_t* id) {
  if (!buf || len < HDR)
    return fail(PHC_ERR_INVALID_INPUT, "frame shorter than its header");
  if (reinterpret_cast<uintptr_t>(buf) % 8)
    return fail(PHC_ERR_INVALID_INPUT, "frame buffer not 8-byte aligned");
  Reader r{static_cast<const unsigned char*>(buf)};
  const uint32_t m = r.get<uint32_t>();
  const uint16_t version = r.get<uint16_t>();
  const uint16_t hbytes = r.get<uint16_t>();
  const uint64_t frame = r.get<uint64_t>();
  const uint64_t rid = r.get<uint64_t>();
  const uint64_t sum = r.get<uint64_t>();
  if (m != magic) return fail(PHC_ERR_INVALID_INPUT, "wrong frame magic");
  if (version != PHC_WIRE_VERSION)
    return fail(PHC_ERR_INVALID_INPUT, "unsupported frame version %u",
                (unsigned)version);
  if (hbytes != HDR)
    return fail(PHC_ERR_INVALID_INPUT, "unexpected header size %u",
                (unsigned)hbytes);
  if (frame != (uint64_t)len)
    return fail(PHC_ERR_INVALID_INPUT,
                "frame length %llu differs from the %lld bytes given",
                (unsigned long long)frame, (long long)len);
  if (fnv1a(static_cast<const unsigned char*>(buf) + HDR, len - HDR) != sum)
    return fail(PHC_ERR_INVALID_INPUT, "frame checksum mismatch");
  *id = rid;
  return PHC_OK;
}

// Body bytes of a request with these counts, or -1.
int64_t request_body(int64_t n, int64_t neq, int64_t total) {
  if (n < 1 || n > MAX_N || neq < 0 || neq >= MAX_COUNT || total < 0 ||
      total >= MAX_COUNT)
    return -1;
  const int64_t cbytes = mul_checked(total, 16, PHC_WIRE_MAX_FRAME);
  const int64_t ecount = mul_checked(total, n, PHC_WIRE_MAX_FRAME);
  if (cbytes < 0 || ecount < 0) return -1;
  const int64_t body =
      REQ_FIXED + OPT_BYTES + pad8(4 * neq) + cbytes + pad8(4 * ecount);
  return HDR + body > PHC_WIRE_MAX_FRAME ? -1 : body;
}

int64_t endpoints_body(int64_t np, int64_t N, int32_t flags) {
  if (np < 0 || np >= MAX_COUNT || N < 0 || N > MAX_N || (flags & ~7))
    return -1;
  const int64_t zb = mul_checked(np, 16 * N, PHC_WIRE_MAX_FRAME);
  if (zb < 0 || np > PHC_WIRE_MAX_FRAME / 8) return -1;
  const int64_t i32 = pad8(4 * np), f64 = 8 * np;
  int64_t body = EP_FIXED + 5 * i32 + 2 * f64 + zb;
  if (flags & PHC_WIRE_EG_RESIDUAL) body += f64;
  if (flags & PHC_WIRE_NLIN) body += i32;
  if (flags & PHC_WIRE_NBE) body += i32;
  return HDR + body > PHC_WIRE_MAX_FRAME ? -1 : body;
}

int64_t total_terms(const phc_user_system* sys) {
  if (!sys || sys->n < 1 || sys->neq < 0) return -1;
  if (sys->neq > 0 && !sys->nterms) return -1;
  int64_t total = 0;
  for (int32_t i = 0; i < sys->neq; ++i) {
    if (sys->nterms[i] < 0) return -1;
    total += sys->nterms[i];
  }
  if (total > 0 && (!sys->coefs || !sys->exps)) return -1;
  return total;
}

void put_options(Writer& w, const phc_solve_options& o) {
  const double d[18] = {
      o.tracker.step_init,         o.tracker.step_max,
      o.tracker.step_min,          o.tracker.corrector_tol,
      o.tracker.contraction,       o.endgame.t_start,
      o.endgame.radius_factor,     o.endgame.radius_min,
      o.endgame.loop_radius,       o.endgame.winding_agreement,
      o.endgame.tol,               o.endgame.residual_tol,
      o.endgame.closure_factor,    o.infinity_tol,
      o.singular_cond,             o.real_tol,
      o.dedup_tol,                 o.singular_dedup_tol};
  const int32_t i[8] = {o.tracker.corrector_maxit, o.tracker.grow_after,
                        o.tracker.max_steps,       o.endgame.nsamples,
                        o.endgame.max_winding,     o.use_endgame,
                        o.seed,                    o.refine_maxit};
  w.bytes(d, sizeof(d));
  w.bytes(i, sizeof(i));
}

void get_options(Reader& r, phc_solve_options* o) {
  memset(o, 0, sizeof(*o));
  double* d[18] = {
      &o->tracker.step_init,       &o->tracker.step_max,
      &o->tracker.step_min,        &o->tracker.corrector_tol,
      &o->tracker.contraction,     &o->endgame.t_start,
      &o->endgame.radius_factor,   &o->endgame.radius_min,
      &o->endgame.loop_radius,     &o->endgame.winding_agreement,
      &o->endgame.tol,             &o->endgame.residual_tol,
      &o->endgame.closure_factor,  &o->infinity_tol,
      &o->singular_cond,           &o->real_tol,
      &o->dedup_tol,               &o->singular_dedup_tol};
  int32_t* i[8] = {&o->tracker.corrector_maxit, &o->tracker.grow_after,
                   &o->tracker.max_steps,       &o->endgame.nsamples,
                   &o->endgame.max_winding,     &o->use_endgame,
                   &o->seed,                    &o->refine_maxit};
  for (double* p : d) *p = r.get<double>();
  for (int32_t* p : i) *p = r.get<int32_t>();
}

}  // namespace

extern "C" {

const char* phc_wire_last_error(void) { return g_err; }

int64_t phc_wire_request_bytes(const phc_user_system* sys) {
  const int64_t total = total_terms(sys);
  if (total < 0) return -1;
  const int64_t body = request_body(sys->n, sys->neq, total);
  return body < 0 ? -1 : HDR + body;
}

int phc_wire_encode_request(uint64_t request_id, const phc_user_system* sys,
                            const phc_solve_options* opts, void* buf,
                            int64_t cap, int64_t* written) {
  if (!opts || !written) return fail(PHC_ERR_INVALID_INPUT, "null argument");
  const int64_t frame = phc_wire_request_bytes(sys);
  if (frame < 0) return fail(PHC_ERR_INVALID_INPUT, "malformed system");
  if (!buf || cap < frame)
    return fail(PHC_ERR_CAPACITY, "buffer of %lld bytes, frame needs %lld",
                (long long)cap, (long long)frame);
  const int64_t total = total_terms(sys);
  unsigned char* out = static_cast<unsigned char*>(buf);
  Writer w{out, HDR};
  w.put<int32_t>(sys->n);
  w.put<int32_t>(sys->neq);
  w.put<int64_t>(total);
  put_options(w, *opts);
  w.bytes(sys->nterms, 4 * (int64_t)sys->neq);
  w.align();
  w.bytes(sys->coefs, 16 * total);
  w.bytes(sys->exps, 4 * total * sys->n);
  w.align();
  if (w.at != frame)
    return fail(PHC_ERR_NUMERICAL, "internal size mismatch");
  write_header(out, MAGIC_REQUEST, frame, request_id);
  *written = frame;
  return PHC_OK;
}

int phc_wire_decode_request(const void* buf, int64_t len,
                            uint64_t* request_id, phc_user_system* sys,
                            phc_solve_options* opts) {
  if (!request_id || !sys || !opts)
    return fail(PHC_ERR_INVALID_INPUT, "null argument");
  uint64_t rid = 0;
  int rc = check_header(buf, len, MAGIC_REQUEST, &rid);
  if (rc) return rc;
  if (len < HDR + REQ_FIXED + OPT_BYTES)
    return fail(PHC_ERR_INVALID_INPUT, "request body too short");
  Reader r{static_cast<const unsigned char*>(buf), HDR};
  const int32_t n = r.get<int32_t>();
  const int32_t neq = r.get<int32_t>();
  const int64_t total = r.get<int64_t>();
  const int64_t body = request_body(n, neq, total);
  if (body < 0)
    return fail(PHC_ERR_INVALID_INPUT, "request counts out of range");
  if (HDR + body != len)
    return fail(PHC_ERR_INVALID_INPUT,
                "request body of %lld bytes, its counts need %lld",
                (long long)(len - HDR), (long long)body);
  get_options(r, opts);
  const int32_t* nterms = r.view<int32_t>(neq);
  int64_t sum = 0;
  for (int32_t i = 0; i < neq; ++i) {
    if (nterms[i] < 0)
      return fail(PHC_ERR_INVALID_INPUT, "negative term count");
    sum += nterms[i];
  }
  if (sum != total)
    return fail(PHC_ERR_INVALID_INPUT, "term counts add up to %lld, not %lld",
                (long long)sum, (long long)total);
  const double* coefs = r.view<double>(2 * total);
  const int32_t* exps = r.view<int32_t>(total * n);
  sys->n = n;
  sys->neq = neq;
  sys->nterms = neq ? nterms : nullptr;
  sys->coefs = total ? coefs : nullptr;
  sys->exps = total ? exps : nullptr;
  *request_id = rid;
  return PHC_OK;
}

int64_t phc_wire_endpoints_bytes(int64_t npaths, int32_t N, int32_t flags) {
  const int64_t body = endpoints_body(npaths, N, flags);
  return body < 0 ? -1 : HDR + body;
}

int phc_wire_encode_endpoints(uint64_t request_id, int64_t npaths, int32_t N,
                              int32_t flags, const phc_endpoints* e,
                              void* buf, int64_t cap, int64_t* written) {
  if (!e || !written) return fail(PHC_ERR_INVALID_INPUT, "null argument");
  const int64_t frame = phc_wire_endpoints_bytes(npaths, N, flags);
  if (frame < 0) return fail(PHC_ERR_INVALID_INPUT, "sizes out of range");
  if (npaths > 0 &&
      (!e->status || !e->t || !e->z || !e->winding || !e->eg_status ||
       !e->eg_error || !e->naccepted || !e->nrejected ||
       ((flags & PHC_WIRE_EG_RESIDUAL) && !e->eg_residual) ||
       ((flags & PHC_WIRE_NLIN) && !e->nlin) ||
       ((flags & PHC_WIRE_NBE) && !e->nbe)))
    return fail(PHC_ERR_INVALID_INPUT, "a record array is missing");
  if (!buf || cap < frame)
    return fail(PHC_ERR_CAPACITY, "buffer of %lld bytes, frame needs %lld",
                (long long)cap, (long long)frame);
  unsigned char* out = static_cast<unsigned char*>(buf);
  Writer w{out, HDR};
  w.put<int64_t>(npaths);
  w.put<int32_t>(N);
  w.put<int32_t>(flags);
  w.put<int32_t>(0);
  w.put<int32_t>(0);
  const int64_t i32 = 4 * npaths, f64 = 8 * npaths;
  w.bytes(e->status, i32); w.align();
  w.bytes(e->t, f64);
  w.bytes(e->z, f64 * 2 * N);
  w.bytes(e->winding, i32); w.align();
  w.bytes(e->eg_status, i32); w.align();
  w.bytes(e->eg_error, f64);
  w.bytes(e->naccepted, i32); w.align();
  w.bytes(e->nrejected, i32); w.align();
  if (flags & PHC_WIRE_EG_RESIDUAL) w.bytes(e->eg_residual, f64);
  if (flags & PHC_WIRE_NLIN) { w.bytes(e->nlin, i32); w.align(); }
  if (flags & PHC_WIRE_NBE) { w.bytes(e->nbe, i32); w.align(); }
  if (w.at != frame)
    return fail(PHC_ERR_NUMERICAL, "internal size mismatch");
  write_header(out, MAGIC_ENDPOINTS, frame, request_id);
  *written = frame;
  return PHC_OK;
}

int64_t phc_wire_error_bytes(const char* message) {
  const int64_t m = message ? (int64_t)strnlen(message, MAX_MESSAGE) : 0;
  return HDR + EP_FIXED + pad8(m);
}

int phc_wire_encode_error(uint64_t request_id, int32_t code,
                          const char* message, void* buf, int64_t cap,
                          int64_t* written) {
  if (!written) return fail(PHC_ERR_INVALID_INPUT, "null argument");
  if (code == PHC_OK)
    return fail(PHC_ERR_INVALID_INPUT, "an error reply needs a nonzero code");
  const int64_t frame = phc_wire_error_bytes(message);
  if (!buf || cap < frame)
    return fail(PHC_ERR_CAPACITY, "buffer of %lld bytes, frame needs %lld",
                (long long)cap, (long long)frame);
  const int32_t m = message ? (int32_t)strnlen(message, MAX_MESSAGE) : 0;
  unsigned char* out = static_cast<unsigned char*>(buf);
  Writer w{out, HDR};
  w.put<int64_t>(0);
  w.put<int32_t>(0);
  w.put<int32_t>(0);
  w.put<int32_t>(code);
  w.put<int32_t>(m);
  w.bytes(message, m);
  w.align();
  write_header(out, MAGIC_ENDPOINTS, frame, request_id);
  *written = frame;
  return PHC_OK;
}

int phc_wire_decode_endpoints(const void* buf, int64_t len,
                              uint64_t* request_id, int64_t* npaths,
                              int32_t* N, int32_t* flags, int32_t* error,
                              const char** message, int32_t* message_bytes,
                              phc_endpoints* out) {
  if (!request_id || !npaths || !N || !flags || !error || !message ||
      !message_bytes || !out)
    return fail(PHC_ERR_INVALID_INPUT, "null argument");
  uint64_t rid = 0;
  int rc = check_header(buf, len, MAGIC_ENDPOINTS, &rid);
  if (rc) return rc;
  if (len < HDR + EP_FIXED)
    return fail(PHC_ERR_INVALID_INPUT, "endpoints body too short");
  Reader r{static_cast<const unsigned char*>(buf), HDR};
  const int64_t np = r.get<int64_t>();
  const int32_t n = r.get<int32_t>();
  const int32_t f = r.get<int32_t>();
  const int32_t err = r.get<int32_t>();
  const int32_t m = r.get<int32_t>();
  memset(out, 0, sizeof(*out));
  *message = nullptr;
  *message_bytes = 0;
  if (err != 0) {
    if (np != 0 || n != 0 || f != 0 || m < 0 || m > MAX_MESSAGE ||
        HDR + EP_FIXED + pad8(m) != len)
      return fail(PHC_ERR_INVALID_INPUT, "malformed error reply");
    *message = reinterpret_cast<const char*>(r.view<char>(m));
    *message_bytes = m;
  } else {
    const int64_t body = endpoints_body(np, n, f);
    if (body < 0 || m != 0)
      return fail(PHC_ERR_INVALID_INPUT, "endpoint counts out of range");
    if (HDR + body != len)
      return fail(PHC_ERR_INVALID_INPUT,
                  "endpoints body of %lld bytes, its counts need %lld",
                  (long long)(len - HDR), (long long)body);
    out->status = const_cast<int32_t*>(r.view<int32_t>(np));
    out->t = const_cast<double*>(r.view<double>(np));
    out->z = const_cast<double*>(r.view<double>(np * 2 * n));
    out->winding = const_cast<int32_t*>(r.view<int32_t>(np));
    out->eg_status = const_cast<int32_t*>(r.view<int32_t>(np));
    out->eg_error = const_cast<double*>(r.view<double>(np));
    out->naccepted = const_cast<int32_t*>(r.view<int32_t>(np));
    out->nrejected = const_cast<int32_t*>(r.view<int32_t>(np));
    if (f & PHC_WIRE_EG_RESIDUAL)
      out->eg_residual = const_cast<double*>(r.view<double>(np));
    if (f & PHC_WIRE_NLIN) out->nlin = const_cast<int32_t*>(r.view<int32_t>(np));
    if (f & PHC_WIRE_NBE) out->nbe = const_cast<int32_t*>(r.view<int32_t>(np));
  }
  *request_id = rid;
  *npaths = np;
  *N = n;
  *flags = f;
  *error = err;
  return PHC_OK;
}

int phc_wire_peek(const void* head, int64_t len, int32_t* kind,
                  int64_t* frame_bytes) {
  if (!head || len < 16 || !kind || !frame_bytes)
    return fail(PHC_ERR_INVALID_INPUT, "need the first 16 bytes of a frame");
  Reader r{static_cast<const unsigned char*>(head)};
  const uint32_t m = r.get<uint32_t>();
  const uint16_t version = r.get<uint16_t>();
  const uint16_t hbytes = r.get<uint16_t>();
  const uint64_t frame = r.get<uint64_t>();
  if (m != MAGIC_REQUEST && m != MAGIC_ENDPOINTS)
    return fail(PHC_ERR_INVALID_INPUT, "wrong frame magic");
  if (version != PHC_WIRE_VERSION || hbytes != HDR)
    return fail(PHC_ERR_INVALID_INPUT, "unsupported frame version");
  if (frame < (uint64_t)HDR || frame > (uint64_t)PHC_WIRE_MAX_FRAME)
    return fail(PHC_ERR_INVALID_INPUT, "frame length out of range");
  *kind = m == MAGIC_REQUEST ? PHC_WIRE_REQUEST : PHC_WIRE_ENDPOINTS;
  *frame_bytes = (int64_t)frame;
  return PHC_OK;
}

}  // extern "C"
