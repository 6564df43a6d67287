/*
 * phc_wire.h -- byte frames for an out-of-process serving front end
 * (libphc_wire.so, host code only).
 *
 * No counterpart in the reference: its solve is an in-process Fortran call.
 * A front end that receives solve requests from other processes needs a
 * byte form of what a caller hands to `add_equation` and `solve`
 * (src/sparse.f90:83-141, src/solve.f90:57-84, 154-163) and of the
 * endpoint records the tracker returns. The request frame carries the
 * system exactly as given -- zero coefficients, negative exponents and
 * surplus equations included -- so that the receiving process builds it
 * with the library's own add_equation and gets the same validation and
 * zero-term dropping as an in-process caller.
 *
 * Layout (little-endian, every section starts at a multiple of 8 bytes):
 *
 *   header, 32 bytes
 *     u32 magic        "PHCR" request, "PHCE" endpoints (bytes in order)
 *     u16 version      1
 *     u16 header bytes 32
 *     u64 frame bytes  header + body
 *     u64 request id   chosen by the client, echoed in the reply
 *     u64 checksum     64-bit FNV-1a of the body
 *   request body
 *     i32 n, i32 neq, i64 total terms
 *     options: 18 f64 then 8 i32, in the order of phc_solve_options with the
 *       reserved fields left out (tracker step_init, step_max, step_min,
 *       corrector_tol, contraction; endgame t_start, radius_factor,
 *       radius_min, loop_radius, winding_agreement, tol, residual_tol,
 *       closure_factor; infinity_tol, singular_cond, real_tol, dedup_tol,
 *       singular_dedup_tol; corrector_maxit, grow_after, max_steps,
 *       nsamples, max_winding, use_endgame, seed, refine_maxit)
 *     i32 nterms[neq]                (padded to 8)
 *     f64 coefs[total][2]
 *     i32 exps[total][n]             (padded to 8)
 *   endpoints body
 *     i64 npaths, i32 N, i32 flags, i32 error, i32 message bytes
 *     error == 0: status i32[np], t f64[np], z f64[np][N][2],
 *       winding i32[np], eg_status i32[np], eg_error f64[np],
 *       naccepted i32[np], nrejected i32[np], then eg_residual f64[np]
 *       if flags & 1, nlin i32[np] if flags & 2, nbe i32[np] if flags & 4
 *       (each i32 array padded to 8)
 *     error != 0: npaths = N = 0 and the message text (padded to 8)
 *
 * Decoding checks the magic, version, sizes (with overflow checks), that
 * the body length matches the counts exactly, the checksum, and that the
 * equation term counts are >= 0 and add up; it never reads past `len`.
 * Decoded arrays are views into the caller's buffer (8-byte aligned), so
 * the buffer must outlive them.
 */
#ifndef PHC_WIRE_H
#define PHC_WIRE_H

#include <stdint.h>
#include "polyhomocont.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PHC_WIRE_VERSION 1
#define PHC_WIRE_HEADER_BYTES 32
/* Largest frame either side accepts (1 GiB). */
#define PHC_WIRE_MAX_FRAME ((int64_t)1 << 30)
#define PHC_WIRE_REQUEST 1
#define PHC_WIRE_ENDPOINTS 2
/* endpoints flags: optional arrays present */
#define PHC_WIRE_EG_RESIDUAL 1
#define PHC_WIRE_NLIN 2
#define PHC_WIRE_NBE 4

/* Frame bytes of a request for `sys`, or -1 if `sys` is malformed
 * (n < 1, neq < 0, a negative term count, missing arrays, or a frame above
 * PHC_WIRE_MAX_FRAME). */
int64_t phc_wire_request_bytes(const phc_user_system *sys);

/* Writes the request frame into buf[0..cap). *written = frame bytes.
 * PHC_ERR_INVALID_INPUT if sys is malformed, PHC_ERR_CAPACITY if cap is too
 * small. */
int phc_wire_encode_request(uint64_t request_id, const phc_user_system *sys,
                            const phc_solve_options *opts, void *buf,
                            int64_t cap, int64_t *written);

/* Decodes a request frame of exactly `len` bytes. *sys receives views into
 * buf (which must be 8-byte aligned), *opts the options (reserved fields
 * 0). PHC_ERR_INVALID_INPUT on any structural error
 * (phc_wire_last_error says which). */
int phc_wire_decode_request(const void *buf, int64_t len,
                            uint64_t *request_id, phc_user_system *sys,
                            phc_solve_options *opts);

/* Frame bytes of an endpoints reply: npaths records of N unknowns with the
 * optional arrays named by flags; -1 if the sizes are out of range. */
int64_t phc_wire_endpoints_bytes(int64_t npaths, int32_t N, int32_t flags);

/* Writes an endpoints reply. flags says which optional arrays of `e` are
 * sent (each must then be non-NULL); the other arrays are required. */
int phc_wire_encode_endpoints(uint64_t request_id, int64_t npaths, int32_t N,
                              int32_t flags, const phc_endpoints *e,
                              void *buf, int64_t cap, int64_t *written);

/* Writes an error reply (npaths 0) carrying a PHC_ERR_* code and a message
 * (at most 4096 bytes are sent). */
int64_t phc_wire_error_bytes(const char *message);
int phc_wire_encode_error(uint64_t request_id, int32_t code,
                          const char *message, void *buf, int64_t cap,
                          int64_t *written);

/* Decodes an endpoints or error reply. On a records reply *error = 0 and
 * *out holds views into buf (absent optional arrays NULL). On an error
 * reply *error is the code, *message points at its text (not
 * NUL-terminated) and *message_bytes is its length. */
int phc_wire_decode_endpoints(const void *buf, int64_t len,
                              uint64_t *request_id, int64_t *npaths,
                              int32_t *N, int32_t *flags, int32_t *error,
                              const char **message, int32_t *message_bytes,
                              phc_endpoints *out);

/* From the first 16 bytes of a frame: its kind (PHC_WIRE_REQUEST or
 * PHC_WIRE_ENDPOINTS) and total length, for readers of a byte stream.
 * PHC_ERR_INVALID_INPUT on an unknown magic or version or an out-of-range
 * length. */
int phc_wire_peek(const void *head, int64_t len, int32_t *kind,
                  int64_t *frame_bytes);

/* Message of the last failed call on this thread (never NULL). */
const char *phc_wire_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PHC_WIRE_H */
