/*
 * onion-gen candidate-key search on MI355X (gfx950): the C ABI.
 *
 * This is the drop-in boundary for the reference's device engine.  Each entry
 * point names the reference interface it replaces (paths relative to the
 * reference repository root).  All functions are extern "C", take plain
 * pointers and sizes, throw nothing, and report errors as a negative status
 * with a message copied into the caller's (err, errcap) buffer.
 *
 * Conventions
 *   - A field element / coordinate is 32 bytes, little-endian, canonical.
 *   - A point "xyzt" is 128 bytes: X, Y, Z, T of extended coordinates
 *     (src/curve.rs:43-49), the start point s0*B of a block.
 *   - Candidate k is the point (s0 + 8k)*B; a survivor's offset is 8k, the
 *     scalar delta from s0 (src/gpu/mod.rs:113-119, src/batch.rs:74-77).
 */
#ifndef ONION_GEN_HIP_H
#define ONION_GEN_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- devices */

/* Replaces gpu::cuda::look (src/gpu/cuda.rs:35-82) for the HIP/ROCm path:
 * the number of visible devices, 0 when there is no driver or no device. */
int og_hip_device_count(void);

/* Replaces the Device record fields (src/gpu/mod.rs:49-70): name, memory
 * (bytes) and compute units of device `ordinal`. */
int og_hip_device_info(int ordinal, char *name, size_t namecap, uint64_t *memory,
                       int *compute_units);

/* ----------------------------------------------------------------- engine */

/* Replaces pub struct Survivor (src/gpu/mod.rs:113-119). */
typedef struct {
  uint64_t offset;    /* scalar delta from s0: 8k */
  uint8_t packed[32]; /* canonical y, little-endian, sign bit clear */
} og_survivor;

/* Replaces tuning::Shape (src/tuning.rs:27-39), extended with `rounds`,
 * `radix` and `slots`; first = pairs per window (half), second = chains
 * (threads; per unit slot for the slot engine's entry). */
typedef struct {
  uint32_t first;
  uint32_t second;
  double rate;
  int64_t at;
  uint32_t rounds;
  uint32_t radix;
  uint32_t slots;  /* the slot engine's unit slots per step; 0 elsewhere */
} og_tuning_shape;

typedef struct og_engine og_engine;
typedef struct og_filterset og_filterset;

/* Replaces gpu::hit_room (src/gpu/mod.rs:306-312): survivor room per launch. */
uint32_t og_hit_room(uint64_t batch, uint32_t bitmap_bits);

/* Replaces gpu::Engine::new (src/gpu/mod.rs:346-353) with the arguments the
 * reference binds: the start point, the chain count (0 = the shape the tuning
 * cache recorded, measured first when there is none), and the prefix bitmap of
 * FilterSet::index_words (src/filter.rs:1857-1859) as 64-bit words. */
int og_hip_engine_create(int ordinal, const uint8_t start_xyzt[128], uint32_t threads,
                         const uint64_t *bitmap, size_t bitmap_words, uint32_t bitmap_bits,
                         og_engine **out, char *err, size_t errcap);

/* Shape of an engine.  threads = 0 takes the whole shape (threads, half,
 * rounds) from the tuning cache, or measures it first (og_hip_engine_tune);
 * otherwise zero half/rounds take their defaults. */
typedef struct {
  uint32_t threads; /* chains (multiple of 256); 0 = the tuned shape */
  uint32_t half;    /* pairs per window: one window is 2*half+1 candidates */
  uint32_t rounds;  /* windows per chain per advance */
  uint32_t hit_cap; /* survivor room per advance; 0 = og_hit_room */
  uint32_t sign_stage; /* general forms: where the sign of x is formed for the
                          candidates the key alone cannot settle.  0: chosen
                          from the filters (the address kernel when under 1 key
                          in 8 needs it), 1: in the search loop, 2: in the
                          address kernel.  Launches of more than one round, or
                          of half below 8, always form it in the loop. */
  uint64_t max_bytes;  /* threads = 0: the device bytes the chosen shape may
                          hold (a remembered shape above it is measured
                          again); 0 = half of the device's free memory */
} og_engine_opts;

/* The extended constructor: the filter set compiled by og_filters_parse,
 * which adds to the reference's bitmap-only device path a literal-prefix
 * compare, an exact table behind the bitmap (src/filter.rs:2645-2701) and the
 * checksum forms (suffix, contains) evaluated on the device with SHA3-256. */
int og_hip_engine_create_ex(int ordinal, const uint8_t start_xyzt[128], const og_engine_opts *opts,
                            const og_filterset *filters, og_engine **out, char *err,
                            size_t errcap);

/* Replaces the automatic choice of Engine::new with retune = true
 * (src/gpu/mod.rs:354-412), bounded by memory as its walk is
 * (src/gpu/mod.rs:314-342): times one launch of every shape of a walk over
 * chains {2^20 .. 2^16} x half {1024 .. 128} (rounds for about 2^28
 * candidates a launch) that fits `max_bytes` of device memory (0 = half of
 * the free memory), with the given filters; keeps, among the shapes within 3%
 * of the fastest, the one with the smallest footprint (og_tuning_choose);
 * records it in the tuning cache under og_tuning_device_key(<name>, "hip", 0)
 * and writes it. */
int og_hip_engine_tune(int ordinal, const uint8_t start_xyzt[128], const og_filterset *filters,
                       uint64_t max_bytes, og_tuning_shape *best, char *err, size_t errcap);

/* Device bytes an engine holds: chain state, scratch planes, tables, ring.
 * No counterpart in the reference (its comment at src/gpu/mod.rs:314-318
 * states the figure). */
uint64_t og_hip_engine_bytes(const og_engine *e);

/* The shape an engine runs: chains, pairs per window, windows per launch. */
void og_hip_engine_shape(const og_engine *e, uint32_t *threads, uint32_t *half, uint32_t *rounds);

/* Where an engine forms the sign of x (og_engine_opts::sign_stage): 1 in the
 * search loop, 2 in the address kernel; -1 for a null engine.  No counterpart
 * in the reference, whose CPU path forms it per candidate (src/batch.rs). */
int og_hip_engine_sign_stage(const og_engine *e);

/* The search kernels an engine runs: 1 the key forms (prefix compare, bitmap
 * and exact table: no parking planes, the three-column probe of y), 2 the
 * pattern forms (the automata, the checksum stage); -1 for a null engine.  No
 * counterpart in the reference. */
int og_hip_engine_kernels(const og_engine *e);

/* The device ordinal an engine runs on. */
int og_hip_engine_device(const og_engine *e);

/* Survivors one advance may return: the reference's hit_room for its batch
 * and bitmap (src/gpu/mod.rs:306-312), raised to four times the expected
 * survivors plus 1024 when the filter set lets more through (a dictionary),
 * at most 2^22.  Size the advance's output by it. */
uint32_t og_hip_engine_room(const og_engine *e);

/* Starts the engine over from another start point (its candidate 0), keeping
 * its shape, filters and buffers: the next advance examines k = 0..batch-1 of
 * the new start.  No counterpart in the reference, which builds an engine per
 * block (src/run.rs:1048-1056); here a lease's blocks reuse one engine. */
int og_hip_engine_restart(og_engine *e, const uint8_t start_xyzt[128], char *err, size_t errcap);

/* Replaces Engine::threads / Engine::batch (src/gpu/mod.rs:446-467). */
uint32_t og_hip_engine_threads(const og_engine *e);
uint64_t og_hip_engine_batch(const og_engine *e);

/* Replaces Engine::advance (src/gpu/mod.rs:470-476, cuda.rs:282-365): examines
 * the next batch() candidates and writes the survivors.  Blocking; not
 * re-entrant.  Overflowing `cap` or the device room is an error, never a
 * silent truncation (src/gpu/cuda.rs:356-363). */
int og_hip_engine_advance(og_engine *e, og_survivor *out, uint32_t cap, uint32_t *n, char *err,
                          size_t errcap);

/* Time, in milliseconds, of the device work of the last advance (HIP events
 * around the launches on the engine's stream). */
double og_hip_engine_last_ms(const og_engine *e);

/* Replaces Drop for the engine (src/gpu/cuda.rs:103-122). */
void og_hip_engine_destroy(og_engine *e);

/* Elementwise GF(2^255-19) arithmetic for parity tests of the field itself
 * (the device forms of src/field.rs:82-255): ops 0..5 in the 8 x 32 form and
 * 10..15 in the 10 x 25.5 form the search kernels run in; each block is mul,
 * sq, add, sub, invert, canonical; op 20 is the address checksum of a as a
 * packed key (SHA3-256(".onion checksum" | a | 03)[0..2] in the low 16 bits of
 * the output); op 16 is (2x - y)(x - y) in the 10-limb
 * form, the loosest operands the search multiplies; ops 17 and 18 canonicalise
 * a product's output as the search kernels do (17: x mod 2^255, split into
 * limbs without reduction, times one; 18: x times y) and op 19 is bit 0 of the
 * canonical (x mod 2^255, split the same way) times y, as the search takes the
 * sign of x.  Ops 21 and 22 are the products by a table constant the search
 * loop makes, the constant taken from the element leading each group of 64
 * (element i & ~63, A = its a, B = its b): 21 is x * B, 22 is x * B + y * A
 * (one double product).  Op 23 is the key forms' probe word for the product of
 * op 16's operands, from four of its columns, in output word 0, and in word 1
 * whether it is exact: the canonical product's low 32 bits are word 0 when
 * word 1 is 1, else word 0 plus 0, 19 or 38.  Op 27 is the same from three
 * columns: bits 0..25 of the canonical product (word 0 masked to 26 bits),
 * exact on the same terms.  Op 24 is op 19 taken as the search takes it,
 * from three columns where they settle it.  Op 25 is
 * (x - y) y in the 10-limb form, the loosest operands the search loops
 * multiply with the 32-bit fold of the top carry.  Op 26 is the inverse of x
 * by the divstep recurrence the search windows use (0 maps to 0).  Inputs are
 * n raw 32-byte little-endian words (any value below 2^256); outputs are
 * canonical encodings.  The _host
 * variant runs the same code on the CPU. */
int og_hip_field_op(int ordinal, int op, const uint8_t *a, const uint8_t *b, uint8_t *out,
                    size_t n, char *err, size_t errcap);
int og_field_op_host(int op, const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);

/* The radix bench (no counterpart in the reference; it decides the field
 * representation for gfx950): `threads` lanes each run two chains of `iters`
 * dependent multiplications in radix 5 (5 x 51 in 64-bit limbs, the
 * fiat-crypto form of src/field.rs:368-432, each limb product four 32-bit
 * multiplies), 8 (8 x 32 saturated limbs, row-wise), 9 (9 x 29 limbs,
 * carry-free columns) or 10 (10 x 25.5 limbs, the shape of
 * src/field4.rs:568-670).  Writes field
 * multiplications per second and, when `results` is non-null, the 2*threads
 * canonical chain ends (64 bytes per lane). */
int og_hip_radix_bench(int ordinal, int radix, uint32_t threads, uint32_t iters,
                       double *muls_per_s, uint8_t *results, char *err, size_t errcap);

/* The issue-rate probe (no counterpart in the reference; it measures the
 * roofline's denominator, the peak rate of the VALU instruction the field
 * multiplication is made of): `threads` lanes each run eight independent
 * chains of `iters` x 8 instructions of one kind - op 0 v_mad_u64_u32, 1
 * v_mul_lo_u32, 2 v_lshrrev_b64, 3 v_add_u32, 4 v_bitop3_b32, 5
 * v_alignbit_b32 (the last two make up the Keccak rounds) - or, op 6, one
 * dependent chain of v_mad_u64_u32 per lane at three waves per SIMD.  Writes
 * lane-instructions per second and the shader clock the SIMDs ran at
 * (s_memtime against the constant-rate s_memrealtime, in Hz; 0 when
 * unknown). */
int og_hip_issue_bench(int ordinal, int op, uint32_t threads, uint32_t iters, double *ops_per_s,
                       double *clock_hz, char *err, size_t errcap);

/* ---------------------------------------------------- allocation-free search */

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that must not allocate,
 * harnesses that forbid hipMalloc).  One advance of an engine of this shape
 * (og_hip_engine_create_ex + og_hip_engine_advance, one window per chain) for
 * a set of exactly one literal prefix: examines candidates k0 .. k0 +
 * threads * (2 * half + 1) - 1 of the start point, with the offset table,
 * chain state, scratch planes and survivor ring in a static device workspace
 * of og_hip_search_noalloc_max() bytes (64 MiB) per device, which each device
 * holds until the process ends.  Calls no hipMalloc, hipFree or allocating
 * library routine.  `out` (cap records) and `n_out` are device memory; the
 * work is ordered on `stream` (a hipStream_t; NULL for the null stream) and
 * the call returns without waiting.  When the stream reaches its end, *n_out
 * holds the survivors found and the first min(*n_out, cap, 16384) are in out
 * (offset = 8k, packed = canonical y); a count above min(cap, 16384) is an
 * overflow, an error as for og_hip_engine_advance.  A shape whose
 * og_hip_search_noalloc_bytes exceeds the workspace is refused with a
 * message, as is any other filter set.  Every call on a device shares its
 * workspace: calls that overlap on one device race. */
int og_hip_search_noalloc(int ordinal, const uint8_t start_xyzt[128], uint64_t k0,
                          uint32_t threads, uint32_t half, const og_filterset *filters,
                          og_survivor *out, uint32_t cap, uint32_t *n_out, void *stream,
                          char *err, size_t errcap);
/* Workspace bytes og_hip_search_noalloc needs for a shape, and the size of the
 * static workspace. */
size_t og_hip_search_noalloc_bytes(uint32_t threads, uint32_t half);
size_t og_hip_search_noalloc_max(void);

/* --------------------------------------------------------------- hit sink */

enum {
  OG_HIT_OK = 0,        /* verified: the secret produces this key, which matched */
  OG_HIT_CLAMPING = 1,  /* the offset broke clamping: dropped (src/key.rs:196-206) */
  OG_HIT_MISMATCH = 2   /* the re-derived key disagrees with the survivor */
};

typedef struct {
  uint64_t offset;             /* 8k, as the survivor's */
  int32_t status;              /* OG_HIT_* */
  uint8_t public_key[32];      /* with the sign bit, when status is OG_HIT_OK */
  uint8_t expanded_secret[64]; /* tor's layout: scalar at the offset | right half */
  char hostname[63];           /* "<56 symbols>.onion", NUL-terminated */
} og_key_record;

/* Replaces the per-hit host work of run::device_worker / record_hit
 * (src/run.rs:1030-1116, 740-896) with key::expanded_secret_at_offset
 * (src/key.rs:167-206) and the independent re-derivation of verify::one
 * (src/verify.rs:89-189), for a whole batch of survivors of the search seeded
 * by `seed`: the secret at each offset (clamping checked on the host), its
 * public key recomputed on the device from the scalar alone, 255 bits compared
 * with the survivor, the hostname built.  Writes n records. */
int og_hip_finish_hits(int ordinal, const uint8_t seed[32], const og_survivor *hits, size_t n,
                       og_key_record *out, char *err, size_t errcap);

/* Replaces output::write_key (src/output.rs:42-71): under `dir`, a directory
 * named after the hostname (mode 0700) holding hs_ed25519_secret_key (96 bytes,
 * mode 0600), hs_ed25519_public_key (64 bytes) and hostname.  An existing
 * directory is never touched.  Returns 0 when created, 1 when it existed (the
 * path is written either way), -1 on an error (message in path_out). */
int og_write_key(const char *dir, const uint8_t public_key[32], const uint8_t expanded_secret[64],
                 char *path_out, size_t cap);

/* ------------------------------------------------------------------ lease */

/* proto::Lease (src/proto.rs:47-57) as the adapter takes it. */
typedef struct {
  uint64_t lease;
  uint8_t seed[32];     /* the root seed (the wire carries it as hex) */
  uint64_t first_block;
  uint64_t blocks;
} og_lease;

/* proto::Found (src/proto.rs:78-84): where a find is. */
typedef struct {
  uint64_t lease;
  uint64_t block;
  uint64_t offset;      /* within the block's chain: 8k */
} og_lease_find;

/* A lease worker's search (src/worker/run.rs:165-229 over run::run with
 * first_block / block_limit, src/run.rs:256-261): for each block b of the lease,
 * the chain of block_seed(seed, b) (src/run.rs:211-220) through `launches` advances
 * of the engine (0 = BATCHES_PER_BLOCK, 64, src/run.rs:29), every survivor
 * verified by og_hip_finish_hits, and confirmed against `filters` when given,
 * before it is reported.  The engine is restarted
 * per block.  Writes up to `cap` finds (an overflow is an error), the candidates
 * examined and the blocks done. */
int og_hip_run_lease(og_engine *e, const og_lease *lease, const og_filterset *filters,
                     uint32_t launches, og_lease_find *out, size_t cap, size_t *n,
                     uint64_t *candidates, uint64_t *blocks_done, char *err, size_t errcap);

/* ---------------------------------------------------------------- filters */

/* Replaces FilterSet::parse_all + FilterSet::new / with_index_bits
 * (src/filter.rs:1930-1950, 1326-1476): newline-separated filters in the
 * reference's syntax ("abc", "prefix:abc", "suffix:..d", "contains:abcde",
 * '?' and '[a-d]' classes, "regex:<expression>" as src/expression.rs parses
 * it).  index_bits = 0 chooses the width as narrowest_acceptable does
 * (src/filter.rs:2279-2299).  Expressions are kept after the other filters,
 * in source order, never absorbed (src/filter.rs:869-876). */
og_filterset *og_filters_parse(const char *text, uint32_t index_bits, char *err, size_t errcap);
void og_filters_free(og_filterset *s);
/* Filters kept after dedupe and absorption. */
size_t og_filters_count(const og_filterset *s);
/* Replaces FilterSet::index_words (src/filter.rs:1857-1859): copies the bitmap
 * into words (when cap suffices) and returns its width in bits, 0 if none. */
uint32_t og_filters_index_words(const og_filterset *s, uint64_t *words, size_t cap);
/* Replaces FilterSet::match_key (src/filter.rs:2098-2150) on a packed key
 * without sign, restricted to literal prefixes: index of the matching filter
 * or -1. */
int og_filters_match_key(const og_filterset *s, const uint8_t packed[32]);
/* Replaces FilterSet::match_whole_address (src/filter.rs:1673-1691): index of
 * a filter the 35 address bytes match, or -1.  Expressions are matched over
 * the 56 address symbols as Expression::find does (src/expression.rs:209-215). */
int og_filters_match_address(const og_filterset *s, const uint8_t address[35]);
/* Replaces FilterSet::describe_structures (src/filter.rs:1700-1845): one line
 * per structure the device reads, and per expression its proxies (the
 * obligatory symbol sequences that guard its automaton) and reach.  Copies at
 * most cap-1 bytes and a terminator; returns the full length. */
size_t og_filters_describe(const og_filterset *s, char *out, size_t cap);
/* The device matcher of the general forms (suffix, contains, class prefixes)
 * and the expressions, run on the host for one key: the two stages of
 * FilterSet::examine_key (src/filter.rs:1625-1691) with Filter::key_prefilter
 * (:782-810) and the expressions' literal guards (:914-1005).  packed is the
 * canonical y (sign bit clear), sign the sign of x.  key_hit: the key's
 * symbols 0..48 settled a match; need: the key stage left a match reaching
 * past symbol 48 possible; hit: the verdict on the whole address.  Returns 0,
 * or -1 when the set has neither general filter nor expression, or does not
 * compile. */
int og_filters_scan_host(const og_filterset *s, const uint8_t packed[32], int sign,
                         int *key_hit, int *need, int *hit);
/* The key stage of og_filters_scan_host for the two candidates of a pair at
 * once, as the search kernels run it (one interleaved walk of the automata
 * for both): key_hit[i] and need[i] as above.  No counterpart in the
 * reference; it exists so that the pair form is checked against the single
 * one.  Returns 0, or -1 as og_filters_scan_host (and for a set without
 * general filters or expressions in the automaton banks). */
int og_filters_scan_pair_host(const og_filterset *s, const uint8_t a[32], const uint8_t b[32],
                              int key_hit[2], int need[2]);
/* The address stage's automata on a 35-byte address (packed key, checksum,
 * version) as the address kernels run them: *match = whether one accepts it,
 * *window = whether the set's address-stage window prefilter lets the walk
 * run (1 for a set without one).  No counterpart in the reference; it exists
 * so that the prefilter is checked never to stop a walk that would match.
 * Returns 0, or -1 for a null argument or a set without automata. */
int og_filters_address_window_host(const og_filterset *s, const uint8_t address[35],
                                   int *window, int *match);

/* The device's exact table behind the bitmap (per prefix length a sorted key
 * array and its directory) run on the host for one canonical key (sign bit
 * clear): *hit = whether the key's first 64 bits, masked to some length,
 * equal a key of that length.  Exact for prefixes of up to 12 symbols, a
 * "maybe" for longer ones.  No counterpart in the reference; it exists so that
 * the table is checked against the set's own match_key.  Returns 0, or -1 for
 * a null argument or a set without the table (not a dictionary of prefixes). */
int og_filters_exact_table_host(const og_filterset *s, const uint8_t packed[32], int *hit);

/* ------------------------------------------------------------------ batch */

/* Many independent searches in one launch.  No counterpart in the reference,
 * whose device path runs one search per engine (src/gpu/mod.rs:292-477); per
 * search the candidate numbering, survivor contract and errors are
 * Engine::advance's.  A batch keeps any number of resident searches on the
 * device (chain state, filter tables, next candidate) and runs `slots` of them
 * per step, round-robin, each on `chains` chains; one og_batch_step replays a
 * captured HIP graph that advances each scheduled search by
 * chains * (2*half+1) * rounds candidates. */
typedef struct og_batch og_batch;

typedef struct {
  uint32_t slots;    /* searches per step; 0 = 64 */
  uint32_t chains;   /* chains per search, a multiple of 256; 0 = 16384 */
  uint32_t half;     /* pairs per window; 0 = 256 */
  uint32_t rounds;   /* windows per chain per step; 0 = 1 */
  uint32_t ring_cap; /* survivor records per step; 0 = 65536 */
  uint32_t policy;   /* which resident searches take the slots: OG_POLICY_* */
} og_batch_opts;

enum {
  /* the searches that have examined the fewest candidates (ties: first
   * admitted first), so a new search runs at the next step */
  OG_POLICY_LEAST_SERVED = 0,
  /* in turn */
  OG_POLICY_ROUND_ROBIN = 1
};

typedef struct {
  uint64_t tag;       /* the tag the search was admitted with */
  uint64_t offset;    /* scalar delta from the search's s0: 8k */
  uint32_t search;    /* the search's id */
  uint8_t packed[32]; /* canonical y, sign bit clear */
} og_batch_hit;

/* With slots, chains and half all 0 the unit shape is the one og_batch_tune
 * recorded for this device (key og_tuning_device_key(<name>, "hip-batch", 0)),
 * else 64 x 16384 x 256. */
int og_batch_create(int ordinal, const og_batch_opts *opts, og_batch **out, char *err,
                    size_t errcap);
/* The slot engine's unit shape chosen by a timed walk, as og_hip_engine_tune
 * chooses the single engine's (src/tuning.rs:27-50, src/gpu/mod.rs:346-428): for
 * each of slots {32, 64, 128} x chains per slot {4096 .. 32768} x half {256, 512}
 * that fits `max_bytes` (0 = half the free memory), as many searches as slots -
 * the given filter sets in turn, each on its own chain - run one warm-up step and
 * then steps for at least 0.25 s.  A step bounds how soon an admitted search runs
 * and a found hit is reported, so only shapes whose steps average at most
 * `max_step_ms` (0 = 25 ms) take part; among those within 3% of the fastest
 * aggregate rate the smallest footprint is kept (og_tuning_choose), recorded
 * under og_tuning_device_key(<name>, "hip-batch", 0), and written to `best` with
 * its rate.  No counterpart in the reference, which has no slot engine. */
int og_batch_tune(int ordinal, const og_filterset *const *filters, size_t nfilters,
                  uint64_t max_bytes, double max_step_ms, og_batch_opts *best, double *rate,
                  char *err, size_t errcap);
/* Device bytes a resident search takes: its chain state and filter tables
 * (the admission footprint).  b may be null for the default chain count. */
size_t og_batch_footprint(const og_batch *b, const og_filterset *filters);
/* Makes a search (start point, filters, caller's tag) resident: its chains are
 * set up on the device and its candidate index starts at 0.  Writes its id. */
int og_batch_admit(og_batch *b, const uint8_t start_xyzt[128], const og_filterset *filters,
                   uint64_t tag, int *id, char *err, size_t errcap);
/* The search records at most `want` hits in all (default: every hit).  Hits
 * past it are still found and counted (og_batch_search_found) but not written
 * to the ring, so a search that needs n hits cannot flood it. */
int og_batch_set_want(og_batch *b, int id, uint64_t want);
/* Hits the search has found, recorded or not. */
uint64_t og_batch_search_found(const og_batch *b, int id);
/* Frees a resident search's device memory; it is not scheduled again. */
int og_batch_retire(og_batch *b, int id);
/* One step: `slots` resident searches, chosen by the policy, each advance;
 * writes the survivors.  Overflowing the ring or `cap` is an error. */
int og_batch_step(og_batch *b, og_batch_hit *out, uint32_t cap, uint32_t *n, char *err,
                  size_t errcap);
int og_batch_search_state(const og_batch *b, int id, uint64_t *tag, uint64_t *examined,
                          uint64_t *hits, int *live);
/* Candidates one step examines per unit slot a search takes. */
uint64_t og_batch_per_step(const og_batch *b);
/* Unit slots a resident search takes per step (its chains are this many times
 * the unit's); 0 when not resident.  With fewer resident searches than slots
 * each is laid out wider from its next candidate, so the device stays full. */
uint32_t og_batch_search_width(const og_batch *b, int id);
/* Candidates examined by every step so far, over all searches. */
uint64_t og_batch_examined(const og_batch *b);
/* Device bytes held by resident searches. */
uint64_t og_batch_resident_bytes(const og_batch *b);
/* Device time of the last step (HIP events around the graph launch). */
double og_batch_last_ms(const og_batch *b);
void og_batch_destroy(og_batch *b);

/* ------------------------------------------------------------------ serve */

/* A serve loop over one batch: requests arrive over time, wait in a bounded
 * queue, become resident first come first served while their device footprint
 * (chain state + filter tables) fits the budget, share the slots round-robin,
 * and are retired the step their n-th hit arrives.  No counterpart in the
 * reference (its nearest analogue is the master's lease queue,
 * src/master/queue.rs:51-175). */
typedef struct {
  double arrival_s;    /* seconds after the start, non-decreasing */
  uint8_t seed[32];    /* the search's seed: s0 = clamp(SHA-512(seed)[0..32]) */
  char filter[64];     /* filter text, NUL-terminated within the 64 bytes
                          (a request without the NUL is rejected) */
  uint32_t want_hits;  /* n >= 1: the request completes at its n-th hit
                          (0 is rejected) */
} og_request;

enum {
  OG_REQ_NOT_ARRIVED = 0,
  OG_REQ_QUEUED = 1,
  OG_REQ_RUNNING = 2,
  OG_REQ_DONE = 3,
  OG_REQ_REJECTED = 4
};

typedef struct {
  uint64_t offset;
  uint8_t packed[32];
} og_found;

typedef struct {
  int32_t status;       /* OG_REQ_* at the end of the run */
  uint32_t hits;        /* hits recorded: at most want_hits */
  uint64_t examined;
  double admitted_s, first_hit_s, done_s;  /* -1 when not reached */
  og_found found[16];   /* the first 16 recorded hits */
  uint64_t seen;        /* hits found over `examined`, recorded or not */
} og_request_result;

typedef struct {
  double wall_s, keys_per_s, device_busy_s;
  uint64_t examined, steps;
  uint32_t max_resident;
  uint32_t completed, running, queued, rejected, not_arrived;
  double p50_time_to_n, p99_time_to_n;          /* over completed requests */
  double p50_time_to_first, p99_time_to_first;  /* over requests with a hit */
} og_serve_stats;

/* Runs the loop for `duration_s` seconds of wall time (or until every request
 * is done).  max_queue bounds the waiting queue; device_budget bounds the
 * device bytes of resident searches (0 = 4 GiB).  A request is rejected
 * (OG_REQ_REJECTED) when it arrives to a full queue, when its filter does not
 * parse, when it asks for no hit, or when its footprint alone exceeds the
 * budget; a rejected request never holds up the queue behind it. */
int og_serve_run(int ordinal, const og_batch_opts *opts, const og_request *reqs, size_t n,
                 double duration_s, uint32_t max_queue, uint64_t device_budget,
                 og_request_result *results, og_serve_stats *stats, char *err, size_t errcap);

/* ----------------------------------------------------------------- tuning */



/* Replaces tuning::device_key (src/tuning.rs:65-74): "device/<name>/<api>/
 * <threads>/<version>" with '/' and newline in the name replaced by '_'.
 * Copies at most cap-1 bytes and a terminator; returns the key's length. */
size_t og_tuning_device_key(const char *name, const char *api, uint32_t threads, char *out,
                            size_t cap);
/* The version the keys end in: the reference crate's CARGO_PKG_VERSION
 * (Cargo.toml:3, "1.0.2") unless the host sets its own, as a Rust host that
 * links this library passes env!("CARGO_PKG_VERSION").  Returns 0, or -1 for
 * an empty version, one over 64 bytes or one holding '/' or a newline. */
int og_tuning_set_version(const char *version);
/* The shape the walk keeps, from the measured rates (0 = not measured) and
 * footprints of n shapes: the fastest within `budget` bytes (0 = no bound),
 * then, among those within `tolerance` of its rate and the budget, the
 * smallest footprint.  Returns its index, or (size_t)-1 when none was
 * measured.  No counterpart in the reference, whose walk keeps the fastest
 * (src/gpu/mod.rs:383-411). */
size_t og_tuning_choose(const double *rate, const uint64_t *bytes, size_t n, uint64_t budget,
                        double tolerance);

/* Replaces tuning::path (src/tuning.rs:80-91); OG_TUNING_PATH overrides. */
int og_tuning_path(char *out, size_t cap);
/* Replaces tuning::get (src/tuning.rs:94-97): 1 when an entry exists. */
int og_tuning_get(const char *key, og_tuning_shape *out);
/* Replaces tuning::put (src/tuning.rs:103-126): atomic rewrite by rename. */
int og_tuning_put(const char *key, const og_tuning_shape *shape);

/* -------------------------------------------------------------------- key */

/* src/key.rs:23-31: SHA-512(seed)[0..32], clamped. */
void og_key_secret_scalar(const uint8_t seed[32], uint8_t sk[32]);
/* src/key.rs:33-40: both halves, the first clamped. */
void og_key_expanded_secret(const uint8_t seed[32], uint8_t out[64]);
/* s0*B of a seed as 128 bytes xyzt (src/batch.rs:165-168 start_point). */
void og_key_start_point(const uint8_t seed[32], uint8_t xyzt[128]);
/* scalar*B packed with its sign bit (src/key.rs:55-62 over curve::scalar_base_mult). */
void og_key_public_from_scalar(const uint8_t scalar[32], uint8_t pk[32]);
/* src/key.rs:167-177: the 64-byte secret of the candidate at `offset`;
 * returns 0 when the offset broke clamping (no key). */
int og_key_expanded_secret_at_offset(const uint8_t seed[32], uint64_t offset, uint8_t out[64]);
/* src/key.rs:126-132 and :116: pk || checksum || 0x03, and its base32 text. */
void og_key_address_bytes(const uint8_t pk[32], uint8_t out[35]);
void og_key_address(const uint8_t pk[32], char out[57]);
/* src/run.rs:211-220 */
void og_run_block_seed(const uint8_t root[32], uint64_t index, uint8_t out[32]);

#ifdef __cplusplus
}
#endif
#endif
