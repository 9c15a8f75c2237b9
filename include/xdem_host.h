/* xdem_host.h — host-side library (libxdem_host.so): seeded input generators and the deck runner.
 *
 * The deck runner reads exaNBody/exaDEM decks (.msp, YAML) with the operator names and slots of
 * the reference (data/config/ and the example decks) and executes the hot-path operators
 * through the C ABI of libxdem_hip.so (include/xdem.h). Operators that are not part of the hot
 * path (logging, dumps, analyses) are recognised and skipped; operators whose physics is not
 * provided (cohesive laws, moving drivers, cylinders, ...) stop the run with an error, as the
 * reference's operators stop on a fatal error.
 */
#ifndef XDEM_HOST_H
#define XDEM_HOST_H

#include <stdint.h>

#include "xdem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Seeded generators (SplitMix64 counter streams): value k of stream (seed, tag). */
void xdem_gen_uniform(uint64_t seed, uint64_t tag, int64_t n, double a, double b, double* out);
void xdem_gen_normal(uint64_t seed, uint64_t tag, int64_t n, double mean, double sigma, double* out);
/* Random sequential addition at radius scale `scale` (e.g. 0.8), largest first, in the box [lo, hi)
 * periodic on every axis, then Jacobi push-apart sweeps at full radius until the largest overlap is
 * below tol or max_sweeps; writes positions, returns the number of sweeps; *max_overlap receives the
 * largest remaining overlap. */
int64_t xdem_gen_pack_periodic(uint64_t seed, int64_t n, const double* rad, const double* lo, const double* hi,
                               double scale, int64_t max_sweeps, double tol, double* x, double* y, double* z,
                               double* max_overlap);

typedef struct xdem_deck xdem_deck;

typedef struct xdem_deck_stats {
  int64_t particles;     /* particles created by the input operators                  */
  int64_t iterations;    /* time steps executed by compute_loop                       */
  int64_t rebuilds;      /* neighbour rebuilds (check_and_update_particles triggers)   */
  int64_t operator_calls;/* native operator executions                                */
  double seconds;        /* wall time of the whole run                                */
} xdem_deck_stats;

/* Loads a deck: the base configuration (main-config.msp found in the search directories) then the
 * user deck, includes resolved relative to the including file and then to each search directory.
 * Returns 0 or a negative XDEM_ERR_* code; *out is set on success. */
int xdem_deck_load(const char* path, const char* const* search_dirs, int32_t n_dirs, xdem_deck** out);
/* Overrides one global slot (as "global: { name: value }" would), e.g. simulation_end_iteration. */
int xdem_deck_set_global(xdem_deck* deck, const char* name, const char* value);
/* Runs the "simulation" operator of the deck on a context created by the caller (the context keeps
 * the final state for downloads). */
int xdem_deck_run(xdem_deck* deck, xdem_ctx* ctx, xdem_deck_stats* stats);
/* Text of the last error of the deck (load, parse or run). */
int xdem_deck_last_error(const xdem_deck* deck, char* buf, int64_t n);
void xdem_deck_destroy(xdem_deck* deck);

#ifdef __cplusplus
}
#endif

#endif /* XDEM_HOST_H */
