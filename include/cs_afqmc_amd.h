/*
 * cs_afqmc_amd -- C ABI of the MI355X (gfx950) phaseless-AFQMC propagation step.
 *
 * The step replaces ad_afqmc's propagator / trial measurement pytree functions
 * (reference files cited per entry point, paths relative to the reference root).
 * Plain pointers and sizes only; no torch or jax types cross this boundary.
 *
 * Conventions
 *   - Every entry point returns AFQMC_OK (0) or a nonzero status; the message of the
 *     last failure on the calling thread is returned by afqmc_last_error().
 *     Numerical pathologies (NaN overlaps, collapsed weights) are not errors: they
 *     follow the reference's zeroing rules (propagation.py:149-160).
 *   - "dev" pointers are device memory on the context's device; they are owned by
 *     the caller.  Host pointers are read during the call only.
 *   - Complex numbers are interleaved (re, im) float64 pairs (complex128).
 *   - Work is enqueued on the context's stream; no entry point except
 *     afqmc_ctx_create / afqmc_ctx_destroy synchronises the host.
 *   - Walkers: restricted  -> one array (W, norb, nocc) complex128 (pass it as
 *     walkers_up, walkers_dn = NULL); unrestricted -> up (W, norb, nocc_up) and
 *     dn (W, norb, nocc_dn).
 *   - Calls that share one context must not overlap (the context owns one workspace).
 */
#ifndef CS_AFQMC_AMD_H
#define CS_AFQMC_AMD_H

#include <stdint.h>

#if defined(__GNUC__) || defined(__clang__)
#define AFQMC_API __attribute__((visibility("default")))
#else
#define AFQMC_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afqmc_ctx afqmc_ctx;

enum {
  AFQMC_OK = 0,
  AFQMC_ERR_ARG = 1,     /* bad argument (null pointer, size out of range) */
  AFQMC_ERR_HIP = 2,     /* HIP runtime error */
  AFQMC_ERR_SIZE = 3,    /* problem larger than the kernels support */
  AFQMC_ERR_STATE = 4    /* call not valid in the context's state */
};

/* walker_type: propagator_restricted + rhf trial, or propagator_unrestricted + uhf
 * trial (mpi_jax.py:314-352 picks the pair from options["walker_type"/"trial"]), or the
 * unrestricted propagator and trial of prop_unrestricted/ with separate alpha and beta
 * Cholesky vectors (prop_unrestricted/propagation.py:361-400, :461-510,
 * wavefunctions_unrestricted.py:453-475): AFQMC_UHF_SPIN_CHOL, with ham->chol holding
 * (2, G, N*N), alpha then beta, and rot_chol / exp_h1 of each spin formed from that
 * spin's vectors.  Each spin's VHS is then built from its own vectors. */
enum { AFQMC_RHF = 1, AFQMC_UHF = 2, AFQMC_UHF_SPIN_CHOL = 3 };

/* Precision of the Cholesky contractions (force bias, VHS, energy) and of the
 * Taylor/one-body products.  Overlaps, determinants, QR, weights and energy
 * accumulation are float64 in every mode. */
enum { AFQMC_PREC_FP64 = 0, AFQMC_PREC_FP32 = 1, AFQMC_PREC_BF16X3 = 2 };

typedef struct afqmc_dims {
  int32_t walker_type;  /* AFQMC_RHF or AFQMC_UHF */
  int32_t norb;         /* N  (<= 256) */
  int32_t nocc_up;      /* n_up (<= 64) */
  int32_t nocc_dn;      /* n_dn (<= 64; == nocc_up for AFQMC_RHF) */
  int32_t nchol;        /* G */
  int32_t n_walkers;    /* W: workspace is sized for this many walkers */
  int32_t n_exp_terms;  /* Taylor terms incl. identity (propagation.py:208, default 6) */
  int32_t precision;    /* AFQMC_PREC_* */
  double dt;            /* time step */
} afqmc_dims;

/* Float64 intermediates, laid out as the reference's ham_data / wave_data after
 * build_measurement_intermediates (wavefunctions.py:602-619, :817-841) and
 * build_propagation_intermediates (propagation.py:292-325, :464-509).  Each pointer
 * may address host memory or device memory of the context's device (copied with
 * hipMemcpyDefault on the creation stream; device arrays must be ready on it). */
typedef struct afqmc_ham_host {
  double h0;               /* ham_data["h0"]                                        */
  double h0_prop;          /* Re ham_data["h0_prop"]                                */
  const double* chol;      /* ham_data["chol"]      (G, N*N); AFQMC_UHF_SPIN_CHOL:
                              (2, G, N*N), alpha then beta                          */
  const double* mo_coeff;  /* wave_data["mo_coeff"] rhf (N,n); uhf up (N,n_up)
                              followed by dn (N,n_dn)                               */
  const double* rot_h1;    /* ham_data["rot_h1"]    rhf (n,N); uhf up then dn       */
  const double* rot_chol;  /* ham_data["rot_chol"]  rhf (G,n,N); uhf up (G,n_up,N)
                              followed by dn (G,n_dn,N)                             */
  const double* mf_shifts; /* ham_data["mf_shifts"] (G) complex128                   */
  const double* exp_h1;    /* ham_data["exp_h1"]    rhf (N,N); uhf (2,N,N)           */
} afqmc_ham_host;

/* Create a context: copies the intermediates to the device (plus the transposed and
 * reduced-precision copies the kernels use) and allocates the step workspace for
 * dims->n_walkers walkers.  stream: a hipStream_t (NULL = default stream).
 * Replaces the jit-time setup of propagator.propagate's closure over ham_data
 * (propagation.py:100-165, static args propagation.py:327-328). */
AFQMC_API int afqmc_ctx_create(afqmc_ctx** out, const afqmc_dims* dims,
                     const afqmc_ham_host* ham, void* stream);
AFQMC_API int afqmc_ctx_destroy(afqmc_ctx* ctx);
AFQMC_API int afqmc_set_stream(afqmc_ctx* ctx, void* stream);
/* Bytes of device memory the context holds (intermediates + workspace). */
AFQMC_API int64_t afqmc_ctx_device_bytes(const afqmc_ctx* ctx);
/* Entries of one stored VHS plane: norb*norb, or norb(norb+1)/2 when the mixed mode
 * keeps V as its upper triangle (symmetric Cholesky vectors, register Taylor kernel).
 * No counterpart in the reference: a layout query for measurement tools. */
AFQMC_API int64_t afqmc_ctx_vhs_entries(const afqmc_ctx* ctx);
AFQMC_API const char* afqmc_last_error(void);

/* One phaseless step, propagator.propagate (propagation.py:100-165) with
 * trial.calc_force_bias (wavefunctions.py:151-165 -> :503-511 / :694-709),
 * _apply_trotprop (propagation.py:255-290 / :359-400) and trial.calc_overlap
 * (wavefunctions.py:82-96 -> :478-480 / :658-667).
 *   walkers_up/dn  dev  updated in place
 *   weights        dev  (W) float64, updated in place
 *   overlaps       dev  (W) complex128, in: old overlaps, out: new overlaps
 *   fields         dev  (W, G) float64 auxiliary fields
 *   scalars        dev  float64[2] = {e_estimate, pop_control_ene_shift};
 *                       [1] is updated as in propagation.py:161-163
 *   force_bias_out dev  (W, G) complex128 or NULL (debug / parity output)
 *   flags          AFQMC_STEP_REUSE_GREEN: the walkers are exactly those the previous
 *                  afqmc_propagate on this context returned, so the Green's function
 *                  it left in the workspace is reused for the force bias (the library
 *                  also requires the same walker pointer and count, else recomputes)
 * n_walkers may be smaller than the context's workspace size. */
enum { AFQMC_STEP_REUSE_GREEN = 1 };
AFQMC_API int afqmc_propagate(afqmc_ctx* ctx, int32_t n_walkers, void* walkers_up,
                    void* walkers_dn, double* weights, void* overlaps,
                    const double* fields, double* scalars, void* force_bias_out,
                    int32_t flags);

/* trial.calc_overlap (wavefunctions.py:82-96 / :61-80): overlaps (W) complex128. */
AFQMC_API int afqmc_overlap(afqmc_ctx* ctx, int32_t n_walkers, const void* walkers_up,
                  const void* walkers_dn, void* overlaps);

/* trial.calc_force_bias (wavefunctions.py:151-165 / :127-149): fb (W, G) complex128. */
AFQMC_API int afqmc_force_bias(afqmc_ctx* ctx, int32_t n_walkers, const void* walkers_up,
                     const void* walkers_dn, void* force_bias);

/* trial.calc_energy (wavefunctions.py:225-241 / :203-223) -> rhf._calc_energy_restricted
 * (wavefunctions.py:529-541) / uhf._calc_energy (:711-743): energies (W) complex128. */
AFQMC_API int afqmc_energy(afqmc_ctx* ctx, int32_t n_walkers, const void* walkers_up,
                 const void* walkers_dn, void* energies);

/* propagator.orthonormalize_walkers (propagation.py:251-253 / :422-424 ->
 * linalg_utils.qr_vmap :51-55 / qr_vmap_uhf :58-65): walkers <- Q of the reduced
 * Householder QR (LAPACK zgeqr2 + zung2r convention), in place. */
AFQMC_API int afqmc_orthonormalize(afqmc_ctx* ctx, int32_t n_walkers, void* walkers_up,
                         void* walkers_dn);

/* stochastic_reconfiguration (sr.py:24-34 / :37-48): systematic resampling with
 * the uniform draw zeta.  walkers gathered in place, weights <- sum|w| / W.
 * idx_out dev (W) int32 or NULL: the resampled indices.
 * flags: 0, or AFQMC_SR_GLOBAL_ORDER for the gather/scatter form that
 * propagator.stochastic_reconfiguration_global uses (propagation.py:243-249 ->
 * sr.py:66-106, here with the whole population on one device), which forms the
 * comb points as ((i + zeta) / W) * total instead of total * (i + zeta) / W. */
enum { AFQMC_SR_GLOBAL_ORDER = 1 };
AFQMC_API int afqmc_sr(afqmc_ctx* ctx, int32_t n_walkers, void* walkers_up, void* walkers_dn,
             double* weights, double zeta, int32_t* idx_out, int32_t flags);

/* afqmc_sr with the uniform draw taken on the device from a device-resident key
 * (uint64[2] = {seed, counter}), as stochastic_reconfiguration_local does with
 * random.split + random.uniform (propagation.py:234-241): zeta is the draw of
 * (seed, counter) and counter advances by one.  Nothing is read back to the host, so a
 * block that ends with SR can be captured into a HIP graph and replayed. */
AFQMC_API int afqmc_sr_key(afqmc_ctx* ctx, int32_t n_walkers, void* walkers_up,
                           void* walkers_dn, double* weights, uint64_t* key, int32_t* idx_out,
                           int32_t flags);

/* The resampling of afqmc_sr without a context and without allocating: no counterpart in
 * the reference.  It is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, foreign runtimes that own all buffers).  Same
 * indices, weights and walker copies as afqmc_sr on the same weights and zeta, but the
 * walkers are gathered out of place:
 *   walkers_up / walkers_dn  dev  inputs, n_walkers x elems_up / elems_dn complex128
 *                                 (elems = norb * nocc of that spin; elems_dn = 0 and
 *                                 walkers_dn = out_dn = NULL for restricted walkers)
 *   out_up / out_dn          dev  outputs of the same sizes, distinct from the inputs
 *   weights                  dev  (W) float64, in place: sum|w| / W
 *   idx_out                  dev  (W) int32 or NULL
 *   stream                       hipStream_t of the caller (NULL = default stream)
 * The cumulative weights and indices live in a fixed array the library keeps on every
 * device it runs on (12 MiB), so n_walkers is at most 2^20 (AFQMC_ERR_SIZE otherwise).
 * Every call on one device shares that array: calls that overlap in time on one device
 * (different streams or threads) race and must be ordered by the caller. */
AFQMC_API int afqmc_sr_static(int32_t n_walkers, int32_t elems_up, int32_t elems_dn,
                              const void* walkers_up, const void* walkers_dn, void* out_up,
                              void* out_dn, double* weights, double zeta, int32_t* idx_out,
                              int32_t flags, void* stream);

/* Block statistics of sampler._block_scan (sampling.py:74-91), on the device:
 *   energies dev (W) complex128 local energies; the real part is clamped to
 *            e_estimate when |E - e_estimate| > sqrt(2 / dt)
 *   weights  dev (W) float64
 *   scalars  dev float64[2] = {e_estimate, pop_control_ene_shift};
 *            [1] <- 0.9 [1] + 0.1 block_energy
 *   out      dev float64[AFQMC_BLOCK_STATS_LEN] = {block_energy, block_weight,
 *            number of zero weights, number of clamped local energies} */
enum { AFQMC_BLOCK_STATS_LEN = 4 };
AFQMC_API int afqmc_block_stats(afqmc_ctx* ctx, int32_t n_walkers, const void* energies,
                                const double* weights, double* scalars, double* out);

/* The driver's estimate update after each sampler call (driver.py:308; the fragment
 * driver lno_afqmc/ccsd_pt2/run_afqmc_nompi.py:89): scalars[0] = e_estimate <-
 * 0.9 e_estimate + 0.1 E, E = sum_b E_b w_b / sum_b w_b over nblocks consecutive rows
 * of afqmc_block_stats output (stats dev, nblocks * AFQMC_BLOCK_STATS_LEN float64),
 * summed in row order; then scalars[1] = pop_control_ene_shift <- the new e_estimate,
 * as the next sampler call starts (sampling.py:329).  Stream-ordered; no host
 * synchronisation. */
AFQMC_API int afqmc_estimate_mix(afqmc_ctx* ctx, double* scalars, const double* stats,
                                 int32_t nblocks);

/* Per-phase device timing with HIP events on the context's stream (rocprof-free
 * measurement for benchmarks).  While enabled, each phase of afqmc_propagate and the
 * block entry points records an event pair; afqmc_prof_collect (after the stream
 * has drained) adds the elapsed milliseconds and launch counts per phase into
 * ms[AFQMC_NPHASE] / counts[AFQMC_NPHASE] and clears the record. */
enum {
  AFQMC_PH_GREEN = 0,    /* Green's function of the current walkers   */
  AFQMC_PH_FB_GEMM = 1,  /* force-bias contraction                    */
  AFQMC_PH_FIELDS = 2,   /* field shifts, shifted fields              */
  AFQMC_PH_VHS_GEMM = 3, /* VHS contraction                           */
  AFQMC_PH_TAYLOR = 4,   /* exp(VHS) Taylor + one-body half steps     */
  AFQMC_PH_OVERLAP = 5,  /* new overlaps                              */
  AFQMC_PH_WEIGHTS = 6,  /* importance function, weights, E shift     */
  AFQMC_PH_QR = 7,
  AFQMC_PH_ENERGY = 8,
  AFQMC_PH_SR = 9,
  AFQMC_PH_RNG = 10,
  AFQMC_NPHASE = 11
};
AFQMC_API int afqmc_prof_enable(afqmc_ctx* ctx, int32_t on);
AFQMC_API int afqmc_prof_collect(afqmc_ctx* ctx, double* ms, int64_t* counts);
/* Restricts the event pairs to the phases whose bit (1 << AFQMC_PH_*) is set in mask;
 * 0 (the default) records every phase.  Each pair costs a few microseconds of device
 * time, so a throughput measurement records only the phase it reports on. */
AFQMC_API int afqmc_prof_phases(afqmc_ctx* ctx, uint32_t mask);

/* Correlated sampling outlier rule, corr_sample.corr_otler_rm (corr_sample.py:551-585):
 * for two populations of n_walkers propagated with the same fields, z_i =
 * |w_i - mean(w)| / std(w) (population std) per population; walkers whose larger
 * z-score exceeds z_thresh get weight 0 in both.  weights1 / weights2 dev (W) float64,
 * updated in place on ctx's stream. */
AFQMC_API int afqmc_cs_outliers(afqmc_ctx* ctx, int32_t n_walkers, double* weights1,
                                double* weights2, double z_thresh);

/* Auxiliary fields for one step: count float64 standard normals written to fields
 * (dev), from a counter-based generator (Philox4x32-10 + Box-Muller) keyed by
 * (seed, counter).  Stands in for jax.random.normal in sampling.py:61-69. */
AFQMC_API int afqmc_fields_normal(afqmc_ctx* ctx, uint64_t seed, uint64_t counter,
                        double* fields, int64_t count);
/* The same draws keyed by a device-resident key (uint64[2] = {seed, counter}): the
 * block's random.split + random.normal of sampling.py:61-69 with the key kept in device
 * memory, so the draw can sit inside a captured graph.  Uses (seed, counter), then
 * advances counter by one (stream-ordered). */
AFQMC_API int afqmc_fields_normal_key(afqmc_ctx* ctx, uint64_t* key, double* fields,
                                      int64_t count);

/* ---- Fragment groups: many independent problems advanced by one grouped step.
 *
 * No single counterpart in the reference: it replaces the loop a fragment driver runs
 * over propagator.propagate (propagation.py:100-165) of each LNO fragment
 * (lno_afqmc/lno_afqmc.py:442-605 runs one AFQMC per fragment), with every fragment's
 * step in the same launches.  Each fragment keeps its own context (integrals,
 * workspace) and caller-owned state; afqmc_group_add binds that state:
 *   walkers_up / walkers_dn, weights, overlaps, scalars: as for afqmc_propagate, with
 *     n_walkers = the context's capacity;
 *   fields_block dev (n_steps, W, nchol) float64: step s of afqmc_group_propagate
 *     reads fields_block + s * W * nchol.
 * All fragments of a group share one precision mode and device.  A fragment's result
 * does not depend on the other members (no split-K, fixed k order in every GEMM).
 * Calls on a group must not overlap in time with calls on its member contexts. */
typedef struct afqmc_group afqmc_group;
AFQMC_API int afqmc_group_create(afqmc_group** out, int32_t max_fragments, void* stream);
AFQMC_API int afqmc_group_destroy(afqmc_group* g);
AFQMC_API int afqmc_group_set_stream(afqmc_group* g, void* stream);
AFQMC_API int afqmc_group_clear(afqmc_group* g);
AFQMC_API int afqmc_group_add(afqmc_group* g, afqmc_ctx* ctx, void* walkers_up,
                              void* walkers_dn, double* weights, void* overlaps,
                              const double* fields_block, int32_t n_steps, double* scalars);
AFQMC_API int afqmc_group_size(const afqmc_group* g);
/* One phaseless step of every member (flags: 0 or AFQMC_STEP_REUSE_GREEN). */
AFQMC_API int afqmc_group_propagate(afqmc_group* g, int32_t step, int32_t flags);
/* Block ends of every member in grouped launches; a member's result does not depend on
 * the other members:
 *   afqmc_group_orthonormalize: afqmc_orthonormalize of each member's walkers
 *     (propagation.py:251-253, linalg_utils.py:51-65 per fragment);
 *   afqmc_group_overlap: afqmc_overlap into each member's bound overlaps
 *     (wavefunctions.py:478-486 / :658-670, trial.calc_overlap per fragment);
 *   afqmc_group_energy: afqmc_energy of each member's walkers into energies[i]
 *     (dev, W complex128, i in the order of afqmc_group_add; wavefunctions.py:529-541 /
 *     :711-743, trial.calc_energy per fragment).
 * The first two give exactly the per-context results.  The energy's force-bias part c_g
 * is contracted in float64 without split-K, so it may differ from afqmc_energy by
 * rounding.
 * They replace one launch sequence per fragment at the end of every block
 * (sampling.py:74-91). */
AFQMC_API int afqmc_group_orthonormalize(afqmc_group* g);
AFQMC_API int afqmc_group_overlap(afqmc_group* g);
AFQMC_API int afqmc_group_energy(afqmc_group* g, void* const* energies);
/* Every member's fields block from its own device key: keys[i] (dev uint64[2] {seed,
 * counter}) draws n_steps x W x nchol normals into member i's bound fields_block, then
 * counter += 1 -- afqmc_fields_normal_key per member (the sampler's per-block draw,
 * sampling.py:60-69), bit for bit, in two launches. */
AFQMC_API int afqmc_group_fields_normal_key(afqmc_group* g, void* const* keys);
/* Per-phase event timing of the grouped step, as afqmc_prof_enable / _phases / _collect
 * for one context (phases AFQMC_PH_GREEN .. AFQMC_PH_WEIGHTS, each phase covering all
 * of its grouped launches).  No counterpart in the reference: measurement only. */
AFQMC_API int afqmc_group_prof_enable(afqmc_group* g, int32_t on);
AFQMC_API int afqmc_group_prof_phases(afqmc_group* g, uint32_t mask);
AFQMC_API int afqmc_group_prof_collect(afqmc_group* g, double* ms, int64_t* counts);
/* Picks the GEMM tile edge (64 or 128) of the force-bias and VHS launches and of the
 * energy's float64 c_g product by timing both edges on the current members; ms_out
 * (6 float64, may be NULL) receives the times [fb 64, fb 128, vhs 64, vhs 128, c_g 64,
 * c_g 128].  The edge does not change results.  Synchronises. */
AFQMC_API int afqmc_group_tune(afqmc_group* g, double* ms_out);

#ifdef __cplusplus
}
#endif

#endif /* CS_AFQMC_AMD_H */
