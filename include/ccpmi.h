/*
 * ccpmi — MI355X-native Brownian-dynamics engine that drops in for the
 * simulation path of ccp (Marco-Mendivil-Carboni/TFM, Program/src).
 *
 * C ABI: plain pointers and sizes, no exceptions cross it. Every call
 * returns CCP_OK (0) or an error code; the message of the last error on a
 * handle is ccp_last_error(h) (ccp_last_error(NULL) for ccp_create). The
 * messages of parameter errors are ccp's own (chrdat.cu:21-63,
 * chrsim.cu:428-435), so scripts that grep them keep working.
 *
 * One handle = one simulation = one HIP stream on the current device; the
 * handle owns all its device memory. Calls on one handle are not
 * thread-safe; distinct handles are independent.
 *
 * Reference interfaces replaced (ccp has no FFI; its seams are the C++ class
 * mmc::chrsim, chrsim.cuh:36-57, and the ccp-perform process boundary,
 * ccp-perform.cu:43-112). Each function cites the member it stands in for.
 */
#ifndef CCPMI_H
#define CCPMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCP_OK 0
#define CCP_ERR_PARAM 1 /* invalid parameter (ccp: mmc::error from chrdat/chrsim ctors) */
#define CCP_ERR_HIP 2   /* HIP runtime failure (ccp: cuda_check, util.cu:83-91) */
#define CCP_ERR_IO 3    /* file open/read/write failure (ccp: check_file, stream checks) */
#define CCP_ERR_STATE 4 /* device error flag: cell overflow, bead outside grid, non-finite */
#define CCP_ERR_ARG 5   /* bad argument to this API (size mismatch, NULL) */

/* Per-bead flag bits. ccp's ptype LADh ("A") is CCP_HET|CCP_LAD, LNDe ("B") is 0. */
#define CCP_HET 1u
#define CCP_LAD 2u

/* Force fields. CCP is ccp's default simulation type (DST); LJ is the
 * optional WCA/LJ/lamina-well set selected by "force_field lj". */
#define CCP_FF_CCP 0
#define CCP_FF_ICG 1
#define CCP_FF_LJ 2

typedef struct ccp_sim ccp_sim;

typedef struct ccp_info {
  uint32_t N;           /* beads */
  uint32_t n_l;         /* lamina binding sites */
  uint32_t n_chains;
  uint32_t force_field; /* CCP_FF_CCP or CCP_FF_LJ */
  uint32_t cps;         /* cells per side of the coarse grid (ccp's sugrid geometry) */
  uint32_t n_c;         /* its cells = cps^3 (size of ccp_cell_list's beg/end) */
  uint32_t cell_capacity; /* coarse grid bucket capacity */
  uint32_t fine_cps;    /* cells per side of the fine grid of the step */
  uint32_t spf;         /* steps per frame */
  uint32_t fpf;         /* frames per file */
  uint32_t contact_bins;     /* contact map side */
  uint32_t contact_bin_size; /* beads per contact bin */
  uint32_t contact_interval; /* steps between contact samples (0 = off) */
  float csl;            /* coarse cell side (cell_size) */
  float fine_csl;       /* fine cell side (fine_cell_size) */
  float R_n, R_o, R_b;
  float T, dt, sd;      /* temperature, timestep, noise standard deviation */
  float contact_cutoff;
  float force_cap;      /* integrator force cap (0 = off; the LJ set's safeguard) */
  uint64_t seed;
} ccp_info;

/* Logger callback: one line per call (no timestamp; the callee adds it). */
typedef void (*ccp_log_fn)(const char *line, void *user);

/* Create a simulation from the text of adjustable-parameters.dat
 * (whitespace-separated "key value" pairs; unknown keys ignored, first
 * occurrence wins). Replaces chrsim::chrsim(parmap&) (chrsim.cu:421-455)
 * with chrdat::chrdat (chrdat.cu:40-88), cngeom (chrdat.cu:13-37) and
 * parmap (util.cu:72-80). Returns NULL on error. */
ccp_sim *ccp_create(const char *param_text);

/* ccp_create with the logger installed before the parameter echo lines
 * (ccp-perform sets its log file before constructing chrsim,
 * ccp-perform.cu:34-48). */
ccp_sim *ccp_create_logged(const char *param_text, ccp_log_fn fn, void *user);

/* Release device and host memory (~chrsim/~chrdat, chrsim.cu:458-467). */
void ccp_destroy(ccp_sim *h);

/* Message of the last error on h, or of the last failed ccp_create when h is NULL. */
const char *ccp_last_error(const ccp_sim *h);

/* Route log lines (parameter echo, progress, events) to fn instead of
 * stdout (mmc::logger, util.cu:15-50). fn NULL restores stdout. */
int ccp_set_logger(ccp_sim *h, ccp_log_fn fn, void *user);

int ccp_get_info(const ccp_sim *h, ccp_info *out);

/* Random LBS sites, bead types, confined random walk and overlap removal
 * (chrsim::generate_initial_condition, chrsim.cu:470-522, with
 * set_lbs_positions :596-631, set_particle_types :634-675,
 * set_particle_positons :678-716, perform_random_walk :719-802,
 * particle_overlaps :805-820). */
int ccp_generate_initial_condition(ccp_sim *h);

/* chrsim::save_checkpoint / load_checkpoint (chrsim.cu:525-555): i_f u32,
 * t f32, types N x i32, r N x 3 f32, lbs n_l x 3 f32, then a trailer that
 * ccp ignores (magic, seed, step, flags) so resumes continue the Philox
 * stream bit-exactly. Files without the trailer load as in ccp. */
int ccp_save_checkpoint(ccp_sim *h, const char *path);
int ccp_load_checkpoint(ccp_sim *h, const char *path);

/* .gro text frame and LBS file (chrdat::write_frame_txt :107-133,
 * read_frame_txt :136-165, write_lbs_txt :229-254, read_lbs_txt :257-284). */
int ccp_write_frame_txt(ccp_sim *h, const char *path);
int ccp_read_frame_txt(ccp_sim *h, const char *path);
int ccp_write_lbs_txt(ccp_sim *h, const char *path);
int ccp_read_lbs_txt(ccp_sim *h, const char *path);

/* chrsim::run_simulation (chrsim.cu:558-593): fpf frames of spf steps; each
 * frame is appended to the minimal .trr file at trr_path (truncated first,
 * chrdat::write_frame_bin :168-201). trr_path NULL runs without output. */
int ccp_run_simulation(ccp_sim *h, const char *trr_path);

/* Frames of spf steps handed to the caller instead of a file: after each
 * frame the host positions (xyz, n beads) go to cb together with the frame
 * index and time, which advance as in ccp_run_simulation. A nonzero return
 * from cb stops the run (status CCP_OK). Replaces chrsim::run_simulation's
 * frame writes (chrsim.cu:558-593) for callers that consume frames in memory. */
typedef int (*ccp_frame_fn)(uint32_t i_f, float t, const float *xyz, size_t n, void *user);
int ccp_run_frames(ccp_sim *h, uint32_t n_frames, ccp_frame_fn cb, void *user);

/* The step loop alone (chrsim.cu:570-580): n stochastic-Heun steps with
 * contact sampling every contact_interval steps; no host transfer. */
int ccp_run_steps(ccp_sim *h, uint64_t n);

/* Wait for the handle's stream and report device error flags. */
int ccp_synchronize(ccp_sim *h);

/* Frame index, time and global step counter (chrdat i_f, t). */
int ccp_get_time(const ccp_sim *h, uint32_t *i_f, float *t, uint64_t *step);
int ccp_set_time(ccp_sim *h, uint32_t i_f, float t, uint64_t step);

/* Host copies of the state (xyz interleaved floats; n = count of beads or
 * sites). Types use ccp's ptype codes: 0 = LADh (A), 1 = LNDe (B). */
int ccp_get_positions(ccp_sim *h, float *xyz, size_t n);
int ccp_set_positions(ccp_sim *h, const float *xyz, size_t n);
int ccp_get_types(ccp_sim *h, int32_t *pt, size_t n);
int ccp_set_types(ccp_sim *h, const int32_t *pt, size_t n);
int ccp_get_flags(ccp_sim *h, uint8_t *flags, size_t n);
int ccp_set_flags(ccp_sim *h, const uint8_t *flags, size_t n);
int ccp_get_lbs(ccp_sim *h, float *xyz, size_t n_l);
int ccp_set_lbs(ccp_sim *h, const float *xyz, size_t n_l);
/* chain offsets, n_chains + 1 entries */
int ccp_get_chains(ccp_sim *h, uint32_t *off, size_t n);

/* One force evaluation F(r) of the current positions, as exec_RK_1 computes
 * it (calc_bf + calc_cf + calc_srf, chrsim.cu:381-384); ff selects
 * CCP_FF_CCP/LJ (the handle's own) or CCP_FF_ICG. */
int ccp_compute_forces(ccp_sim *h, int ff, float *f_xyz, size_t n);

/* The cell list of the current positions in ccp's layout
 * (sugrid::generate_arrays, sugrid.cu:125-133): sci/soi sorted keys and
 * bead indices (n = N), beg/end per cell (n_c entries, empty cells 0/0). */
int ccp_cell_list(ccp_sim *h, uint32_t *sci, uint32_t *soi, uint32_t *beg, uint32_t *end, size_t n, size_t n_c);

/* The cell list the step itself uses for the current positions (the grid
 * of cell_size, or fine_cell_size when split, whose lookup multiplies by
 * 1/cell_size): cstart (fine_cps^3 + 1 entries) the first slot of every
 * cell, soi (N) the bead index of every slot. The last Heun stage bins and
 * orders it; after ccp_set_positions it is built afresh. Stands in for the
 * sorted arrays of sugrid::generate_arrays (sugrid.cu:125-133) as the step
 * sees them. */
int ccp_step_grid(ccp_sim *h, uint32_t *cstart, size_t n_c1, uint32_t *soi, size_t n);

/* The cell list of ccp_cell_list for a device position array, with no
 * device allocation: xyz is n interleaved float triples in device memory,
 * csl and cps ccp's cell side and cells per side (sugrid(N, csl, cps),
 * sugrid.cuh:15-54); sci/soi (n entries) and beg/end (cps^3 entries) are
 * device arrays that receive ccp's layout, bit for bit what ccp_cell_list
 * gives for the same positions and geometry. Launches on `stream` (a
 * hipStream_t; NULL is the null stream) and waits for it. No counterpart
 * in ccp: it serves callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc). Its
 * only device memory beyond the arguments is a fixed array of 256 KiB per
 * GPU, so cps is at most 255 (CCP_ERR_ARG beyond); calls that overlap in
 * time on one GPU share that array and race. A bead outside the grid gives
 * CCP_ERR_STATE. Errors: ccp_last_error(NULL). */
int ccp_cell_list_noalloc(const float *xyz, uint32_t n, float csl, uint32_t cps, uint32_t *sci, uint32_t *soi,
                          uint32_t *beg, uint32_t *end, void *stream);

/* The noise of step `step` for every bead, in units of one standard
 * deviation (begin_iter, chrsim.cu:348-356), and the redraw count. */
int ccp_noise(ccp_sim *h, uint64_t step, float *z_xyz, uint32_t *attempts, size_t n);

/* Count non-bonded bead pairs (|i-j| > 1) closer than dmin
 * (chrsim::particle_overlaps, chrsim.cu:805-820, on the device). */
int ccp_count_overlaps(ccp_sim *h, float dmin, uint64_t *count);

/* Contact map: ordered pairs i != j with |r_i - r_j| < contact_cutoff,
 * binned by bead index (bin = i / contact_bin_size) into a
 * contact_bins^2 uint32 map accumulated with integer atomics. */
int ccp_contact_sample(ccp_sim *h);
int ccp_contact_reset(ccp_sim *h);
int ccp_contact_map(ccp_sim *h, uint32_t *map, size_t n_bins2, uint32_t *n_samples);

/* ccp's per-frame observables of the current positions, computed on the
 * device (chrana::calc_observables, chrana.cu:372-480): scalars[0] centre of
 * mass distance, [1] radius of gyration squared, [2] nematic order
 * parameter, [3] nucleus chromatin fraction; rcd (3 x n_bins, row-major)
 * the radial chromatin density of A beads, B beads and all beads (ccp uses
 * 64 bins). */
int ccp_observables(ccp_sim *h, double scalars[4], double *rcd, size_t n_bins);

/* ccp's mean distance d(s) and contact probability P(s) x 1000 (d < 2) of
 * chromosome `chromosome`, s = 1..n (calc_sd_cp, chrana.cu:13-45, with the
 * ranges of chrana.cu:93-97,484-491): FP32 sums as ccp. When N = 30386 the
 * chromosome is 0..5 of ccp's CHR_LIM, averaged with its homologue, and n
 * is half its length; otherwise chromosome 0 is beads [0, N) and n = N/2. */
int ccp_distance_contact(ccp_sim *h, uint32_t chromosome, float *sd, float *cp, size_t n);

/* ccp's per-frame contact map of the current positions (calc_cm,
 * chrana.cu:48-84): cms (cms+1)/2 floats of the packed lower triangle, with
 * cms = N/4 (N/8 when N = 30386, averaging the two chromosome copies). */
int ccp_frame_contact_map(ccp_sim *h, float *cm, size_t lcm);

/* ---- ccp's analysis (chrana, chrana.cuh:20-95; driver ccp-analyze.cu) ----
 * An analysis handle reads the current state of a simulation handle frame by
 * frame: ccp's per-frame observables (dcm, rg2, nop, ncf, radial density in
 * 64 bins) as time series, d(s)/P(s) of each chromosome and the packed
 * 4-bead contact map as per-simulation sums on the device. Per simulation
 * it writes analysis-XXX.dat; at the end analysis-fin.dat and
 * contact-map.bin, in ccp's layouts. */
typedef struct ccp_analysis ccp_analysis;

/* chrana::chrana (chrana.cu:88-120). Fails if the contact map of N would
 * exceed 2^31 pixels. The analysis keeps a pointer to h. */
int ccp_analysis_create(ccp_sim *h, ccp_analysis **out);
void ccp_analysis_destroy(ccp_analysis *a);

/* chrana::calc_observables (chrana.cu:372-512) on the positions and time
 * of the simulation handle. */
int ccp_analysis_add_frame(ccp_analysis *a);

/* chrana::add_trajectory_file (chrana.cu:136-152): read frames_per_file
 * .trr frames (chrdat::read_frame_bin, chrdat.cu:204-226) into the
 * simulation handle, adding each. */
int ccp_analysis_add_trajectory_file(ccp_analysis *a, const char *path);

/* calc_last_is_stat + save_last_is_stat + clear_is_var
 * (chrana.cu:154-283): statistics of the frames added since the last call,
 * written to path (analysis-XXX.dat layout). */
int ccp_analysis_end_simulation(ccp_analysis *a, const char *path);

/* calc_cs_final_stat + save_cs_final_stat + save_cm_avg_bin
 * (chrana.cu:286-369): combined statistics over the simulations ended so
 * far, written to fin_path (analysis-fin.dat) and cm_path (contact-map.bin:
 * lcm float64 averages of the packed lower triangle). */
int ccp_analysis_finish(ccp_analysis *a, const char *fin_path, const char *cm_path);

/* Error message of the last failed call on a. */
const char *ccp_analysis_last_error(const ccp_analysis *a);

/* The bead flags (CCP_HET, CCP_LAD) that ccp_generate_initial_condition of
 * a simulation with these parameters would assign (set_particle_types,
 * chrsim.cu:634-675, or the het/LAD generator), without creating the
 * simulation or touching a GPU. Errors: ccp_last_error(NULL). */
int ccp_sequence_pattern(const char *param_text, uint8_t *flags, size_t n);

/* ccp's time-series statistics of n float values (calc_stats for tdstat,
 * sostat.cu:141-183): out = {avg, var, sem, f_n_b, i_t, thermalised}. */
int ccp_series_stats(const float *v, size_t n, double out[6]);

/* Inverse-variance weighted combination of n (avg, sem) results
 * (sostat.cu:186-211): out = {avg, var, sem}. */
int ccp_weighted_stats(const double *avg, const double *sem, size_t n, double out[3]);

/* ---- batches of independent systems (no counterpart in ccp) ----
 * Many systems with the WCA/LJ set, each with its own bead count, nucleus
 * size, eps_het, eps_lam, het/LAD pattern and seed, stepped together: one
 * force launch and one order launch per Heun stage cover all of them, and
 * a hipGraph replays blocks of steps. Each system evolves exactly as a
 * ccp_sim created from ccp_batch_system_params(b, s) would alone. Every
 * contact_interval steps each system's contact map (contact_bins^2 uint32,
 * ceil(n / contact_bins) beads per bin) and radial density (radial_bins
 * equal-volume shells) are sampled. */
typedef struct ccp_batch ccp_batch;

typedef struct ccp_batch_system {
  uint32_t n_beads;
  uint32_t n_chains;      /* equal chains; 0 means 1 */
  float volume_fraction;  /* sets the nucleus radius by ccp's formula (chrdat.cu:57-58) */
  float eps_het;          /* het-het LJ well (kT) */
  float eps_lam;          /* lamina well of LAD beads (kT) */
  float het_fraction;
  float lad_fraction;
  uint64_t seed;          /* pattern, walk and noise key */
  uint32_t replica;       /* 0, or k > 0: same pattern as replica 0, its own walk and noise */
  uint32_t pad;
} ccp_batch_system;

/* common_params: parameter text shared by all systems (timestep,
 * temperature, kernel shape, cell_capacity, contact_cutoff,
 * contact_interval, contact_bins, radial_bins, het_block_mean,
 * lad_placement, graph_steps, ...). NULL on error (ccp_batch_last_error(NULL)). */
ccp_batch *ccp_batch_create(const char *common_params, const ccp_batch_system *systems, uint32_t n_sys);
void ccp_batch_destroy(ccp_batch *b);
/* The layout ccp_batch_create would give these systems, computed on the host
 * with no device call (the same parameter checks and errors): total beads,
 * total cells, bucket bytes at the initial cell capacity, and each system's
 * first cell in the shared cell range (n_sys + 1 entries, the last the
 * total). Any output may be NULL. Errors: ccp_batch_last_error(NULL). */
int ccp_batch_layout(const char *common_params, const ccp_batch_system *systems, uint32_t n_sys,
                     uint32_t *n_beads, uint64_t *n_cells, uint64_t *bucket_bytes, uint64_t *cell_offsets);
const char *ccp_batch_last_error(const ccp_batch *b);
/* offsets: n_sys + 1 first bead indices (system s holds beads [off[s], off[s+1])) */
int ccp_batch_info(const ccp_batch *b, uint32_t *n_sys, uint32_t *n_beads, uint32_t *n_cells, uint32_t *offsets);
/* the parameter text of system s alone (for a ccp_create of the same system) */
int ccp_batch_system_params(const ccp_batch *b, uint32_t s, char *buf, size_t len);
/* each system's initial condition as ccp_generate_initial_condition makes it */
int ccp_batch_generate_initial_conditions(ccp_batch *b);
int ccp_batch_run_steps(ccp_batch *b, uint64_t n);
int ccp_batch_synchronize(ccp_batch *b);
/* all beads: xyz (3 n floats) and CCP_HET/CCP_LAD flags (n bytes); either may be NULL */
int ccp_batch_get_state(ccp_batch *b, float *xyz, uint8_t *flags, size_t n);
int ccp_batch_set_state(ccp_batch *b, const float *xyz, const uint8_t *flags, size_t n);
int ccp_batch_get_step(const ccp_batch *b, uint64_t *step);
int ccp_batch_set_step(ccp_batch *b, uint64_t step);
/* n = n_sys * contact_bins^2 summed counts; n_samples = samples taken */
int ccp_batch_contact_maps(ccp_batch *b, uint32_t *maps, size_t n, uint32_t *n_samples);
/* n = n_sys * radial_bins mean bead volume fraction per shell */
int ccp_batch_radial_density(ccp_batch *b, double *rd, size_t n);
int ccp_batch_reset_samples(ccp_batch *b);
/* device time per step of n eager steps (HIP events), for the autotuner */
int ccp_batch_time_steps(ccp_batch *b, uint32_t n, float *ms_per_step);

/* Average device time per launch (ms) of each kernel of the step over
 * n_steps steps (at most 100000) that advance the simulation, measured with
 * HIP events recorded between the launches on the handle's stream (the
 * kernels run back to back; the host waits once at the end): out[0] force
 * stage 1, [1] order (scan + sort + gather) after stage 1, [2] force stage
 * 2, [3] order after stage 2, [4] whole step. Stands in for the timing of
 * ccp-profile.cu:55-87, which times whole run_simulation calls. */
int ccp_time_kernels(ccp_sim *h, uint32_t n_steps, float out[5]);

/* Capture every hipGraph block that ccp_run_steps replays (graph_steps
 * steps and each power of two below it) ahead of a timed run, so that the
 * run launches graphs only. No counterpart in ccp (run_simulation,
 * chrsim.cu:558-593, launches each kernel from the host). */
int ccp_capture_graphs(ccp_sim *h);

#ifdef __cplusplus
}
#endif

#endif
