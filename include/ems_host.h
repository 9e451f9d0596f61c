/*
 * ems_host.h -- the host side above the boundary, in C++: the parts of the
 * reference's compile/setup and run loop that produce what include/ems_gpu.h
 * consumes, and drive it.
 *
 * The reference keeps these on its (Rust) host: a Rust caller would use its
 * own `Coefficients::from_edges_rounded`, `CpmlBoundary::new` and
 * `run::drive` and bind only ems_gpu.h. They are restated here so the
 * engine can be driven without the Rust toolchain, with the same names and
 * argument meanings.
 *
 * Arrays use the engine layout (i*ny + j)*nz + k; cells use
 * (i*(ny-1) + j)*(nz-1) + k. Returns 0 or a negative EMS_E_* code.
 */
#ifndef EMS_HOST_H
#define EMS_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "ems_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* courant_limit (src/engine/coeffs.rs:249-255). */
double ems_courant_limit(const double *x, int nx, const double *y, int ny, const double *z,
                         int nz);

/* CellMaterials -> EdgeMaterials::from_cells (src/engine/edge_material.rs:376-417)
 * -> Coefficients::from_edges_rounded (src/engine/coeffs.rs:879-966).
 * cell_mat: material id per cell (NULL = all material 0); mat_table: per
 * material 12 doubles eps_r[3], mu_r[3], kappa[3], sigma_m[3] (Material,
 * src/engine/material.rs:22-31); pec: PEC mask per E edge, 3*N bytes (axis
 * major) or NULL; face_pec[6]: FaceCondition::Pec of x_min..z_max;
 * rounding 0 Exact, 1 OpenEms (CoeffRounding, coeffs.rs:706-761).
 * lanes: out, 12*N floats (ca 0-2, cb 3-5, da 6-8, db 9-11).
 * threads: worker threads for the x-planes (results do not depend on it). */
int ems_coefficients_from_cells(const double *x, int nx, const double *y, int ny,
                                const double *z, int nz, const uint16_t *cell_mat,
                                const double *mat_table, int n_mat, const uint8_t *pec,
                                const int face_pec[6], double dt, int rounding, int threads,
                                float *lanes);

/* EdgeMaterials::from_cells (src/engine/edge_material.rs:376-417): the
 * relative permittivity of every E edge and permeability of every H dual
 * edge (3*N doubles each, axis major; 0 where the edge is padding), from
 * the same cells and table as ems_coefficients_from_cells. */
int ems_edge_materials_from_cells(const double *x, int nx, const double *y, int ny,
                                  const double *z, int nz, const uint16_t *cell_mat,
                                  const double *mat_table, int n_mat, int threads,
                                  double *eps_r, double *mu_r);

/* A solid for the cell rasteriser (no reference counterpart: the reference
 * takes its solids from its external csx crate, whose rasteriser is not
 * part of it). kind 0: the box [p0, p3) x [p1, p4) x [p2, p5); 1: the
 * sphere of centre (p0, p1, p2) and radius p3; 2: the cylinder along
 * `axis` of radius p2 around (p0, p1) in the other two axes (in increasing
 * axis order), between p3 and p4 along it. Coordinates in metres. */
typedef struct {
    int32_t kind, axis;
    double p[6];
} ems_solid;

/* Sets cell_mat[cell] = id (cell (i*ncy + j)*ncz + k of the nx-1 by ny-1
 * by nz-1 cells) for every cell whose centre (the midpoints of its node
 * lines, 0.5 (l[i] + l[i+1])) lies inside the solid: box lo <= c < hi per
 * axis; sphere (cx - p0)^2 + (cy - p1)^2 + (cz - p2)^2 < r^2; cylinder the
 * same over its two radial axes, and lo <= c < hi along its axis. Only the
 * cells of the solid's bounding box are visited. Returns the number of
 * cells set, or a negative EMS_E_* code. Fills the cell materials that
 * ems_coefficients_from_cells and ems_edge_materials_from_cells average
 * onto the edges (a quarter of each edge's dual face per touching cell). */
int64_t ems_raster_cells(const double *x, int nx, const double *y, int ny, const double *z,
                         int nz, const ems_solid *solid, uint16_t id, uint16_t *cell_mat);

/* Lane::pool (src/engine/coeffs.rs:560-598): pools the nx*ny z-rows of a
 * dense lane if at most max(count/4, 1) are distinct. Returns 1 and fills
 * rows_out (<= count/4 rows, first-seen order) / index_out (nx*ny) when
 * pooled, 0 when the lane stays dense. */
int ems_pool_lane(const float *dense, int nx, int ny, int nz, float *rows_out,
                  uint32_t *index_out, uint64_t *nrows_out);

/* Cpml settings (src/fdtd.rs:177-296). */
typedef struct {
    int32_t cells;
    double grading_order, sigma_factor, kappa_max, alpha_max;
    int32_t convolution; /* 0 Constant, 1 Linear (default) */
} ems_cpml_params;

/* CpmlCoefficients::in_medium (src/engine/boundary/cpml.rs:350-388):
 * out = rho, sigma, kappa, alpha, b, a, direct. */
void ems_cpml_coefficients(const ems_cpml_params *p, double rho, double delta, double dt,
                           double n, double out[7]);

/* layer_index (cpml.rs:1054-1120): mean refractive index of the `cells`
 * layer on face (0 x_min .. 5 z_max), from dense lanes (12*N floats). */
double ems_cpml_layer_index(const double *x, int nx, const double *y, int ny, const double *z,
                            int nz, const float *lanes, double dt, int face, int cells);

/* CpmlBoundary::e_profile / h_profile (cpml.rs:1005-1018): the grading of
 * the layer on `face` at its E positions (e_side 1: the layer's lines, outer
 * face included, interface excluded) or H positions (cell midpoints), in
 * increasing index order; 7 doubles each as ems_cpml_coefficients. */
int ems_cpml_profile(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                     int face, const ems_cpml_params *p, double dt, double n_face, int e_side,
                     double *out, int *count);

/* Layer::new (cpml.rs:1122-1207): the four slabs of one face layer, in the
 * order E(a1), E(a2), H(a1), H(a2) with (a1, a2) the face's transverse
 * axes. b/a/k of slab q are written to coeff[q*3*np + {0,1,2}*np ..] with
 * np = nodes along the normal; the slab pointers point into coeff. */
int ems_cpml_layer(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                   int face, const ems_cpml_params *p, double dt, double n_face,
                   ems_cpml_slab out[4], float *coeff);

/* Waveform::gaussian(f0, fc).value(t) (src/waveform.rs:442-453). */
double ems_gaussian(double f0, double fc, double t);
/* Its duration 2T (waveform.rs:506-513). */
double ems_gaussian_duration(double fc);
/* out[n] = gaussian((n + offset)*dt) in f64, n < count. */
void ems_gaussian_values(double f0, double fc, double dt, double offset, uint64_t count,
                         double *out);
/* amp[n] = (gaussian((n + offset)*dt) as f32), n < count. */
void ems_gaussian_table(double f0, double fc, double dt, double offset, uint64_t count,
                        float *amp);

/* Waveform::openems_gaussian(f0, fc).value(t, dt) (src/waveform.rs:454-466):
 * tau = 3/(2 pi fc), T = 3 tau; with dt > 0 openEMS's sampled signal (sample
 * k = floor(t/dt + 1/4) non-zero for 1 <= k <= floor(2T/dt)). */
double ems_openems_gaussian(double f0, double fc, double t, double dt);
/* Its duration 2T (waveform.rs:506-513). */
double ems_openems_gaussian_duration(double fc);
/* The f32 amplitude table of an EdgeSource of `kind` (0 SoftE .. 3 HardH)
 * driven with openEMS timing (openems_signal / openems_lag,
 * src/sim/map.rs:278-305): amp[n] = the openEMS Gaussian at the kind's
 * sample time of step n minus the lag (dt/2 soft, dt hard) and
 * floor(delay/dt) steps, 0 before its first sample. */
int ems_openems_source_table(double f0, double fc, double delay, int kind, double dt,
                             uint64_t count, float *amp);

/* fold_parallel (src/engine/lumped.rs:275-282): adds conductance g [S] and
 * capacitance c [F] in parallel to an E edge with coefficients (ca, cb).
 * EMS_E_INVALID for cb = 0 (a PEC edge). */
int ems_fold_parallel(float ca, float cb, double dt, double g, double c, float *ca_out,
                      float *cb_out);

/* dft (src/post/dft.rs:218-222, RunningDft :75-190): X(f) = sum x(t_n)
 * e^{-j2 pi f t_n} dt with t_n = t0 + n dt; out[2k], out[2k+1] = Re, Im of
 * X(freqs[k]). */
void ems_dft(double t0, double dt, const double *x, uint64_t n, const double *freqs, int nf,
             double *out);
/* RunningDft's anchor phasors e^{-j 2 pi f_k t_n} at sample n from the
 * exact phase (src/post/dft.rs:116-122, phase_cycles :204-212), and its
 * per-sample rotations e^{-j 2 pi f_k dt} (:93-97); (re, im) per frequency. */
void ems_dft_phasors(const double *freqs, int nf, double t0, double dt, uint64_t n, double *out);
void ems_dft_steps(const double *freqs, int nf, double dt, double *out);

/* The conformal PEC's stability estimate (stability, src/engine/
 * conformal.rs:2858-2925; apply_s :2927-2995; faces :3012-3055): s =
 * sqrt(lambda_max(S)) / 2 by `iterations` power iterations, S = sqrt(cb)
 * (PK)^T dt (P^T L P)^-1 (PK) sqrt(cb) over the whole grid (the kernel's
 * faces with db != 0, merge groups with dt/L, blocks with the factor of
 * their M), from the fixed xorshift start vector plus 1 on the seed
 * edges; f64, one thread, the reference's loop and summation orders.
 * Per block row: the terms of its reduced current (`row_*`, block-local
 * rows). Returns s, or a negative code for invalid input. */
typedef struct {
    int32_t nodes[3];
    const float *cb[3], *db[3];   /* dense lanes, N each */
    int32_t n_groups;             /* merge groups with their own update */
    const double *g_dt_over_l;
    const uint32_t *g_term_start; /* n_groups + 1 */
    const int32_t *g_term_axis;
    const uint64_t *g_term_idx;
    const double *g_term_w;
    int32_t n_blocks;
    const uint32_t *blk_start;    /* n_blocks + 1 rows */
    const double *blk_dt;
    const uint32_t *row_term_start; /* rows + 1 */
    const int32_t *row_term_axis;
    const uint64_t *row_term_idx;
    const double *row_term_w;
    const uint32_t *env_first;    /* the factor of M by rows */
    const uint64_t *env_off;      /* rows + 1 */
    const double *env_val;
    int32_t n_seeds;
    const int32_t *seed_axis;
    const uint64_t *seed_idx;
    int32_t iterations;
} ems_stability_desc;
double ems_conformal_stability(const ems_stability_desc *d);

/* RunLimits + EndCriteria (src/engine/run.rs:22-139, src/fdtd.rs:428-605). */
typedef struct {
    uint64_t max_timesteps;
    int32_t has_criteria;
    double energy_ratio;
    uint64_t check_interval;
    int32_t has_excitation_end;
    double excitation_end;
    /* Settle (fdtd.rs:440-490); has_settle 0 = none */
    int32_t has_settle;
    double settle_frequency, settle_tolerance, settle_window, settle_max_extension;
    /* RunLimits::check_after_interval, extra_check, energy_measure
     * (run.rs:38-52): check the state reached after every interval (and
     * once more after extra_check steps) instead of the one before a step;
     * the measure 0 Invariant, 1 OpenEms (see ems_gpu_energy_in). A
     * stepper without its own state evaluation (the scripted one) ignores
     * the first two and the measure, as the reference's does. */
    int32_t check_after_interval, has_extra_check;
    uint64_t extra_check;
    int32_t energy_measure;
} ems_run_limits;

/* RunOutcome / StopReason (run.rs:141-214): reason 0 EnergyDecay,
 * 1 MaxTimesteps, 2 Aborted, 3 Diverged. */
typedef struct {
    int32_t reason;
    double ratio_db;
    uint64_t steps, timestep;
    double time, energy, max_energy, elapsed_s;
    int32_t has_criterion_steps;
    uint64_t criterion_steps;
} ems_run_outcome;

/* Engine::run_until (run.rs:216-273, drive 424-540) on a device engine.
 * Probe ids whose signals the settle test watches: probes[n_probes]. */
int ems_run_until(ems_gpu *h, const ems_run_limits *limits, const int *probes, int n_probes,
                  ems_run_outcome *out);

/* Engine::run_until for every model of a batch, in lockstep: each model
 * follows its own run::drive (run.rs:424-540) -- the same energy checks,
 * stop decisions and step counts as ems_run_until on it alone -- and
 * leaves the batch when it stops. out[q] per model. Settle, the checks
 * after the interval and openEMS's measure are not supported here. */
int ems_run_until_batch(ems_gpu_batch *b, int n_models, ems_gpu *const *models,
                        const ems_run_limits *limits, ems_run_outcome *out);

/* The same loop over a scripted stepper (run.rs:548-582): energy
 * w[min(n, nw-1)] at timestep n, dt = 1, and one probe sampled every step
 * with value signal[k] (k < ns; none if ns = 0). For testing the loop. */
int ems_drive_script(const double *w, uint64_t nw, const double *signal, uint64_t ns,
                     const ems_run_limits *limits, ems_run_outcome *out);

/* ---- setup producers (the reference's install steps on the host) ---- */

/* One pole of a dispersive medium on one edge (Dispersive::push,
 * src/engine/dispersive.rs:1870-1880): field 0 electric (E edge), 1
 * magnetic (H dual edge); kind 0 Lorentz (omega_p, omega_0, gamma), 1 Drude
 * (omega_p, gamma), 2 Debye (delta_eps, tau); every parameter finite and
 * >= 0. A pole with p[0] = 0 is null and ignored. */
typedef struct {
    int32_t field, axis, node[3], kind;
    double p[3];
} ems_pole_entry;

/* Dispersive::fill_cells (dispersive.rs:1896-1986): every edge of `field`
 * touching the cell box cells = {x0, x1, y0, y1, z0, z1} (half-open cell
 * ranges) gets each of the n_poles poles scaled by the share of its
 * averaging area (E) or length (H) inside the box (Pole::scaled: omega_p
 * by sqrt(share), a Debye delta_eps by the share), in the order axis x, y,
 * z, then the edge's nodes along its axis and the two others, then the
 * poles. Up to cap entries go to out (NULL to count); *count = all. */
int ems_dispersive_fill_cells(const double *x, int nx, const double *y, int ny, const double *z,
                              int nz, const int32_t cells[6], int field, const ems_pole_entry *poles,
                              int n_poles, ems_pole_entry *out, uint64_t cap, uint64_t *count);

/* The install plan of a dispersive medium (Dispersive::install,
 * dispersive.rs:2063-2319), i.e. what ems_gpu_add_dispersive takes: per
 * side (0 electric, 1 magnetic) the edges in plane order (x, field, axis,
 * y, z) with the sum of their poles' 1/z and the end of their poles in
 * cidx, and the distinct pole constants (a, 1/C, 1/z or k_b, L/4 dt or
 * L/2, kind 0 trapezoidal, 1 leapfrog; leapfrog for Lorentz and Drude
 * poles with scheme 1) at the edge's capacitance eps0 A/l (inductance
 * mu0 A/l). Edges whose cb (db) is 0 are skipped and counted. The caller
 * sizes every array for n entries (nodes 3 per edge). */
typedef struct {
    int32_t n_edges[2];
    int32_t *axis[2], *nodes[2];
    double *sum_inv_z[2];
    uint32_t *pole_end[2];
    uint64_t n_poles[2];
    uint32_t *cidx[2];
    int32_t n_consts;
    double *a, *inv_c, *inv_z, *half_l;
    int32_t *kind;
    int32_t skipped;
} ems_pole_plan;
int ems_dispersive_plan(const double *x, int nx, const double *y, int ny, const double *z, int nz,
                        double dt, const float *lanes, int scheme, const ems_pole_entry *entries,
                        int n, ems_pole_plan *out);

/* LumpedElement::install (src/engine/lumped.rs:829-1012): the edges'
 * shares of an element's R, L, C (lumped.rs:829-878), the parallel R and C
 * folded into their ca, cb (fold_parallel), and the RLC branch parameters.
 * The element's n_edges edges (nodes, 3 per edge) run along axis; with
 * columns > 0 the element's r, l, c (NaN: absent) are shared by `series`
 * edges in series times `columns` in parallel (a box), otherwise every
 * edge has its own edge_r, edge_l, edge_c (NaN: absent, the same set on
 * every edge). topology 0 Parallel (R and C folded, L a branch), 1 Series
 * (R, L, C in a branch when L or C is present); explicit_inductor: a
 * parallel inductor integrated explicitly (branch 2). */
typedef struct {
    int32_t axis, n_edges;
    const int32_t *nodes;
    int32_t series, columns;
    double r, l, c;
    const double *edge_r, *edge_l, *edge_c;
    int32_t topology, explicit_inductor;
} ems_lumped_element;
/* Out, n_edges each: cb_e (the edge's folded cb), g_src (the Thevenin
 * conductance, 0 in a series branch), rec_g, rec_c (the edge's parallel
 * G and C), and with a branch two_l_dt = 2L/dt, inv_c, z = 2L/dt + R +
 * dt/(2C); branch 0 none, 1 the trapezoidal branch, 2 the explicit
 * inductor; src_in_branch for a series branch. */
typedef struct {
    double *cb_e, *g_src, *rec_g, *rec_c, *two_l_dt, *inv_c, *z;
    int32_t branch, src_in_branch;
} ems_lumped_install;
/* ca, cb: the dense lanes of the element's axis (nodes[0]*nodes[1]*nodes[2]
 * floats each), folded in place. EMS_E_INVALID for an edge outside the
 * grid, a PEC edge (cb = 0) or a value that is not finite and > 0. */
int ems_lumped_plan(const int32_t nodes[3], double dt, const ems_lumped_element *el, float *ca,
                    float *cb, ems_lumped_install *out);

#ifdef __cplusplus
}
#endif
#endif
