/*
 * ems_gpu.h -- the drop-in boundary: an MI355X (gfx950) FDTD step engine
 * behind a C ABI.
 *
 * It replaces the reference engine's step scheduler and kernels, i.e.
 * everything under `Engine::advance` -> `pool::drive`
 * (src/engine/mod.rs:833-992, src/engine/pool.rs:878-940). The host keeps
 * mesh compilation (`Compiled::build_driving`, src/compile/mod.rs:926-936)
 * and hands over the compiled per-edge arrays once; the engine owns all
 * device memory and one HIP stream.
 *
 * Conventions (mirroring src/engine/mod.rs:1-62):
 *   - nodes n = [nx, ny, nz]; every field and coefficient array has one entry
 *     per node at (i*ny + j)*nz + k (z contiguous);
 *   - u[a] are the E-edge voltages, i[a] the H dual-edge currents; after n
 *     completed steps u holds t = n*dt and i holds (n - 1/2)*dt;
 *   - coefficient lanes: ca 0-2, cb 3-5, da 6-8, db 9-11 (coeffs.rs:385-391);
 *   - stepping is infallible in the reference; here every call returns 0 on
 *     success and a negative EMS_E_* code otherwise, with the message from
 *     ems_gpu_last_error. Device faults are sticky.
 *   - not thread-safe per handle; distinct handles may be used concurrently.
 *
 * Arithmetic: f32 fields and coefficients, f64 probes and energy, compiled
 * without FP contraction: fields, probes and energies are bit-identical to
 * the reference CPU engine given the same arrays.
 */
#ifndef EMS_GPU_H
#define EMS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMS_OK 0
#define EMS_E_INVALID (-1)     /* ems::Error::Invalid (src/error.rs) */
#define EMS_E_UNSUPPORTED (-2) /* ems::Error::Unsupported */
#define EMS_E_IO (-3)          /* ems::Error::Io */
#define EMS_E_DEVICE (-4)      /* HIP error (no reference counterpart) */

typedef struct ems_gpu ems_gpu;

/* Grid of the engine: EngineBuilder::build_from (src/engine/mod.rs:325-382)
 * after TimeStep::resolve. curl_order: 0 Paired (default), 1 OpenEms
 * (CurlOrder, src/engine/kernel.rs:81-107). */
typedef struct {
    int32_t nodes[3];
    double dt;
    int32_t curl_order;
} ems_grid_desc;

/* Engine::builder(...).build() -> one device engine. */
int ems_gpu_create(const ems_grid_desc *desc, ems_gpu **out);
void ems_gpu_destroy(ems_gpu *h);
const char *ems_gpu_last_error(const ems_gpu *h);

/* Coefficient lane `lane` (0..11): `nrows` z-rows of nz floats, and the
 * row of each (i, j) pair (index i*ny + j) among them, or NULL for a dense
 * lane (nrows = nx*ny). Mirrors Lane::{Dense, Pooled} and CoeffRows::row
 * (src/engine/coeffs.rs:393-599, src/engine/rows.rs:45-50). Lanes default to
 * all zero. */
int ems_gpu_set_lane(ems_gpu *h, int lane, const float *rows, uint64_t nrows,
                     const uint32_t *row_index);

/* One CPML slab (src/engine/boundary/cpml.rs:394-437): e_side 1 corrects u
 * from the i currents (after the E kernel), 0 corrects i from u (after the
 * H kernel). `comp` is the corrected component, `normal` the layer normal;
 * the node box lo..hi and b, a, k (= 1/kappa - 1 - shift) per depth index
 * node[normal] - lo[normal]. Slabs are applied in the order given, as the
 * reference applies its layers in registration order (cpml.rs:815-850). */
typedef struct {
    int32_t e_side, comp, normal;
    int32_t lo[3], hi[3];
    const float *b, *a, *k;
} ems_cpml_slab;
int ems_gpu_add_cpml(ems_gpu *h, const ems_cpml_slab *slabs, int n);

/* MurBoundary (src/engine/boundary/mur.rs:207-640): the first-order Mur
 * absorbing boundary on outer faces built as PEC. Edges (face edges with one
 * 1-D update each): axis, node[3], interior neighbour inner[3] one node
 * inward along the face normal, coefficient k = (v dt - d)/(v dt + d) as
 * f32, and a hold step (no update before it, MurBoundary::hold_faces;
 * NULL = none). Corners (edges on two Mur faces): the two interior
 * neighbours and coefficients. Before the E update u_b^n and u_1^n are
 * saved; after it (and after the CPML and dispersive corrections) every
 * edge takes u_1^n + k (u_1^{n+1} - u_b^n) (openems_order: (u_1^n - k u_b^n)
 * + k u_1^{n+1}), then every corner the average of its two updates
 * (openems_corners: u_1^n - k0 u_b^n of the first). Refused after an E
 * source, a voltage probe, a lumped element, a plane wave or cut-face
 * currents: the reference's slab mode (mur.rs:553-562), which this
 * placement follows, needs the boundary registered before them. */
typedef struct {
    int32_t n_edges;
    const int32_t *edge_axis, *edge_node, *edge_inner;
    const float *edge_k;
    const uint64_t *edge_hold;
    int32_t n_corners;
    const int32_t *corner_axis, *corner_node, *corner_inner0, *corner_inner1;
    const float *corner_k0, *corner_k1;
    const uint64_t *corner_hold;
    int32_t openems_order, openems_corners;
} ems_mur_desc;
int ems_gpu_add_mur(ems_gpu *h, const ems_mur_desc *d, int *id);

/* OpenEmsPml (src/engine/boundary/upml.rs:541-760): openEMS's PML_n layers,
 * a two-stage uniaxial PML through a flux F per layer edge. The caller has
 * rewritten the coefficients of every layer edge (ca = 0, cb = dt/((A/l)
 * (1 + a_p)); da, db likewise), so the H and E updates leave the flux
 * increment t in the field (OpenEmsPml::build, upml.rs:602-700). Per edge:
 * axis, node[3] and four f32 coefficients c = ((1 - a_p)/(1 + a_p),
 * (1 - a_q)/(1 + a_q), (1 + a_a)/((1 + a_q) m), (1 - a_a)/((1 + a_q) m)),
 * the E edges first, then the H dual edges. After the H update every E
 * edge takes its start value v from u and every H edge completes its
 * update; after the E update the roles swap (Parts::post_h_slab /
 * post_e_slab, upml.rs:442-470). The update (Run::apply, upml.rs:404-426),
 * every f32 operation flushed:
 *   F' = c0 F + t;  x = (c1 v - c3 F) + c2 F';  F = F', v = x, field = x.
 * It runs first among the slab corrections, as the reference registers it
 * (compile/mod.rs:1224-1236), so it is refused after dispersive media, Mur
 * boundaries, plane waves or cut-face currents; it never shares an engine
 * with CPML slabs (refused either way round). An edge listed twice and a
 * second layer operator are refused. The steps run unfused (the H update
 * has to be complete before the E update reads it). */
typedef struct {
    int32_t n_e;
    const int32_t *e_axis, *e_node; /* node: 3 per edge */
    const float *e_coef;            /* 4 per edge */
    int32_t n_h;
    const int32_t *h_axis, *h_node;
    const float *h_coef;
} ems_upml_desc;
int ems_gpu_add_upml(ems_gpu *h, const ems_upml_desc *d, int *id);

/* The layer state in the given edge order: side 0 the E edges, 1 the H
 * dual edges; flux and value each hold the side's count of floats. */
int ems_gpu_upml_state(ems_gpu *h, int id, int side, float *flux, float *value);

/* EdgeSource (src/engine/excite.rs:10-173): kind 0 SoftE, 1 HardE,
 * 2 SoftH, 3 HardH; edges (axis, node[3], weight f32). The host tabulates
 * the amplitude `a(t) as f32` at each step's sample time (SoftE (n+1/2)dt,
 * HardE (n+1)dt, SoftH n*dt, HardH (n+1/2)dt); steps at or past n_amp have
 * amplitude 0. *id receives the operator id. */
int ems_gpu_add_source(ems_gpu *h, int kind, int n_edges, const int32_t *axis,
                       const int32_t *nodes, const float *w, const float *amp, uint64_t n_amp,
                       int *id);

/* Probe (src/engine/probe.rs:240-513): quantity 0 voltage (u after the E
 * update, t = (n+1)dt), 1 current (i after the H update, t = (n+1/2)dt);
 * value = sum w*f64(x) in entry order, every `interval` steps. With
 * EMS_PROBE_F32_SUM or-ed into the quantity, Probe::f32_sum (probe.rs:398-404,
 * the sum of openEMS's current probes): f64(sum of f32(w)*x in f32 from 0). */
#define EMS_PROBE_F32_SUM 4
int ems_gpu_add_probe(ems_gpu *h, int quantity, int n_entries, const int32_t *axis,
                      const int32_t *nodes, const double *w, uint64_t interval, int *id);

/* An installed lumped element (LumpedElement, src/engine/lumped.rs:367-406:
 * Installed + Branch + CurrentRecord after `install` :880-1012, which the
 * host performs, folding R and C into the ca/cb lanes with fold_parallel
 * :275-282 before calling this), optionally a LumpedPort
 * (src/engine/port.rs:94-297). Edges run along `axis` from nodes[3*q].
 *   cb, g_src     per edge: the final cb (as f64) and the conductance in
 *                 series with the source (0 where the source sits in the
 *                 branch);
 *   series        M, the edges in series;
 *   n_terms       source terms, weights term_weight[t*n_edges + q]
 *                 (1/M for a box source); n_drive branch-drive terms
 *                 likewise (branch_drive_source); values by ems_gpu_lumped_table;
 *   branch        0 none, 1 the trapezoidal series R+L+C branch, 2 the
 *                 explicit inductor; two_l_dt = 2L/dt, inv_c = 1/C (0 = no
 *                 capacitor), z = 2L/dt + R + dt/(2C) per edge;
 *   rec_g, rec_c  per-edge g and C of the terminal-current record (NULL =
 *                 no record; `current` :1042-1047);
 *   port_columns  > 0: a LumpedPort with N columns: every step records
 *                 V = sum(u)/N and I = -record (port.rs:281-297).
 * pre_e (u^n of each edge) and post_e (:1123-1249) run on the device every
 * step from the next one on, in registration order with the other
 * post_e operators. The openEMS lumped order and cylindrical edge copies
 * are not supported. */
typedef struct {
    int32_t axis, n_edges;
    const int32_t *nodes;
    const double *cb, *g_src;
    int32_t series, src_in_branch;
    int32_t n_terms;
    const double *term_weight;
    int32_t n_drive;
    const double *drive_weight;
    int32_t branch;
    const double *two_l_dt, *inv_c, *z;
    const double *rec_g, *rec_c;
    int32_t port_columns;
} ems_lumped_desc;
int ems_gpu_add_lumped(ems_gpu *h, const ems_lumped_desc *d, int *id);

/* Source::value (lumped.rs:241-247) of term `term` (drive terms numbered
 * after the source terms) at h_time = (n + 1/2)dt, tabulated for steps
 * n < count (absolute step index); 0 beyond. Replaces the previous table. */
int ems_gpu_lumped_table(ems_gpu *h, int id, int term, const double *values, uint64_t count);

/* Recorded series of a lumped element: which 0 the terminal current
 * (LumpedElement::current), 1 the port voltage, 2 the port current
 * (LumpedPort::voltage / current). *count = samples recorded; the first
 * was taken at the step *first_step (NULL allowed). */
int ems_gpu_lumped_values(ems_gpu *h, int id, int which, double *out, uint64_t cap,
                          uint64_t *count, uint64_t *first_step);

/* A running DFT on the device (RunningDft, src/post/dft.rs:75-190) of a
 * lumped element's record series `which` (as ems_gpu_lumped_values) at nf
 * frequencies, with sample n at t0 + n dt counted from the element's first
 * record: X(f) = sum x(t_n) e^{-j 2 pi f t_n} dt, the phasor advanced by
 * one complex product per sample and re-anchored from the exact phase
 * every 1024 samples (the anchors are ems_dft_phasors'), so the result is
 * ems_dft's on the same series, bit for bit, and the series never leaves
 * the device. ems_gpu_dft_update queues the accumulation of the samples
 * recorded since the last update (any number of updates give the same
 * result); ems_gpu_dft_result waits for it and writes 2*nf doubles (Re,
 * Im per frequency) and the samples accumulated. No reference counterpart
 * for the device placement: the reference runs LumpedPort::analyze_with
 * (src/port/mod.rs:501-544) on the host over the recorded series. */
int ems_gpu_add_dft(ems_gpu *h, int lumped_id, int which, const double *freqs, int nf, double t0,
                    double dt, int *dft_id);
int ems_gpu_dft_update(ems_gpu *h, int dft_id);
int ems_gpu_dft_result(ems_gpu *h, int dft_id, double *out, uint64_t *samples);

/* Branch state j^n, q^n and the last j-bar per edge (branch_energy,
 * branch_current, lumped.rs:1033-1087). */
int ems_gpu_lumped_branch(ems_gpu *h, int id, double *j, double *q, double *j_avg);

/* A dispersive medium (Dispersive, src/engine/dispersive.rs:1790-2794),
 * after the host's `install` (:2063-2319): the distinct pole constants and
 * the two sides (electric poles on E edges, magnetic poles on dual edges).
 * A side lists its edges in (x plane, axis, j, k) order: axis, start node,
 * sum of 1/z over its trapezoidal poles, and the end of its poles in the
 * pole arrays (CSR); each pole an index into the constants:
 *   a      2L/dt (trapezoidal) or k_a (leapfrog),
 *   inv_c  1/C_p,
 *   inv_z  1/z (trapezoidal) or k_b (leapfrog),
 *   half_l L/2 (energy),
 *   kind   0 trapezoidal, 1 leapfrog (Branch, dispersive.rs:1641-1663).
 * The device keeps j, q (f64) per pole and the saved field per edge, and
 * corrects the E side after the E update and the H side after the H
 * update (Side::correct / step_edge, :1183-1343, :1392-1470), reading
 * cb / db from the lanes every step. A medium with magnetic poles runs
 * the unfused schedule. Media must be registered after every CPML slab
 * (the order the reference recommends, dispersive.rs module docs). */
typedef struct {
    int32_t n;
    const double *a, *inv_c, *inv_z, *half_l;
    const int32_t *kind;
} ems_pole_consts;
typedef struct {
    int32_t n_edges;
    const int32_t *axis, *nodes;
    const double *sum_inv_z;
    const uint32_t *pole_end;
    uint64_t n_poles;
    const uint32_t *cidx;
} ems_pole_side;
int ems_gpu_add_dispersive(ems_gpu *h, const ems_pole_side *electric,
                           const ems_pole_side *magnetic, const ems_pole_consts *consts,
                           int *id);
/* Branch state of one side (field 0 electric, 1 magnetic) in pole order:
 * j^n, q^n per pole (Dispersive::electric_energy / magnetic_energy). */
int ems_gpu_dispersive_state(ems_gpu *h, int id, int field, double *j, double *q);

/* A total-field / scattered-field plane wave (PlaneWave,
 * src/engine/planewave.rs:440-1561) in a homogeneous background, after the
 * host's PlaneWave::new (:684-1049): the auxiliary 1-D grid's coefficients
 * (ca, cb on n_aux E nodes; da, db on n_aux - 1 H nodes) and the two
 * correction lists (Correction, :424-431): E edges on the box surface
 * corrected after the E update from the auxiliary H, H edges just outside
 * the box corrected after the H update from the auxiliary E, each
 *   field[axis][node] += f32(coef * sum_q w[4q'+q] * aux[m + q]).
 * Corrections of one edge apply in list order. Every step the device then
 * advances the auxiliary grid (:1525-1528, :1553-1559), e[0] from the table
 * of ems_gpu_planewave_table. A plane wave runs the unfused schedule (its
 * post_h hook sits between the H and E updates), at its place among the
 * other serial operators. */
typedef struct {
    int32_t n;
    const int32_t *axis, *nodes;
    const double *coef;
    const uint32_t *m;
    const double *w;
} ems_pw_corrections;
typedef struct {
    int32_t n_aux;
    const double *ca, *cb, *da, *db;
    ems_pw_corrections e_corr, h_corr;
} ems_pw_desc;
int ems_gpu_add_planewave(ems_gpu *h, const ems_pw_desc *d, int *id);
/* values[n] = gain * a((n + 1)dt + lead), the auxiliary source after the E
 * update of step n, for n < count (0 beyond). Replaces the table. */
int ems_gpu_planewave_table(ems_gpu *h, int id, const double *values, uint64_t count);
/* The auxiliary grid's state: e (n_aux values), h (n_aux - 1) (for
 * PlaneWave::incident_e / incident_h, :1446-1470). */
int ems_gpu_planewave_aux(ems_gpu *h, int id, double *e, double *hh);

/* Reduced currents of small cut faces: the MergedFaces operator of the
 * conformal PEC (src/engine/conformal.rs:2313-2700; post_h :2655-2671,
 * post_e :2673-2682, energy :2684-2698), built on the host (merge groups,
 * interpolation hosts and deviation currents, their blocks' envelope
 * Cholesky factors). After every H update: the circulations of u per
 * reduced current (terms ts[g]..ts[g+1]), the update of each merge group
 * g < n_scalar (g' = flush(f32(da g - db c))) and of each block (rows
 * bs[b]..bs[b+1] in elimination order: x = B g - dt c, L L^T x' = x,
 * g' = flush(f32(x'))), then every driven face's current sum w g' as f32.
 * After every E update the circulations again (the energy look-ahead).
 * Columns of B, M and the factor are block-local positions; env_off has
 * one entry per row plus one. Runs the unfused schedule: the circulations
 * and the update (they read u and the operator's own state only) run on a
 * second stream of the engine beside the H update, the driven faces are
 * written at the operator's place among the serial operators of the mid
 * gap; the engine energy adds 1/2 g^T M g^(n+3/2). */
typedef struct {
    int32_t n_groups, n_scalar, n_blocks, n_faces;
    double dt;
    const double *da, *db, *l;  /* per group (da, db used for g < n_scalar) */
    const uint32_t *term_start; /* n_groups + 1 */
    const int32_t *term_axis;
    const uint64_t *term_idx;   /* flat node index of the E edge */
    const double *term_w;
    const uint32_t *blk_start;  /* n_blocks + 1, into the rows */
    const uint32_t *members;    /* per row: its reduced current */
    const uint32_t *env_first;  /* per row: first column of its envelope */
    const uint64_t *env_off;    /* rows + 1 */
    const double *env_val;
    const uint32_t *brow_start; /* rows + 1 */
    const uint32_t *bcol;
    const double *bval;
    const uint32_t *mrow_start; /* rows + 1 */
    const uint32_t *mcol;
    const double *mval;
    /* the factor by columns: for row j the rows k > j whose envelope holds
     * j, in decreasing k, and L[k][j] (the backward substitution taken
     * column by column; the same operations in the same order) */
    const uint32_t *col_start;  /* rows + 1 */
    const uint32_t *col_row;
    const double *col_val;
    /* the CPML's H-side correction of driven faces inside layers
     * (MergedFaces::add_pml, conformal.rs:2582-2638): per term
     * T = sign (u_axis[idx0] - u_axis[idx1]), psi <- b psi + a T (kept after
     * the H update, not after E), contribution k T + psi, added to the
     * circulation of each reduced current driving the face, w times, in
     * term order (per current: pml_start, pml_term, pml_w) */
    int32_t n_pml;
    const int32_t *pml_axis;
    const uint64_t *pml_idx0, *pml_idx1;
    const double *pml_coef;     /* 4 per term: sign, b, a, k */
    const uint32_t *pml_start;  /* n_groups + 1 */
    const uint32_t *pml_term;
    const double *pml_w;
    const int32_t *face_axis;
    const uint64_t *face_idx;   /* flat node index of the driven H edge */
    const uint32_t *face_start; /* n_faces + 1 */
    const uint32_t *face_group;
    const double *face_w;
} ems_merged_desc;
int ems_gpu_add_merged(ems_gpu *h, const ems_merged_desc *d, int *id);
/* ems_gpu_energy of many engines at once (no reference counterpart; the
 * batch run loop's energy checks), each out[q] the value ems_gpu_energy
 * gives, for active[q] != 0 (all when active is NULL). Two or more active
 * engines of one curl order and row-width class (nz up to 32, 64, 128,
 * 256, more) without operator energies (dispersive media, cut-face currents,
 * lumped branches) are evaluated together: each engine's queued work is
 * waited for, then one launch per energy kernel over all of them on the
 * first one's stream. Otherwise every evaluation is queued on its engine's
 * stream before any is waited for. */
int ems_gpu_energy_many(ems_gpu *const *models, int n, const uint8_t *active, double *out);

/* The reduced currents g (n_groups values) and the circulations of the
 * last H or E update (either may be NULL). */
int ems_gpu_merged_state(ems_gpu *h, int id, double *cur, double *circ);

/* Batches (no reference counterpart; for geometry sweeps of many small
 * models): engines stepped together, one fused launch over the tiles of
 * every active model and one launch for all their post-E gaps per step,
 * captured in graphs on the batch's own stream. Every model computes what
 * its own engine would, bit for bit. The models must all run the fused
 * schedule, with one curl order, and be in step with each other (same
 * step parity); a handle listed twice is refused. The batch holds the
 * handles; it does not own them. ems_run_until_batch with an odd check
 * interval can stop models at checks an odd number of steps apart: a later
 * ems_gpu_batch_run over all of them is then refused with EMS_E_INVALID
 * ("models of a batch step in lockstep"); run such models in separate
 * batches, or step the odd ones once on their own. */
typedef struct ems_gpu_batch ems_gpu_batch;
int ems_gpu_batch_create(ems_gpu *const *models, int n, ems_gpu_batch **out);
void ems_gpu_batch_destroy(ems_gpu_batch *b);
const char *ems_gpu_batch_last_error(const ems_gpu_batch *b);
/* Kernel for every model of the batch: mode 1 the tiled fused kernel, 2
 * the same without the next-plane prefetch, 4 packed short rows (every
 * model nz <= 64), 5 packed without the prefetch; tile height (3, 7, 15;
 * tiled modes only), x chunks per model (0 automatic). */
int ems_gpu_batch_set_schedule(ems_gpu_batch *b, int mode, int tile_rows, int x_chunks);
/* n steps of the models with active[q] != 0 (NULL: all); returns after
 * they are done. */
int ems_gpu_batch_run(ems_gpu_batch *b, const uint8_t *active, uint64_t n);
/* Whether ems_gpu_batch_run replays captured HIP graphs (on, the default)
 * or queues the same launches one by one (no reference counterpart: for
 * tools that trace launches but not graph replays). Same results. */
int ems_gpu_batch_set_graphs(ems_gpu_batch *b, int on);

/* Engine::run (src/engine/mod.rs:825-828). */
int ems_gpu_run(ems_gpu *h, uint64_t n);
/* Engine::step_with_energy (src/engine/mod.rs:808-823): one step, *w = the
 * energy invariant W^n of the state before it. */
int ems_gpu_step_with_energy(ems_gpu *h, double *w);
/* Engine::energy (src/engine/mod.rs:722-726): read-only W^n. */
int ems_gpu_energy(ems_gpu *h, double *w);
/* Engine::energy_in (src/engine/mod.rs:747-752) and Stepper::energy_now
 * (src/engine/run.rs:296-299): measure 0 EnergyMeasure::Invariant (=
 * ems_gpu_energy); 1 EnergyMeasure::OpenEms (Engine::energy_openems,
 * mod.rs:727-741): eps0 sum u^2 + mu0 sum w i'^2 over every edge with
 * i' the read-only look-ahead, the dual edges of the outer planes normal
 * to their own axis weighted 4 and the E of the three upper faces left out
 * (kernel.rs:364-423, 132-173, 218-244), the PML layers' corrections of
 * their H edges (upml.rs:505-537), plus the H sources' look-ahead share
 * (dump.rs:1735-1782). Cartesian grids (the only ones held). */
int ems_gpu_energy_in(ems_gpu *h, int measure, double *w);

/* No counterpart in the reference: ems_gpu_energy for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same value, bit for bit, but it calls no
 * hipMalloc, hipFree or hipHostMalloc: its scratch (row sums, plane chains,
 * plane values) lives in one static device array per GPU of
 * EMS_ENERGY_NOALLOC_DOUBLES doubles (16 MiB), so grids with
 * nx*(3*ny - 2) + 13*nx + 1 above that are refused with EMS_E_INVALID, and
 * engines holding dispersive media, reduced currents or lumped branches
 * (whose energies are read back into host buffers) with EMS_E_UNSUPPORTED.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several host threads) race; each call waits for its own result. */
#define EMS_ENERGY_NOALLOC_DOUBLES (2u * 1024u * 1024u)
int ems_gpu_energy_noalloc(ems_gpu *h, double *w);
uint64_t ems_gpu_timestep(const ems_gpu *h);

/* Engine::u_component / i_component / fields_mut (src/engine/mod.rs:659-690):
 * field 0 = u, 1 = i; nx*ny*nz floats. */
int ems_gpu_read_component(ems_gpu *h, int field, int axis, float *out);
int ems_gpu_write_component(ems_gpu *h, int field, int axis, const float *in);

/* Probe::values (src/engine/probe.rs:354-357): copies up to cap samples;
 * *count receives the number recorded so far. */
int ems_gpu_probe_values(ems_gpu *h, int id, double *out, uint64_t cap, uint64_t *count);

/* Probe::times (src/engine/probe.rs:339-342, SampleTimes::get :93-104):
 * sample k at ((step_k as f64) + offset)*dt, offset 1 for voltages and 1/2
 * for currents. */
int ems_gpu_probe_times(ems_gpu *h, int id, double *out, uint64_t cap, uint64_t *count);

/* Engine::dt (src/engine/mod.rs). */
double ems_gpu_dt(const ems_gpu *h);

/* Step schedule (no reference counterpart; the reference picks fused steps
 * by itself, pool.rs:131-164): fused 1 runs H, CPML-H, E and CPML-E of a
 * step in one pass over HBM with double-buffered fields (bit-identical to
 * the unfused schedule; 2 is the same without the next-plane prefetch; 3
 * the full-row variant: one wave per row, nz <= 576, tile_rows 3 or 7;
 * 4 packed short rows: 64 / nz whole k-rows per wave, nz <= 64, else the
 * tiled kernel; 5 the same without the prefetch); tile_rows 3, 7 or 15
 * (0 keeps the current);
 * x_chunks 0 = automatic. Steps that need the serial gap between the H and
 * E updates (H sources) always run unfused. */
int ems_gpu_set_schedule(ems_gpu *h, int fused, int tile_rows, int x_chunks);

/* CPML slabs of the tiled fused step (no reference counterpart): with
 * on = 1 (default 0) the slabs normal to z that follow every other slab
 * of their side in registration order are corrected outside the fused
 * kernel, H-side by increments formed from u^n before it, E-side after it
 * (the same values, bit for bit); 0 keeps every slab inside. */
int ems_gpu_set_cpml_split(ems_gpu *h, int on);
/* Tile classes of the tiled fused step (no reference counterpart): with
 * on != 0 the tiles that a CPML slab reaches run in one launch and the
 * others in a second one, with a kernel instantiation without the slab
 * code (fewer registers, more waves per SIMD). Same results either way. */
int ems_gpu_set_tile_classes(ems_gpu *h, int on);
/* The unfused H and E kernels (no reference counterpart). `which` bit 0
 * (H) and bit 1 (E): the side's threads march a chunk of x-planes (8 by
 * default; the schedule's x chunks where set) with the next plane's inputs
 * loaded before the present plane is computed, else two planes without
 * that. Bit 2 (H) and bit 3 (E), with the side's bit 0 or 1: the side's
 * CPML slabs, but a suffix of z-normal ones, are corrected inside the
 * update instead of in launches of their own. Default 11 (H and E load
 * ahead, E corrects its slabs inside: measured fastest). Same results
 * either way; 0 to 15, else EMS_E_INVALID. */
int ems_gpu_set_unfused_prefetch(ems_gpu *h, int which);

/* Tile order of the fused kernels (no reference counterpart): groups of
 * `group` consecutive tile rows go round-robin to the 8 XCDs (their L2s),
 * 0 = plain block order. Default 2. */
int ems_gpu_set_xcd_group(ems_gpu *h, int group);

/* Waits for all queued device work. */
int ems_gpu_synchronize(ems_gpu *h);

/* Measurement hook (no reference counterpart): runs `steps` steps with HIP
 * events around the main stencil kernels on the engine's stream and
 * returns their mean durations in ms: out[0] the fused step kernel (or the
 * H update when unfused), out[1] the E update (0 when fused), out[2] the
 * whole step. */
int ems_gpu_time_kernels(ems_gpu *h, uint64_t steps, double out[3]);

/* Algorithmic HBM bytes of one launch of each stencil kernel, from the lanes
 * and slabs the engine holds (no reference counterpart): out[0] the fused
 * step kernel (or the H update when unfused), out[1] the E update (0 when
 * fused), out[2] all CPML slab launches of one unfused step. */
int ems_gpu_traffic_model(ems_gpu *h, double out[3]);

/* Measurement hook (no reference counterpart): device-to-device copy of
 * `bytes` (read + write), `reps` times; *gbs = bytes moved / time. */
int ems_gpu_copy_bandwidth(ems_gpu *h, uint64_t bytes, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif
