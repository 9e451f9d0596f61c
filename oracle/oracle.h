/*
 * oracle.h -- CPU restatement of the ems FDTD engine step (TEST INFRASTRUCTURE).
 *
 * This library is the parity oracle and the CPU baseline. It is NOT part of
 * the product: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, always as the checker or the timed CPU
 * reference, never as the thing measured or shipped.
 *
 * Every function restates the reference algorithm with the same f32/f64
 * split and the same expression order (compiled with -ffp-contract=off, no
 * fast-math), citing the reference file:line it follows. The engine state
 * is built from the same arrays the GPU boundary (include/ems_gpu.h)
 * receives, so both consume identical compiled operators.
 *
 * Pinning (see the test files named in each comment):
 *   - kernel (A1/A2): closed forms -- PEC cavity modes, Yee dispersion,
 *     lossy decay rate, invariant drift (tests/test_oracle_physics.py);
 *   - energy order (A3): determinism over thread counts + the in-step value
 *     equals the read-only value bit for bit (tests/test_oracle_engine.py);
 *   - coefficients / Courant (A5): the reference's unit-test numbers
 *     (tests/test_oracle_setup.py);
 *   - CPML coefficients (A6): exact grading/convolution ratios of the
 *     reference unit tests; CPML apply (A7): reflection levels;
 *   - run loop (A12): the scripted energy sequences of run.rs tests.
 */
#ifndef EMS_ORACLE_H
#define EMS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_engine orc_engine;

/* Engine over nx*ny*nz nodes, layout (i*ny+j)*nz+k.  curl_order: 0 Paired,
 * 1 OpenEms (kernel.rs:81-107).  threads >= 1: x-slab worker threads
 * (fields.rs:38-59). */
orc_engine *orc_create(int nx, int ny, int nz, double dt, int curl_order, int threads);
void orc_destroy(orc_engine *e);
const char *orc_last_error(orc_engine *e);

/* Coefficient lane (ca 0-2, cb 3-5, da 6-8, db 9-11): `nrows` rows of nz
 * values and a row index per (i,j) (NULL = dense, nrows = nx*ny). */
int orc_set_lane(orc_engine *e, int lane, const float *rows, uint64_t nrows,
                 const uint32_t *index);

/* One CPML slab (cpml.rs:394-420): side 1 = E (corrects u from i), 0 = H.
 * Slabs are applied in the order added, E slabs after the E kernel, H
 * slabs after the H kernel (cpml.rs:815-850). b/a/k per depth index
 * node[normal]-lo[normal]. */
int orc_add_cpml(orc_engine *e, int e_side, int comp, int normal, const int lo[3],
                 const int hi[3], const float *b, const float *a, const float *k);

/* Edge source (excite.rs:10-115): kind 0 SoftE, 1 HardE, 2 SoftH, 3 HardH.
 * amp[n] = (a(t_n) as f32) for the step n's sample time; beyond n_amp the
 * amplitude is 0. Returns operator id. */
int orc_add_source(orc_engine *e, int kind, int n_edges, const int *axis,
                   const int *nodes, const float *w, const float *amp, uint64_t n_amp);

/* Probe (probe.rs:364-414): quantity 0 voltage (u, after E), 1 current (i,
 * after H). Returns operator id. */
int orc_add_probe(orc_engine *e, int quantity, int n_entries, const int *axis,
                  const int *nodes, const double *w, uint64_t interval);

/* Lumped element (lumped.rs:367-406 after install, :1106-1249) over edges
 * along `axis` starting at nodes[3*q]: final cb and source conductance per
 * edge, M = series, source (n_terms) and branch drive (n_drive) weight rows
 * of n_edges each, branch 0 none / 1 trapezoidal series R+L+C / 2 explicit
 * inductor with 2L/dt, 1/C, Z per edge, terminal-current record g, c per
 * edge (rec_g NULL: none), port_columns > 0: a LumpedPort (port.rs:94-297)
 * recording V = sum(u)/columns and I = -record. Returns the operator id. */
typedef struct {
    int axis, n_edges;
    const int *nodes;
    const double *cb, *g_src;
    int series, src_in_branch;
    int n_terms;
    const double *term_weight;
    int n_drive;
    const double *drive_weight;
    int branch;
    const double *two_l_dt, *inv_c, *z;
    const double *rec_g, *rec_c;
    int port_columns;
} orc_lumped_desc;
int orc_add_lumped(orc_engine *e, const orc_lumped_desc *d);
/* values[n] = Source::value((n + 1/2)dt) of term `term` (drive terms after
 * the source terms) for steps n < count, 0 beyond. */
int orc_lumped_table(orc_engine *e, int id, int term, const double *values, uint64_t count);
/* which 0: terminal current record, 1: port V, 2: port I. Returns the count. */
uint64_t orc_lumped_values(orc_engine *e, int id, int which, double *out, uint64_t cap);
int orc_lumped_branch(orc_engine *e, int id, double *j, double *q, double *javg);

/* Dispersive medium (dispersive.rs) after the host's install: pole
 * constants and the electric / magnetic sides, laid out as ems_pole_consts
 * and ems_pole_side of include/ems_gpu.h. The corrections run as slab
 * hooks after the CPML ones. Returns the operator id. */
typedef struct {
    int n;
    const double *a, *inv_c, *inv_z, *half_l;
    const int *kind;
} orc_pole_consts;
typedef struct {
    int n_edges;
    const int *axis, *nodes;
    const double *sum_inv_z;
    const uint32_t *pole_end;
    uint64_t n_poles;
    const uint32_t *cidx;
} orc_pole_side;
int orc_add_dispersive(orc_engine *e, const orc_pole_side *electric,
                       const orc_pole_side *magnetic, const orc_pole_consts *consts);
int orc_dispersive_state(orc_engine *e, int id, int field, double *j, double *q);

/* Plane wave (planewave.rs) after the host's PlaneWave::new: laid out as
 * ems_pw_desc of include/ems_gpu.h; post_h / post_e hooks in the gaps.
 * The table gives the auxiliary source e[0] after the E update of step n. */
typedef struct {
    int n;
    const int *axis, *nodes;
    const double *coef;
    const uint32_t *m;
    const double *w;
} orc_pw_corrections;
typedef struct {
    int n_aux;
    const double *ca, *cb, *da, *db;
    orc_pw_corrections e_corr, h_corr;
} orc_pw_desc;
int orc_add_planewave(orc_engine *e, const orc_pw_desc *d);
int orc_planewave_table(orc_engine *e, int id, const double *values, uint64_t count);
int orc_planewave_aux(orc_engine *e, int id, double *ev, double *hv);

/* Reduced currents of small cut faces (MergedFaces, conformal.rs:2313-2700),
 * tables built on the host: per reduced current its circulation terms;
 * merge groups (g < n_scalar) with their own da, db; blocks of rows in
 * elimination order with the envelope Cholesky factor of A, and the rows of
 * B and M (block-local columns); the driven faces and their weights. */
typedef struct {
    int32_t n_groups, n_scalar, n_blocks, n_faces;
    double dt;
    const double *da, *db, *l;
    const uint32_t *term_start;
    const int32_t *term_axis;
    const uint64_t *term_idx;
    const double *term_w;
    const uint32_t *blk_start, *members, *env_first;
    const uint64_t *env_off;
    const double *env_val;
    const uint32_t *brow_start, *bcol;
    const double *bval;
    const uint32_t *mrow_start, *mcol;
    const double *mval;
    const uint32_t *col_start, *col_row; /* unused here (a layout of the device's) */
    const double *col_val;
    int32_t n_pml; /* CPML terms of driven faces in layers, 4 coefficients each */
    const int32_t *pml_axis;
    const uint64_t *pml_idx0, *pml_idx1;
    const double *pml_coef;
    const uint32_t *pml_start, *pml_term; /* the terms' targets by reduced current */
    const double *pml_w;
    const int32_t *face_axis;
    const uint64_t *face_idx;
    const uint32_t *face_start, *face_group;
    const double *face_w;
} orc_merged_desc;
int orc_add_merged(orc_engine *e, const orc_merged_desc *d);
int orc_merged_state(orc_engine *e, int id, double *cur, double *circ);

void orc_run(orc_engine *e, uint64_t n);
/* FusedSteps (pool.rs:131-164): 0 Auto (fused when 24 bytes per node of a
 * thread's slab reach 4 MiB and the step allows it), 1 Always, 2 Never.
 * Results are bit-identical either way. */
void orc_set_fused(orc_engine *e, int mode);
/* MurBoundary (boundary/mur.rs): face edges (axis, node[3], inner[3], k,
 * hold) and corners (axis, node[3], inner0[3], inner1[3], k0, k1, hold);
 * holds may be NULL. Returns the operator id or -1. */
int orc_add_mur(orc_engine *e, int ne, const int *eax, const int *enode, const int *einner,
                const float *ek, const uint64_t *ehold, int nc, const int *cax, const int *cnode,
                const int *cin0, const int *cin1, const float *ck0, const float *ck1,
                const uint64_t *chold, int oe_order, int oe_corners);
/* OpenEmsPml (boundary/upml.rs:541-760): the layer edges, E then H dual
 * edges, each with axis, node[3] and four coefficients
 * ((1-a_p)/(1+a_p), (1-a_q)/(1+a_q), (1+a_a)/((1+a_q)m), (1-a_a)/((1+a_q)m));
 * the caller has already rewritten their update coefficients. Refused
 * with CPML slabs, after dispersive media, Mur boundaries, plane waves,
 * cut-face currents or another layer operator. Returns the id or -1. */
int orc_add_upml(orc_engine *e, int ne, const int *eax, const int *enode, const float *eco,
                 int nh, const int *hax, const int *hnode, const float *hco);
/* The layer state (flux, value) of side 0 (E) or 1 (H), in edge order. */
int orc_upml_state(orc_engine *e, int id, int side, float *flux, float *value);
/* Engine::energy_in(EnergyMeasure::OpenEms) (mod.rs:727-752): openEMS's
 * measure eps0 sum u^2 + mu0 sum w i'^2 of the current state (Cartesian
 * boundary rules), the layers' corrections and the H sources' look-ahead
 * energy; read-only. */
double orc_energy_openems(orc_engine *e);
/* The same without the layers' corrections (layers = 0), for the tests. */
double orc_energy_openems_ex(orc_engine *e, int layers);
double orc_step_with_energy(orc_engine *e);
double orc_energy(orc_engine *e);
uint64_t orc_timestep(orc_engine *e);

/* field 0 = u (E voltages), 1 = i (H currents). */
int orc_read(orc_engine *e, int field, int axis, float *out);
int orc_write(orc_engine *e, int field, int axis, const float *in);
uint64_t orc_probe_count(orc_engine *e, int id);
uint64_t orc_probe_values(orc_engine *e, int id, double *out, uint64_t cap);

/* ---- setup restatements, for cross-checking the product's host side ---- */

/* courant_limit (coeffs.rs:249-255) for per-axis line arrays. */
double orc_courant_limit(const double *x, int nx, const double *y, int ny, const double *z,
                         int nz);

/* Coefficients::from_edges_rounded (coeffs.rs:879-966) with Exact rounding,
 * Cartesian grid: per-edge eps_r/kappa (E, 3 arrays each) and mu_r/sigma_m
 * (H), pec mask per E axis (may be NULL), pec_face[6] (x_min..z_max).
 * Writes 12 dense lanes of N floats into `lanes` (12*N). */
int orc_coefficients(const double *x, int nx, const double *y, int ny, const double *z,
                     int nz, const double *eps_r, const double *kappa, const double *mu_r,
                     const double *sigma_m, const uint8_t *pec, const int pec_face[6],
                     double dt, float *lanes);

/* CpmlCoefficients::in_medium (cpml.rs:350-388): out = [rho,sigma,kappa,
 * alpha,b,a,direct]. convolution 0 Constant, 1 Linear. */
void orc_cpml_coefficients(double m, double sigma_factor, double kappa_max, double alpha_max,
                           int convolution, double rho, double delta, double dt, double n,
                           double out[7]);

/* run::drive (run.rs:424-540) over a scripted stepper whose energy at
 * timestep n is w[min(n, nw-1)] (dt = 1): the run-loop restatement used to
 * pin the end-criterion logic. Returns the stop reason (0 EnergyDecay,
 * 1 MaxTimesteps, 2 Aborted, 3 Diverged) and the steps in *steps. */
int orc_drive_script(const double *w, uint64_t nw, uint64_t max_steps, int has_criteria,
                     double energy_ratio, uint64_t interval, int has_excitation_end,
                     double excitation_end, uint64_t *steps);

#ifdef __cplusplus
}
#endif
#endif
