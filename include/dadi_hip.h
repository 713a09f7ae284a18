/*
 * dadi_hip.h -- C ABI of libdadi_hip.so, the MI355X (gfx950) implementation of
 * dadi's ADI time stepping of the allele-frequency density phi in 1-5
 * populations and of the phi -> frequency-spectrum sampling integral.
 *
 * Conventions
 *   - Every pointer named *_dev is a device pointer; the caller owns all
 *     memory (the Python layer allocates it as torch tensors).  The library
 *     never allocates device memory; its only device storage is the static
 *     scratch array of dh_integrate_factored_static.
 *   - Every entry point returns 0 on success and a negative DH_E* code on
 *     failure; dh_last_error() then holds a message.  Errors are per host
 *     thread.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls only enqueue work; they do not synchronise.
 *   - All arithmetic is IEEE FP64 with no contraction into FMA, in the
 *     reference C's operation order, so results are bit-identical to
 *     dadi/integration{1..5}D.c and dadi/tridiag.c on the same inputs.
 */
#ifndef DADI_HIP_H
#define DADI_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_MAXD 5

#define DH_OK 0
#define DH_EINVAL -1   /* bad argument */
#define DH_EHIP -2     /* HIP runtime error */
#define DH_ERANGE -3   /* size beyond what the kernel supports */

/* Coefficient/RHS order of the implicit system (see DH_ORDER_* users):
 *  TV   = dadi/integration{1..5}D.c implicit_* : b = 1/dt + fluxes, r = phi/dt
 *  PRE  = dadi/Integration.py:1037-1242 + implicit_precalc_{2,3}D* :
 *         b = (fluxes + boundary) + 1/dt, r = (1/dt)*phi
 *  PRE1 = dadi/Integration.py:986-1035 (1D constant parameters):
 *         b = (fluxes + boundary) + 1/dt, r = phi/dt                      */
#define DH_ORDER_TV 0
#define DH_ORDER_PRE 1
#define DH_ORDER_PRE1 2

/* A run of identical time steps: every step of the segment injects
 * `inject` at index 1 of each injecting axis (dadi/Integration.py:70-145),
 * then sweeps every non-frozen axis in order with step `dt`.  Parameters are
 * per population k; mig[k][j] is the migration rate into population k from
 * the j-th other population counted in increasing order (axis 1 of a
 * 3-population phi: m21, m23). beta_p, beta_r, beta_q scale V in 1D only:
 * V = 1/nu x (1-x) * beta_p * beta_r / beta_q, evaluated left to right:
 *   autosome: beta_p = pow(beta+1,2), beta_r = 1, beta_q = 4*beta
 *             (dadi/integration_shared.c:8-10);
 *   X (chrom = 1): beta_p = 2*beta+4, beta_r = beta+1, beta_q = 9*beta, and the
 *             selection term is gamma*4/3*(0.5+h+x(1-2h))*x(1-x)
 *             (dadi/Integration.py:1244-1247). */
typedef struct dh_segment {
    long long nsteps;
    double dt;
    double inject;
    double beta_p, beta_q;
    double nu[DH_MAXD];
    double gamma[DH_MAXD];
    double h[DH_MAXD];
    double mig[DH_MAXD][DH_MAXD - 1];
    double beta_r;
    int chrom;      /* 0: autosome; 1: X chromosome (1 population only) */
    int reserved;
} dh_segment;

/* One integration ("instance").  phi_dev holds phi in C order, updated in
 * place; its last axis may be padded: consecutive rows of the last axis are
 * ld doubles apart (ld >= n; 0 means n, i.e. dense).  A pitch that is a
 * multiple of 8 keeps every row 64-byte aligned.  cur_seg/cur_step record
 * progress: a launch resumes from them and writes them back, so a schedule
 * can be run in several quanta. */
typedef struct dh_instance {
    double *phi_dev;
    const double *xx_dev;
    const dh_segment *seg_dev;
    long long nseg;
    long long cur_seg;
    long long cur_step;
    int D;
    int n;
    int order;
    int frozen;  /* bit k set: population k frozen (its axis is not swept) */
    int inject;  /* bit k set: population k receives new mutations */
    int status;  /* 1 when the schedule has completed */
    int ld;      /* row pitch of the last axis in doubles (0 = n) */
    int reserved; /* 0 from the caller; bit 0 is set by dh_integrate_resident for an
                     instance it integrates with exact division from then on */
} dh_instance;

/* Launch shape of the workgroup-per-instance kernel. */
typedef struct dh_launch {
    int waves;             /* 64-lane waves per workgroup, 1..8 */
    int groups;            /* persistent workgroups; 0 = as many as fit, capped by ninst */
    long long max_steps;   /* steps per instance in this call; <= 0 = to completion */
    int chunk;             /* Thomas scratch: 0 = full gamma column per lane in LDS;
                              8 = bet checkpoint every 8 rows, gammas recomputed in
                              registers (less LDS, more VALU, same bits) */
    int recompute;         /* with chunk 8: 1 = also checkpoint u' and recompute it in the
                              back substitution from the untouched phi rows instead of
                              writing u' over phi (24 instead of 32 bytes of phi traffic
                              per point, same bits) */
    double *gamma_dev;     /* NULL, or (with chunk 0, recompute 0) a device workspace for
                              the gamma columns: no recomputation and no gamma LDS, at
                              16 more bytes of memory traffic per point and sweep; it must
                              hold groups * waves * 64 * nmax doubles (dh_nd_groups) and
                              may not be shared by launches that overlap */
    long long gamma_doubles;
    int delj;              /* 1: Chang and Cooper's delta per interval (dadi's
                              use_delj_trick) instead of 1/2; needs chunk 0 (no
                              recompute, no workspace) and is not available to
                              dh_integrate_packed.  Device exp: results agree with
                              the reference to rounding, not bit for bit. */
    int max_us;            /* > 0: time budget of this call (dh_integrate_nd,
                              dh_integrate_1d_wave): a workgroup takes no further time
                              step of an instance once this many microseconds have
                              passed since it took the instance up (at least one step
                              is taken); progress is recorded as with max_steps.
                              0 = none. */
} dh_launch;

/* Library / device set-up; must be called once per process before use.
 * Replaces the pycuda context creation of dadi/cuda/__init__.py:3-10. */
int dh_init(int device);
/* Copies the last error message of this thread into buf; returns its length. */
int dh_last_error(char *buf, int len);
/* ABI self-description, checked by the Python loader. */
int dh_abi_version(void);
size_t dh_sizeof_segment(void);
size_t dh_sizeof_instance(void);
/* LDS footprint per workgroup of the nD kernel for (D, n, waves, chunk,
 * recompute, gamma in a global workspace or not), bytes. */
size_t dh_lds_bytes(int D, int n, int waves, int chunk, int recompute, int gamma_global,
                    int delj);
/* Number of persistent workgroups dh_integrate_nd launches for this launch
 * shape (`groups` if > 0, else as many as fit), capped at ninst; sizes the
 * gamma workspace.  Negative DH_E* code on error. */
int dh_nd_groups(int D, int nmax, int waves, int chunk, int recompute, int gamma_global,
                 int groups, int ninst);
/* Largest LDS a workgroup may use on the initialised device (0 before dh_init). */
size_t dh_max_lds(void);

/*
 * Runs ninst instances, one workgroup per instance at a time, pulled from a
 * device-side queue.  All instances must share D and order.  counter_dev is
 * one int of device memory used as the queue head (reset by this call).
 * Each workgroup keeps the per-axis coefficient vectors in LDS and one Thomas
 * scratch column per lane.  Replaces, per time step, the Python-driven loop of
 * dadi/Integration.py:1108-1116 / 355-386 / 514-553 / 682-726 / 886-936 and
 * the reference GPU path dadi/cuda/Integration.py:89-995.
 */
int dh_integrate_nd(dh_instance *table_dev, int ninst, int D, int order, int nmax,
                    int *counter_dev, const dh_launch *cfg, void *stream);

/*
 * Like dh_integrate_nd, with `slots` instances per workgroup stepped in
 * lockstep so that their tridiagonal lines fill the waves (a 2-population phi
 * on 100 points has 100 lines per sweep; three of them fill 300 of 320 lanes
 * of a 5-wave workgroup).  Every instance must have D populations, n points,
 * the same order, grid (xx) and row pitch; each keeps its own schedule, and a
 * finished slot takes the next instance from the queue.  cfg->chunk must be
 * 8 (checkpointed gammas) with no recompute and no workspace.  Same results,
 * bit for bit, as dh_integrate_nd.
 */
int dh_integrate_packed(dh_instance *table_dev, int ninst, int D, int order, int n, int slots,
                        int *counter_dev, const dh_launch *cfg, void *stream);
/* LDS per workgroup of dh_integrate_packed, bytes. */
size_t dh_packed_lds_bytes(int D, int n, int waves, int slots);

/*
 * Runs ninst 1-population instances, one lane per instance, lanes pulling
 * instances from a device-side queue independently (lane-level continuous
 * batching).  scratch_dev must hold 6 * nmax * lanes doubles (8 with
 * cfg->delj) where lanes = 64 * waves * groups; nmax is the largest n in the
 * table.
 * Replaces dadi/Integration.py:986-1035 / 219-258 with
 * dadi/tridiag.c:36-40 and dadi/integration1D.c:7-66.
 */
int dh_integrate_1d(dh_instance *table_dev, int ninst, int order, int nmax,
                    int *counter_dev, double *scratch_dev, long long scratch_doubles,
                    const dh_launch *cfg, void *stream);
/*
 * Runs ninst 1-population instances, one wave per instance with phi and the
 * tridiagonal system in LDS (latency path: a step costs a few dependent
 * operations per grid point).  Within a segment every step has the same
 * matrix, so its Thomas factorisation (gamma_j, bet_j of dadi/tridiag.c) is
 * formed once per segment with the per-step operations and only the
 * right-hand side is solved per step; results are the bits of
 * dh_integrate_1d.  Needs dh_1d_wave_lds_bytes(nmax) of LDS per workgroup
 * (at most dh_max_lds()); no scratch.  cfg->groups = 0 picks the grid;
 * cfg->max_steps and cfg->delj as for dh_integrate_1d.
 * Replaces the same reference code as dh_integrate_1d
 * (dadi/Integration.py:986-1035 / 219-258, dadi/tridiag.c:36-40,
 * dadi/integration1D.c:7-66).
 */
int dh_integrate_1d_wave(dh_instance *table_dev, int ninst, int order, int nmax,
                         int *counter_dev, const dh_launch *cfg, void *stream);
/*
 * Runs ninst instances of 2-5 populations, one workgroup per instance at a
 * time (latency path).  Within a segment every step has the same tridiagonal
 * matrices, so the Thomas factorisation of every line of every axis (a_j,
 * bet_j, 1/bet_j, gamma_j; dadi/tridiag.c:22-29 with the coefficients of
 * dadi/integration_shared.c:72-93) is formed once per segment with the
 * per-step operations into a workspace, and a step only solves the right-hand
 * sides: a few dependent operations per row instead of the coefficient build
 * and three divisions.  Results are the bits of dh_integrate_nd.  More memory
 * traffic per point than dh_integrate_nd, so it is for batches whose time is
 * set by their longest integrations rather than by bandwidth: a single
 * integration, a serving loop, 2D batches of up to a few thousand
 * integrations (on MI355X faster than dh_integrate_packed up to 2048, slower
 * at 8192).  Launching two workgroups per compute unit pays for large
 * batches on small grids (workspaces staying in the last-level cache).
 * cfg->gamma_dev / gamma_doubles is the
 * workspace: groups * dh_factored_slice_doubles(D, nmax) doubles, groups =
 * cfg->groups if > 0, else min(ninst, compute units).  cfg->waves 1..8;
 * cfg->max_steps and cfg->max_us as for dh_integrate_nd; no delta trick.
 * Replaces the same reference code as dh_integrate_nd.
 */
int dh_integrate_factored(dh_instance *table_dev, int ninst, int D, int order, int nmax,
                          int *counter_dev, const dh_launch *cfg, void *stream);
/* Workspace doubles one workgroup of dh_integrate_factored needs (0 if D or nmax is out of range). */
long long dh_factored_slice_doubles(int D, int nmax);
/*
 * dh_integrate_factored without any caller-provided device memory besides
 * the instance table, phi, grids and schedules: the factorisation workspace
 * and the queue counter live in a static device array of the library
 * (dh_static_scratch_doubles() doubles of workspace per GPU, 32 MiB, held
 * until the process ends).  Allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid allocation).
 * cfg->gamma_dev must be NULL.  cfg->groups = 0 launches min(ninst, compute
 * units, workgroups whose workspaces fit); an explicit cfg->groups, or a
 * single workgroup's dh_factored_slice_doubles(D, nmax), beyond the scratch
 * is rejected with DH_ERANGE.  Results are those of dh_integrate_factored.
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one
 * GPU (other streams or host threads) race and must be serialised by the
 * caller.
 */
int dh_integrate_factored_static(dh_instance *table_dev, int ninst, int D, int order,
                                 int nmax, const dh_launch *cfg, void *stream);
/* Workspace doubles of the static scratch of dh_integrate_factored_static. */
long long dh_static_scratch_doubles(void);

/*
 * Runs ninst 2-population instances, one workgroup per instance at a time,
 * pulled from a device-side queue, with phi resident in registers: a
 * workgroup reads phi when it takes an instance up and writes it back when it
 * puts it down, and every time step in between runs on chip (each lane holds
 * one line of phi; the lines are transposed through LDS between the two
 * sweeps).  Results are the bits of dh_integrate_nd.  A kernel serves 8
 * consecutive grid sizes: with R = dh_resident_rows(n), every instance must
 * have between R - 7 and R points; order DH_ORDER_TV or DH_ORDER_PRE.  An
 * instance of another shape is left untouched with status -1.
 * Two kernels run in turn on the stream: the first divides by bet_j through
 * one shared reciprocal per row (the operations of IEEE division where no
 * operand or quotient is near the ends of the exponent range, checked at
 * run time); an instance where a check fails is left as it was taken up,
 * marked (bit 0 of `reserved`) and integrated by the second, exactly
 * dividing, kernel, in this call and every later one.  A time budget
 * (cfg->max_us) applies to each of the two kernels.  cfg->groups (0 = one
 * workgroup per pair of SIMDs) and cfg->max_steps / cfg->max_us as for
 * dh_integrate_nd; no workspace, no delta trick.
 * Replaces, like dh_integrate_nd, the per-step loop of
 * dadi/Integration.py:1108-1116 / 355-386 with implicit_precalc_2Dx/2Dy or
 * implicit_2Dx/2Dy (dadi/integration2D.c:19-217) and dadi/tridiag.c:14-34.
 */
int dh_integrate_resident(dh_instance *table_dev, int ninst, int order, int n,
                          int *counter_dev, const dh_launch *cfg, void *stream);
/* Rows R of the dh_integrate_resident kernel for n points (it serves
 * n = R-7 .. R; R is a multiple of 8), or 0 when there is none (n outside
 * 33..104). */
int dh_resident_rows(int n);
/*
 * Latency form of dh_integrate_resident, same arguments and results (the bits
 * of dh_integrate_nd), for 33 to 128 points (R = dh_resident_latency_rows(n);
 * every instance between R - 7 and R points): phi resident in registers, and at the start
 * of each constant-parameter segment each lane forms its line's Thomas
 * factorisation once (a_j, bet_j, its reciprocal, gamma_{j+1}) into the
 * workspace cfg->gamma_dev, so that a time step runs only the right-hand-side
 * recurrences: three quarters of dh_integrate_resident's step time for one
 * integration (18 against 27 us at 100 points), more memory traffic per step
 * (32 bytes of records per point and sweep, from the L2), so it is meant for
 * single integrations, batches too small to fill the GPU and grids above 104
 * points, where dh_integrate_resident has no kernel.  cfg->gamma_dev must hold
 * groups * dh_resident_latency_slice_doubles(n) doubles (groups = cfg->groups,
 * or 0: min(ninst, workgroups that fit)) and may not be shared by launches
 * that overlap.  An instance whose checks fail is redone at once, from the
 * progress it was taken up with, by the same workgroup in the exactly
 * dividing form (IEEE divisions, the same factorisation), and is not marked;
 * instances marked by an earlier call (bit 0 of `reserved`, e.g. by
 * dh_integrate_resident) run on that form in the same call.
 * Replaces what dh_integrate_resident replaces.
 */
int dh_integrate_resident_latency(dh_instance *table_dev, int ninst, int order, int n,
                                  int *counter_dev, const dh_launch *cfg, void *stream);
/* Workspace doubles per workgroup of dh_integrate_resident_latency for n
 * points, or 0 when there is no kernel (n outside 33..128). */
long long dh_resident_latency_slice_doubles(int n);
/* Rows R of the dh_integrate_resident_latency kernel for n points (it
 * serves n = R-7 .. R; R is a multiple of 8), or 0 (n outside 33..128). */
int dh_resident_latency_rows(int n);
/* LDS per workgroup of dh_integrate_1d_wave, bytes. */
size_t dh_1d_wave_lds_bytes(int nmax);


/*
 * One implicit half-step along `axis` for a single instance, spread over as
 * many workgroups as the axis has 64-line tiles.  The step is taken from the
 * schedule: s = *step_base_dev + step_offset; seg_of_step_dev[s] names the
 * segment; nothing happens when s >= total_steps.  first_axis = 1 makes the
 * sweep add the step's injected mass first.  Used by the plain one-call path
 * for large phi, captured per step into a hipGraph.  Replaces one
 * implicit_<D>D<axis> call (dadi/integration{1..5}D.c) /
 * implicit_precalc_{2,3}D<axis>.
 */
int dh_sweep_axis(const dh_instance *inst_host, int axis, int first_axis,
                  const long long *step_base_dev, long long step_offset,
                  const int *seg_of_step_dev, long long total_steps,
                  int waves, void *stream);
/* Adds `by` to *step_base_dev (one thread), closing a captured group of steps. */
int dh_advance_steps(long long *step_base_dev, long long by, void *stream);
/* Captures `nsteps_per_graph` steps of dh_sweep_axis launches (all active axes)
 * plus dh_advance_steps into a hipGraph and replays it until total_steps are
 * done.  The graph is cached per instance descriptor and reused. */
int dh_integrate_graph(const dh_instance *inst_host, long long *step_base_dev,
                       const int *seg_of_step_dev, long long total_steps,
                       int nsteps_per_graph, int waves, void *stream);
/* As dh_integrate_graph, for steps first_step .. first_step+nsteps-1 only
 * (capped at total_steps): a long integration advanced in quanta, each call
 * resuming where the previous one stopped (the serving loop's multi-workgroup
 * path for large phi).  nsteps must be a multiple of nsteps_per_graph unless
 * the range reaches total_steps.  Same graph as dh_integrate_graph, same
 * results: splitting an integration into ranges does not change any bit.
 * Replaces the same per-step loop as dh_integrate_graph
 * (dadi/Integration.py:682-726, 886-936). */
int dh_integrate_graph_range(const dh_instance *inst_host, long long *step_base_dev,
                             const int *seg_of_step_dev, long long total_steps,
                             long long first_step, long long nsteps, int nsteps_per_graph,
                             int waves, void *stream);

/* One integration of a multi-workgroup batch (dh_integrate_batch). */
typedef struct dh_sweep_item {
    dh_instance inst;              /* phi, grid, schedule, n, order, masks, pitch */
    const int *seg_of_step_dev;    /* segment index of each time step */
    long long total_steps;
    int first_axis;                /* first non-frozen axis: receives the step's injection */
    int wg_begin;                  /* first workgroup of this integration in a sweep launch */
    int wgs;                       /* workgroups sweeping its lines */
    int step_shift;                /* this integration's step = batch step + step_shift */
} dh_sweep_item;

/*
 * A batch of integrations of D >= 2 populations advanced together, each sweep
 * of each step spread over many workgroups: one launch per axis per step
 * covers every integration (workgroup w sweeps lines of integration
 * wg_item_dev[w]; items[i].wgs workgroups share integration i's lines), the
 * steps of all integrations run in lockstep from a device step counter, and
 * an integration past its last step is skipped (integration i takes step
 * s + items[i].step_shift at batch step s, so integrations at different points
 * of their schedules can share the launches).  nsteps_per_graph steps
 * (all axes plus the counter advance) are captured into a hipGraph and
 * replayed for steps first_step .. first_step+nsteps-1 (nsteps a multiple of
 * nsteps_per_graph unless the range reaches the longest schedule).  For
 * batches too small to fill the GPU with one workgroup per integration (a
 * 3-population phi on 60 points sweeps 3,600 lines).  Same results as
 * dh_integrate_nd.  Replaces, for each integration, the per-step loop of
 * dadi/Integration.py:514-553 / 682-726 / 886-936 like dh_integrate_graph.
 */
int dh_integrate_batch(const dh_sweep_item *items_dev, const int *wg_item_dev, int ninst,
                       int total_wgs, int D, int order, int nmax, int waves,
                       long long *step_base_dev, long long max_total_steps,
                       long long first_step, long long nsteps, int nsteps_per_graph,
                       void *stream);
size_t dh_sizeof_sweep_item(void);

/*
 * Sampling phi -> expected spectrum (dadi/Spectrum_mod.py:1307-1348,
 * 1509-1548, 1774-1871): contracts the last axis of in_dev, shaped
 * (P, n), against the binomial sampling weights of sample size m and writes
 * out_dev shaped (m+1, P) (the new sample axis first), i.e.
 *   out[k][p] = sum_i c1[p][i]*db1[k][i] + (k+1)/((m+1)(m+2)) * sum_i s[p][i]*db2[k][i]
 * with s = dphi/dx and c1 = (phi - s*x)/(m+1) of the piecewise-linear phi.
 * db1_dev/db2_dev are (m+1, n-1) tables of regularised incomplete-beta
 * differences; work_dev needs 2*P*(n-1) doubles.  Calling it D times
 * (last axis first) yields the (m0+1, ..., m_{D-1}+1) spectrum.
 */
int dh_sample_axis(const double *in_dev, long long P, int n, const double *xx_dev,
                   int m, const double *db1_dev, const double *db2_dev,
                   double *work_dev, double *out_dev, void *stream);

/*
 * Direct (trapezoid-rule) sampling along the last axis: out (m1, P) =
 * A (m1, n) contracted with in (P, n) over n, i.e. out[k][p] = sum_i A[k][i] *
 * in[p][i].  A[k][i] is the binomial sampling factor of entry k at grid point
 * i (times x(1-x) for a heterozygote-ascertained axis) times the trapezoid
 * weight of point i.  Replaces the direct samplers
 * dadi/Spectrum_mod.py:1263-1279, 1350-1380, 1550-1592, 1594-1643 (used for
 * het_ascertained and force_direct); calling it D times, last axis first,
 * yields the spectrum.
 */
int dh_contract_axis(const double *in_dev, long long P, int n, const double *A_dev, int m1,
                     double *out_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
