/*
 * wsph.h — C-ABI of libwsph.so, the MI355X (gfx950) IISPH step.
 *
 * Every entry point takes plain pointers and sizes; array arguments marked
 * "device" are HIP device pointers (a torch ROCm tensor's data_ptr()), arrays
 * marked "host" are ordinary host memory. Every call returns an int status
 * (WSPH_OK = 0, negative on error) and leaves a message readable with
 * wsph_last_error() (thread-local), the way the reference raises
 * RuntimeError from IISPH.raise_error (wp_sph/solver/IISPH.py:1484-1487).
 *
 * The reference has no FFI: its solver is the Python class
 * wp_sph.solver.IISPH (IISPH.py:736-1501) whose methods each issue
 * wp.launch(kernel, dim, inputs, outputs) calls. Each entry point below
 * names the reference code it replaces.
 */
#ifndef WSPH_H
#define WSPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WSPH_ABI_VERSION 2

/* The mean density error (update_rho_error, IISPH.py:542-568) is summed in
 * 64-bit fixed point so that it is independent of summation order: each
 * term |rho_star / rho0 - 1| (float32) is clamped at WSPH_RESIDUAL_TERM_CLAMP and
 * truncated to a multiple of 1/WSPH_RESIDUAL_FIXED_SCALE. */
#define WSPH_RESIDUAL_FIXED_SCALE 268435456.0 /* 2^28 */
#define WSPH_RESIDUAL_TERM_CLAMP 1024.0f

/* ---- status codes ---------------------------------------------------- */
enum {
  WSPH_OK = 0,
  WSPH_E_ARG = -1,        /* invalid argument or call order */
  WSPH_E_HIP = -2,        /* a HIP runtime call failed */
  WSPH_E_CAPACITY = -3,   /* neighbour list or grid capacity exceeded */
  WSPH_E_NAN = -4,        /* non-finite position or velocity after a step */
  WSPH_E_BARRIER = -5,    /* persistent kernel grid barrier timed out */
  WSPH_E_UNSUPPORTED = -6 /* device cannot run the requested mode */
};

/* ---- enumerations ------------------------------------------------------ */
enum { WSPH_LAYOUT_LINEAR = 0, WSPH_LAYOUT_MORTON = 1 };
/* pressure start value each step: 0 as IISPH.py:166-174, the previous
 * step's pressure, a caller-supplied field, or the attached pressure
 * predictor network (wsph_set_predictor). The previous step's pressure is
 * carried as zero for particles whose a_ii fell inside update_p's guard band
 * |a_ii| < INV_SMALL (IISPH.py:535).
 * WSPH_WARM_LEARNED: at step start a feature pass (8 features per particle,
 * see wsph_get_pairs) feeds the network, evaluated inline on the matrix
 * cores; the start value is max(0, p_carry + P_ref * net), the carried value
 * where |a_ii / a_ref| < 0.05 (weakly coupled particles), 0 inside the guard
 * band. The previous step's pressure arbitrates the network: after each step
 * both start values are compared with the converged pressure, and when the
 * network's error was the larger one the next step starts from the carried
 * pressure instead (the network is still evaluated and compared). */
enum { WSPH_WARM_ZERO = 0, WSPH_WARM_PREVIOUS = 1, WSPH_WARM_EXTERNAL = 2,
       WSPH_WARM_LEARNED = 3 };
/* dt: CFL rule of IISPH.py:1368-1370, or fixed */
enum { WSPH_DT_CFL = 0, WSPH_DT_FIXED = 1 };
/* Jacobi driver: one launch per kernel with a host convergence test per
 * iteration (the reference's loop, IISPH.py:1200-1296); one persistent
 * cooperative kernel with a device-side residual; or launches queued ahead
 * in chunks whose stop test runs on the device (no host round trip per
 * iteration) */
enum { WSPH_JACOBI_HOST = 0, WSPH_JACOBI_PERSISTENT = 1, WSPH_JACOBI_DEVICE = 2 };

/* fields readable with wsph_get_field (original particle order) */
enum {
  WSPH_F_X = 0,       /* float[n][3] positions              IISPH.py:993 */
  WSPH_F_V = 1,       /* float[n][3] velocities             IISPH.py:783 */
  WSPH_F_P = 2,       /* float[n]    pressure               IISPH.py:788 */
  WSPH_F_RHO = 3,     /* float[n]    density                IISPH.py:785 */
  WSPH_F_RHO_ADV = 4, /* float[n]    advected density       IISPH.py:786 */
  WSPH_F_V_ADV = 5,   /* float[n][3] advected velocity      IISPH.py:784 */
  WSPH_F_D_II = 6,    /* float[n][3] d_ii                   IISPH.py:793 */
  WSPH_F_A_II = 7,    /* float[n]    a_ii                   IISPH.py:792 */
  WSPH_F_AP = 8,      /* float[n]    A p                    IISPH.py:796 */
  WSPH_F_S = 9,       /* float[n][3] sum_j d_ij p_j         IISPH.py:797 */
  WSPH_F_FF_COUNT = 10, /* int[n]    fluid neighbour count  IISPH.py:798 */
  WSPH_F_FS_COUNT = 11, /* int[n]    boundary neighbour count IISPH.py:802 */
  WSPH_F_PSI = 12,    /* float[nb]   boundary psi           IISPH.py:789 */
  WSPH_F_CELL_KEY = 13, /* int[n]    linear hash key of each particle */
  WSPH_F_BX = 14,     /* float[nb][3] boundary positions    IISPH.py:989 */
  WSPH_F_BV = 15,     /* float[nb][3] boundary velocities   IISPH.py:991 */
  WSPH_F_COUNT
};

/* Solver parameters. The physics fields restate wp_sph/main.py:25-61; the
 * host converts each double to float once, as the reference's kernel
 * constants are float32. */
typedef struct wsph_params {
  double spacing;   /* DIAMETER                        main.py:26 */
  double h;         /* SMOOTHING_LENGTH = 2 DIAMETER   main.py:57 */
  double rho0;      /* RHO_0                           main.py:29 */
  double mass;      /* FLUID_MASS = RHO_0 d^3          main.py:59 */
  double gravity;   /* GRAVITY, acting along -y        main.py:25, IISPH.py:277 */
  double omega;     /* OMEGA (relaxed Jacobi)          main.py:27 */
  double eta;       /* ETA (mean density error bound)  main.py:28 */
  double inv_small; /* INV_SMALL guard on a_ii         main.py:30 */
  double vis_mu;    /* VIS_MU Laplacian viscosity      main.py:61, IISPH.py:273-278 */
  double xsph;      /* XSPH velocity smoothing coefficient (0 = off) */
  double dt;        /* fixed dt, or the first dt under CFL (IISPH.py:747) */
  double dt_max;    /* TIME_STEP_MAX = 0.05 d          main.py:58 */
  double cfl;       /* 0.4, IISPH.py:1369 */
  int dt_mode;      /* WSPH_DT_* */
  int min_iters;    /* 2   (IISPH.py:1201) */
  int max_iters;    /* 101 (IISPH.py:1202) */
  int force_iters;  /* >0: run exactly this many Jacobi iterations */
  int warm_start;   /* WSPH_WARM_* */
  int grid_dims[3]; /* fluid hash grid dims (IISPH.py:771-775) */
  int bgrid_dims[3];/* boundary hash grid dims (IISPH.py:776-780) */
  int cell_layout;  /* WSPH_LAYOUT_* : order of cells in memory */
  int max_ff;       /* fluid-fluid neighbour capacity per particle */
  int max_fs;       /* fluid-boundary neighbour capacity per particle */
  int jacobi_mode;  /* WSPH_JACOBI_* */
  int block;        /* threads per block of the neighbour passes (64..1024) */
  int coop_blocks_per_cu; /* persistent kernel blocks per CU (0 = as many as fit) */
  int chunk_iters;  /* WSPH_JACOBI_DEVICE: iterations queued per chunk (0 = 2) */
  int graph_iters;  /* WSPH_JACOBI_DEVICE, > 0: wsph_step replays a captured hipGraph of
                       the step's grid build, advection passes, start value and first
                       graph_iters Jacobi iterations; later iterations run as chunks */
} wsph_params;

/* Per-step statistics (what IISPH.step computes and prints, IISPH.py:1200-1371) */
typedef struct wsph_step_stats {
  int iterations;   /* Jacobi iterations run this step (2..101) */
  int converged;    /* final residual <= eta */
  int ap_evals;     /* A p evaluations (iterations, +1 for a non-zero start) */
  int max_ff_count; /* largest fluid-fluid neighbour count */
  int max_fs_count; /* largest fluid-boundary neighbour count */
  int status;       /* device-reported status of this step */
  double residual;  /* mean density error at exit (IISPH.py:1489-1494) */
  double dt;        /* dt used by this step */
  double sim_time;  /* simulated time after this step */
  double v_max;     /* max |v| after the kick (IISPH.py:688) */
} wsph_step_stats;

/* Per-step record of the statistics history kept on the device. */
typedef struct wsph_step_record {
  int step;             /* step index since wsph_create */
  int iterations;       /* Jacobi iterations */
  int converged;        /* residual <= eta */
  int start_source;     /* start value used: WSPH_WARM_* (LEARNED only when the
                           network's value was used) */
  double sim_time;      /* simulated time after the step */
  double residual;      /* mean density error at the end of the Jacobi loop */
  double kinetic_energy;/* 0.5 m sum |v|^2 after the step */
  double mean_height;   /* mean fluid y */
  double surface_height;/* highest fluid particle y (free-surface height) */
  double v_max;         /* max |v| */
  /* sqrt(mean_i (w_i (start value_i - converged p_i))^2) [Pa] with
   * w_i = min(|a_ii / a_ref|, 1), a particle's weight in the density
   * residual, for the network's prediction and for the carried previous
   * pressure; computed in WSPH_WARM_LEARNED mode, else 0 */
  double start_err_learned;
  double start_err_carry;
} wsph_step_record;

typedef struct wsph_solver wsph_solver;

const char *wsph_last_error(void);
int wsph_abi_version(void);

/* Neighbour predicate used by the grid queries: for float32 d2 >= 0,
 * sqrtf(d2) < h (the test of IISPH.py:72 and :129) holds exactly when
 * d2 < wsph_radius_sq_threshold(h). */
float wsph_radius_sq_threshold(float h);

/* Blocks of block_threads threads that fit on one compute unit at the same
 * time (no counterpart in the reference; it sizes the grid of the persistent
 * Jacobi kernel, whose grid-wide barrier needs every block resident): the
 * minimum over the wave slots (waves_per_simd on each of the 4 SIMDs, each
 * block counted as taking ceil(waves / 4) slots on every SIMD: exact for
 * blocks of 4 or more waves, a lower bound for smaller ones), the VGPRs
 * (512 per lane and SIMD, allocated in units of 8, vgprs per lane of the
 * kernel), the SGPRs, the LDS (lds_bytes per block, static plus dynamic,
 * allocated in units of 512 bytes, against cu_lds_bytes) and 16 workgroups
 * per CU. SGPRs: a wave holds its sgprs (the kernel's total, VCC included)
 * rounded up to 16, plus 16 more, out of 800 per SIMD; measured on MI355X,
 * kernels of 90 and 106 SGPRs fit 7 and 6 waves per SIMD where the runtime's
 * occupancy query reports 8 and 7. The runtime does not report a kernel's
 * SGPRs: sgprs <= 0 takes the most a kernel can hold (112), 6 waves per
 * SIMD. Host arithmetic only; WSPH_E_ARG for figures out of range. The
 * persistent driver reads the kernel's VGPRs and LDS with
 * hipFuncGetAttributes and launches this many blocks per CU (or fewer when
 * wsph_params.coop_blocks_per_cu asks for fewer). */
int wsph_blocks_per_cu(int block_threads, int vgprs, int sgprs, int lds_bytes, int cu_lds_bytes,
                       int waves_per_simd);

/* Self-test of the kernels' rounding helpers on the current device (no
 * counterpart in the reference, which leaves float32 division and sqrt to
 * Warp's codegen). The neighbour loops replace the compiler's correctly
 * rounded 1.0f / b, sqrtf and a / b (for many a over one b) by shorter exact
 * sequences; this runs them against the compiler's versions and writes the
 * mismatch counts to out[4] (host): out[0] reciprocal over all 2^32 float32
 * patterns, out[1] square root over every non-negative float32 up to +inf,
 * out[2] division over 2^29 pairs covering every significand of the divisor,
 * out[3] the neighbour-pair square root (exact from 2^-96 up, below 2^-47
 * under it). All four are 0 on a correct device; the parity tests check that. */
int wsph_selftest_arith(unsigned long long *out);

/* Create a solver (replaces IISPH.__init__'s allocations, IISPH.py:770-805). */
int wsph_create(const wsph_params *params, wsph_solver **out);
/* Release every device buffer the solver allocated. */
int wsph_destroy(wsph_solver *s);

/* Upload the W and grad-W/r tables (host float32), replaces the module-level
 * W_table / grad_W_table of sph_funcs.py:51-89. */
int wsph_set_tables(wsph_solver *s, const float *w_table, int w_size,
                    const float *grad_table, int grad_size);

/* Upload boundary particles (device float[nb][3] each; vb may be NULL for
 * zero velocity), build the boundary hash grid and compute psi
 * (IISPH.py:811-821 -> compute_boundary_density, IISPH.py:23-47).
 * At most 2^28 - 1 particles (WSPH_E_CAPACITY), here and in wsph_set_fluid. */
int wsph_set_boundary(wsph_solver *s, const float *xb, const float *vb, int nb);

/* Rigid motion of the boundary particles with original index >= first (the
 * moving model that follows the walls; the reference's boundary_hide_n,
 * IISPH.py:927). The step that starts at sim_time 0 sets their velocity to
 * v[0..2] (set_rigid_v, IISPH.py:652-660, launched at IISPH.py:1061-1081,
 * before predict_rho_adv); every step moves them by velocity * dt after the
 * fluid drift (update_rigid_x, IISPH.py:663-671, launched at
 * IISPH.py:1355-1364); and every neighbour search rebuilds the boundary hash
 * grid from the moved positions (DYNAMIC_SCENE, IISPH.py:1415-1416). psi
 * keeps its set-up values, as in the reference. Call after
 * wsph_set_boundary, which clears it; first < 0 clears it; not available
 * with batched scenes (WSPH_E_UNSUPPORTED). */
int wsph_set_rigid(wsph_solver *s, int first, const float *v);

/* Upload fluid state (device float[n][3]; v may be NULL for rest), replaces
 * the wp.array copies at IISPH.py:990-994. */
int wsph_set_fluid(wsph_solver *s, const float *x, const float *v, int n);

/* Set the pressure start value for the next step (device float[n], original
 * order). Used with WSPH_WARM_EXTERNAL. */
int wsph_set_pressure(wsph_solver *s, const float *p);

/* One full time step: IISPH.step (IISPH.py:1002-1371). With the
 * device-driven Jacobi loop the call returns once the loop has ended on the
 * device and the integration is enqueued; meanwhile the calling thread polls
 * two words in pinned host memory (busy-waiting with a pause instruction),
 * so that no event sits between the loop's kernels. */
int wsph_step(wsph_solver *s, wsph_step_stats *stats);

/* The four phases of IISPH.step as the reference times them
 * (IISPH.py:1004, 1007, 1154, 1298). wsph_step runs all four. */
int wsph_neighbor_search(wsph_solver *s);        /* IISPH.py:1408-1482 */
int wsph_predict_advection(wsph_solver *s);      /* IISPH.py:1009-1152 */
int wsph_pressure_solve(wsph_solver *s, wsph_step_stats *stats); /* IISPH.py:1155-1296 */
int wsph_integrate(wsph_solver *s, wsph_step_stats *stats);      /* IISPH.py:1299-1371 */

/* Copy a field into dst (device), scattered back to original particle order. */
int wsph_get_field(wsph_solver *s, int field, void *dst);

/* Neighbour lists in original indices and the reference's enumeration
 * order (count_/store_same_neighbor IISPH.py:52-104 when which = 0,
 * count_/store_diff_neighbor IISPH.py:107-163 when which = 1).
 * cnt: device int[n]; idx: device int[n][cap], row i holding particle i's
 * neighbours in order. */
int wsph_get_neighbors(wsph_solver *s, int which, int *cnt, int *idx, int cap);

/* Simulation clock (sim_time, dt) kept on the device (IISPH.py:1367-1371). */
int wsph_get_time(wsph_solver *s, double *sim_time, double *dt);

/* Stand-alone hash-key kernel: linear key cz*dx*dy + cy*dx + cx of the
 * uniform hash grid (warp.HashGrid, used at IISPH.py:771-780, 1414) for n
 * device points, written to device int keys[n]. */
int wsph_hash_keys(const float *x, int n, float cell_width, const int dims[3],
                   int *keys);

/* Phase timing with HIP events on the solver's stream, summed over every
 * step since wsph_set_timing(s, 1) (which resets the sums): out[0..3] ms of
 * neighbour search, advection, pressure solve and integration, out[4] ms of
 * whole steps, out[5] persistent-kernel grid size, out[6] steps recorded,
 * out[7] and out[8] ms of the J1 and J2 launches (A p, IISPH.py:431-518) of
 * each step's first Jacobi iteration (device-driven loop without a captured
 * graph), out[9] steps in those two sums, out[10] ms of the same J1 + J2
 * pairs from the event before J1 to the event after J2, out[11] pairs timed.
 * wsph_set_timing(s, 2) records only the two events around the pair, on
 * every 4th step (out[0..4] and out[7..9] stay 0): each event between two
 * kernels costs a few microseconds of device time, which mode 1 adds to every
 * step and to each of the J1 and J2 times. 0 switches timing off. */
int wsph_last_timings(wsph_solver *s, double *out, int nout);
int wsph_set_timing(wsph_solver *s, int enable);

/* Cumulative counters since wsph_create, kept on the device so that a run
 * of steps needs no host synchronisation: out[0] steps, out[1] Jacobi
 * iterations, out[2] steps that ended above eta, out[3] A p evaluations,
 * out[4] OR of the device status bits, out[5] most blocks seen at a
 * persistent-kernel barrier wait that timed out, out[6] persistent grid. */
int wsph_get_counters(wsph_solver *s, long long *out, int nout);

/* The last min(max_records, steps, 8192) step records, oldest first. The
 * reference prints none of these (it has no statistics export); they are
 * kept on the device so that long runs need no per-step synchronisation. */
int wsph_get_history(wsph_solver *s, wsph_step_record *out, int max_records, int *count);

/* ---- scene generation (host only, no GPU work) --------------------------
 * Poisson-disk sample of a boundary surface with minimum distance `radius`
 * (Bridson's algorithm run on the surface: candidates at distance
 * [radius, 2 radius) along the tangent plane are projected onto the surface;
 * a splitmix64 stream seeded with `seed` drives it, so the points depend only
 * on the arguments). WSPH_SURF_PLANE: params = {x0, z0, x1, z1, y}, the
 * rectangle [x0, x1] x [z0, z1] at height y; WSPH_SURF_TORUS: params =
 * {cx, cy, cz, R, r}, axis along y. Writes *count and, when out is not NULL,
 * the points as float[count][3] (host; WSPH_E_CAPACITY if count > max_points).
 * Replaces the reference's mesh-sampled boundaries (model.py:14-60) for the
 * torus-and-plane scene. */
enum { WSPH_SURF_PLANE = 0, WSPH_SURF_TORUS = 1 };
int wsph_poisson_surface(int shape, const double *params, double radius, uint64_t seed,
                         int max_points, float *out, int *count);

/* Lattice points inside a closed triangle mesh: the containment test of the
 * reference's model sampler (_sample_points, model.py:14-47, which keeps the
 * np.arange lattice points that trimesh's Trimesh.contains accepts).
 * inside[(i * ny + j) * nz + k] = 1 when (ax[i], ay[j], az[k]) is inside,
 * else 0: ray parity along +z, columns through an edge or vertex of the
 * projected mesh resolved by the rasterisation top-left rule, so watertight
 * meshes are classified consistently; a point on the surface counts the
 * crossings strictly above it. ax, ay and az must be strictly ascending
 * (WSPH_E_ARG otherwise), as np.arange lattices are. Device arrays:
 * verts double[nv][3], tris int[nt][3] (vertex indices), ax/ay/az
 * double[nx]/[ny]/[nz], inside uint8[nx * ny * nz]. Blocks until done. */
int wsph_mesh_inside(const double *verts, int nv, const int *tris, int nt, const double *ax, int nx,
                     const double *ay, int ny, const double *az, int nz, unsigned char *inside);
/* wsph_mesh_inside for callers that cannot allocate (code being captured
 * into a HIP graph, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference. Same arguments and the same result, bit
 * for bit; it calls no hipMalloc/hipFree, creates no stream or event and
 * allocates no host memory. Every scratch buffer comes from one fixed
 * 64 MiB device array per GPU, held until the process ends: 128 bytes per
 * triangle, 16 bytes per 16 x 16 column tile, 12 bytes per lattice column
 * (nx * ny), 4 bytes per triangle-tile pair, 16 bytes per work item of up to
 * 32 triangles and 8 bytes per column crossing, each region rounded up to
 * 256 bytes. A call that needs more returns WSPH_E_CAPACITY, before any
 * device work when the sizes alone exceed it. Runs on the null stream and
 * blocks until done; records no timings. Calls that overlap on one GPU (two
 * host threads) race on its array: serialise them. */
int wsph_mesh_inside_noalloc(const double *verts, int nv, const int *tris, int nt,
                             const double *ax, int nx, const double *ay, int ny,
                             const double *az, int nz, unsigned char *inside);
/* Kernel durations (ms, HIP events) of the calling thread's last
 * wsph_mesh_inside, out[0..4]: triangle set-up, binning of the triangles into
 * 16 x 16 column tiles, crossing count pass, crossing fill pass, sort and
 * classification; out[5] the triangle-tile pairs and out[6] the column
 * crossings of that call. No reference counterpart. */
int wsph_mesh_timings(double *out);

/* ---- batched scenes (no reference counterpart) -------------------------
 * Many independent scenes solved together: call after wsph_create and before
 * the particle calls (wsph_set_boundary / wsph_set_fluid), whose
 * arrays then hold the scenes' particles concatenated in scene order. Each
 * scene keeps its own fluid and boundary hash grids (dims per scene, 3 ints
 * each), so particles of different scenes never interact and every scene
 * evolves exactly as it would alone. Each scene runs its own Jacobi stop
 * test (IISPH.py:1200-1207) and stops iterating when it has converged while
 * the others continue. Requires WSPH_DT_FIXED (one dt for all scenes) and
 * the device or persistent Jacobi driver; params.grid_dims are unused. */
int wsph_set_scenes(wsph_solver *s, int nscenes, const int *fluid_counts,
                    const int *boundary_counts, const int *grid_dims, const int *bgrid_dims);
/* Per-scene results of the last step (host arrays of nscenes entries, each
 * optional): Jacobi iterations, residual <= eta, residual at exit. */
int wsph_get_scene_stats(wsph_solver *s, int *iterations, int *converged, double *residual);

/* ---- learned pressure start value (no reference counterpart) -----------
 * Per-particle network 8 -> 64 -> 64 -> 1 with SiLU, evaluated on the
 * matrix cores with BF16 operands and FP32 accumulation. Parameters are a
 * float32 vector of wsph_mlp_nparams() = 4801 values: W1[64][8], b1[64],
 * W2[64][64], b2[64], W3[64], b3. */
typedef struct wsph_mlp wsph_mlp;
int wsph_mlp_create(int max_batch, wsph_mlp **out);
int wsph_mlp_destroy(wsph_mlp *m);
int wsph_mlp_nparams(void);
int wsph_mlp_set_params(wsph_mlp *m, const float *params);   /* host; resets Adam */
int wsph_mlp_get_params(wsph_mlp *m, float *params);         /* host */
/* y[i] = net(x[i][0..7]) for n rows (device pointers) */
int wsph_mlp_forward(wsph_mlp *m, const float *x, int n, float *y);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the loss 0.5 mean (net(x) - t)^2
 * over n rows (device x[n][8], t[n]); lr <= 0 only evaluates. loss and grad
 * (host, optional) receive the loss and the gradient before the update. */
int wsph_mlp_train_step(wsph_mlp *m, const float *x, const float *t, int n, float lr,
                        float *loss, float *grad);

/* Attach a predictor network (kept by reference, not owned; NULL detaches)
 * for WSPH_WARM_LEARNED. No reference counterpart. */
int wsph_set_predictor(wsph_solver *s, wsph_mlp *m);
/* Keep each step's feature rows and regression targets (pair logging) for
 * wsph_get_pairs (on != 0), at 36 bytes per particle of device memory. */
int wsph_set_pair_logging(wsph_solver *s, int on);
/* Training pairs of the last step, in original particle order:
 * features[n][8] (device) and targets[n] (device) with
 * target = (converged p - carried p) / P_ref, the quantity the network
 * predicts. P_ref = rho0 * max(|g|, 1) * 0.1. Feature columns: 0 density
 * error (rho_adv / rho0 - 1) * 100; 1 a_ii / a_ref, with a_ref the a_ii of a
 * particle inside a resting lattice of the solver's spacing at the current
 * dt; 2 fluid neighbour count / 32; 3 carried pressure / P_ref (columns
 * 0-3 and the target clamped to +-64); 4-7 the kernel-weighted means of
 * columns 0-3 over the particle and its fluid neighbours,
 * sum_j W_ij f_j / sum_j W_ij. Requires pair logging or
 * WSPH_WARM_LEARNED; fails before the first step. */
int wsph_get_pairs(wsph_solver *s, float *features, float *targets);

/* Block until the solver's stream is idle. */
int wsph_sync(wsph_solver *s);

#ifdef __cplusplus
}
#endif
#endif /* WSPH_H */
