/*
 * dmipy_walk.h -- C ABI of the MI355X Monte-Carlo spin walk.
 *
 * The reference has no FFI: its seam is the Python call
 *   simulate(n_walkers, diffusivity, waveform, geometry, seed, T2, T1, ...)
 *     dmipy_sim/core.py:193-762
 *   simulate_cpmg(n_walkers, diffusivity, waveform, geometry, T2, seed, ...)
 *     dmipy_sim/core.py:861-969
 * whose hot part is the jax.vmap(walker) . lax.scan(n_t) . lax.scan(n_sub) walk
 * (core.py:700-727, body physics.py:293-415) over the geometry methods
 * reflect / reflect_with_log_weight / permeate / classify_position
 * (geometries.py:93-139, 197-424, 516-813, 1746-2214, 2445-2771) and the ensemble
 * reduction _ens (core.py:426-429).  Everything from the walker keys to _ens is
 * replaced by the entry points below; the Python layer (dmipy_sim_amd) keeps the
 * reference's call signatures and calls them through ctypes.
 *
 * Conventions
 *  - Plain C types only; all pointers in the descriptor structs are HOST pointers
 *    owned by the caller for the duration of the call, except where a field says
 *    "device".
 *  - Vectors are in the geometry's native frame (cylinder axis = z); the caller
 *    rotates lab vectors in and out.
 *  - Every entry point returns 0 on success or a dmw_status code; the message of
 *    the last failure on the calling thread is dmw_last_error().  DMW_ENOMEM
 *    messages contain "out of memory" so the reference's OOM back-off
 *    (core.py:366-394) keys on them unchanged.
 *  - Calls on one thread are serialised on one HIP stream per thread; the library
 *    allocates device memory itself and frees it before returning (plans keep
 *    theirs until dmw_plan_destroy).
 */
#ifndef DMIPY_WALK_H
#define DMIPY_WALK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dmw_geom_kind {
  DMW_FREE = 0,              /* FreeDiffusion      geometries.py:34-45     */
  DMW_BOX1D = 1,             /* Box1D              geometries.py:48-140    */
  DMW_CYLINDER = 2,          /* Cylinder           geometries.py:427-844   */
  DMW_SPHERE = 3,            /* Sphere             geometries.py:143-424   */
  DMW_PACKED_CYLINDERS = 4,  /* PackedCylinders    geometries.py:1604-2242 */
  DMW_PACKED_SPHERES = 5,    /* PackedSpheres      geometries.py:2343-2786 */
  DMW_ELLIPSOID = 6,         /* Ellipsoid          geometries.py:1137-1425 */
  DMW_PERMEABLE_SLAB = 7,    /* PermeableSlab1D    geometries.py:3221-3300 */
  DMW_PERMEABLE_SHELL = 8,   /* PermeableShell, kind='sphere'   geometries.py:3303-3431 */
  DMW_PERMEABLE_SHELL_CYL = 9, /* PermeableShell, kind='cylinder' (axis z in the native frame) */
  DMW_MYELINATED_CYLINDER = 10, /* MyelinatedCylinder geometries.py:847-1136, step physics.py:420-679 */
  DMW_PACKED_MYELINATED = 11   /* PackedMyelinatedCylinders geometries.py:2860-3220,
                                  step physics.py:682-1062 */
};

enum dmw_wall {
  DMW_WALL_REFLECT = 0,      /* geometry.reflect                 physics.py:360-415 */
  DMW_WALL_SURFACE = 1,      /* geometry.reflect_with_log_weight physics.py:319-358 */
  DMW_WALL_PERMEATE = 2      /* geometry.permeate                physics.py:276-317 */
};

enum dmw_status {
  DMW_OK = 0,
  DMW_EINVAL = 1,            /* bad argument              -> ValueError          */
  DMW_EUNSUPPORTED = 2,      /* unsupported combination   -> NotImplementedError */
  DMW_ENOMEM = 3,            /* hipErrorOutOfMemory       -> RuntimeError("out of memory") */
  DMW_ENODEV = 4,            /* no usable GPU             -> RuntimeError        */
  DMW_EHIP = 5               /* any other HIP failure     -> RuntimeError        */
};

/* MyelinatedCylinder compartments 0 = intra-axonal, 1 = myelin sheath, 2 =
 * extra-axonal (make_myelin_step_fn, physics.py:420-679).  One step per waveform
 * sample (run->n_sub must be 1); run->step_len and run->inv_T2 are unused.  The
 * signal is sum_w spin_weight[c0(w)] e^{log_w} e^{i phi} / sum_w spin_weight[c0(w)]
 * (_ens, core.py:426-429, 486-530), c0 the start compartment: by walker index
 * when walkers are placed in-kernel (geometries.py:921-1008 allots them in
 * index order), else classify_position of run->r0 (geometries.py:1035-1052). */
typedef struct dmw_myelin {
  float step_len[3];         /* float32 sqrt(6 D_c dt) per compartment */
  float two_kappa_over_D[4]; /* float32 2 kappa / max(D_leaving, 1e-30): inner membrane
                                left from 0, from 1; outer membrane left from 1, from 2 */
  float dt_over_T2[3];       /* float32 dt / T2_c (applied when run->has_t2) */
  float spin_weight[3];      /* water fractions (proton densities) */
  int64_t n_start[2];        /* walkers [0, n_start[0]) start intra-axonal,
                                [n_start[0], n_start[1]) in the sheath, the rest outside */
} dmw_myelin;

/* PackedMyelinatedCylinders (DMW_PACKED_MYELINATED): n_obj axons in the periodic
 * square of side `cell`, geometry centers (n_obj, 2) and radii = OUTER radii; the
 * compartment of a walker is (class, axon) with class 0 intra, 1 myelin, 2 extra
 * (an extra-axonal walker refers to axon 0, as the reference's k_cyl does).
 * run->n_sub sub-steps per sample (the surface rule when rho > 0), run->step_len
 * and run->inv_T2 unused.  The phase uses the unwrapped position; the signal is
 * sum_w s(w) e^{log_w} e^{i phi} / sum_w s(w), s = myelin_weight for walkers
 * starting in myelin, else 1 (core.py:548-556).  Walkers are placed in-kernel:
 * [start[j], start[j+1]) intra of axon j for j < n_obj, myelin of axon j - n_obj
 * for n_obj <= j < 2 n_obj, the rest [start[2 n_obj], n_walkers) extra-axonal. */
typedef struct dmw_packed_myelin {
  const float* inner_radii;      /* (n_obj) */
  const float* step_len;         /* (n_obj, 3) float32 sqrt(6 D dt_sub): intra, myelin, extra */
  const float* two_kappa_over_D; /* (n_obj, 4) float32 2 kappa / max(D_leaving, 1e-30): inner
                                    membrane left from intra, from myelin; outer membrane
                                    left from myelin, from extra */
  float dt_over_T2[3];           /* float32 dt_sub / T2 per class (applied when run->has_t2) */
  float rho_over_D;              /* wall relaxivity / D_ref (0 = none) */
  float myelin_weight;           /* spin weight of myelin water */
  const int64_t* start;          /* (2 n_obj + 1) ascending walker-index boundaries */
} dmw_packed_myelin;

/* Geometry and wall constants (float32 values rounded as the reference rounds
 * them, see make_step_fn physics.py:229-290 and the EPS/NUDGE definitions
 * geometries.py:503-504, 1694-1696). */
typedef struct dmw_geometry {
  int32_t kind;              /* enum dmw_geom_kind */
  int32_t wall;              /* enum dmw_wall */
  int32_t n_obj;             /* packed substrates: number of objects */
  float size;                /* Cylinder/Sphere radius, Box1D length */
  float cell;                /* packed substrates: periodic cell side L */
  float eps;                 /* EPS_detect */
  float nudge;               /* NUDGE */
  float rho_over_D;          /* surface relaxivity / D (0: none) */
  float kappa_over_D;        /* permeability / D */
  const float* centers;      /* (n_obj, 2) cylinders or (n_obj, 3) spheres */
  const float* radii;        /* (n_obj) */
  float axes[3];             /* Ellipsoid: float32 1/a^2, 1/b^2, 1/c^2 */
  float aux;                 /* Ellipsoid: 1 - NUDGE/min(a,b,c) (clamp); shells and the
                                myelinated cylinder: R_out.  size = R_in for those, the
                                slab length for the slab */
  const dmw_myelin* myelin;  /* DMW_MYELINATED_CYLINDER only, else NULL */
  const dmw_packed_myelin* packed_myelin; /* DMW_PACKED_MYELINATED only, else NULL */
} dmw_geometry;

/* Gradient waveform, Waveform.G of dmipy_sim/waveforms.py:39-90. */
typedef struct dmw_waveform {
  int32_t n_meas;
  int32_t n_t;
  const float* G;            /* (n_meas, n_t, 3) T/m */
  const float* chi;          /* (n_t) transverse-coherence flag, or NULL = all 1 */
  int32_t n_echo;            /* > 0: sample the ensemble at echo_idx (CPMG) */
  const int32_t* echo_idx;   /* (n_echo) sample indices, ascending */
} dmw_waveform;

/* Walker set and step constants of one walk. */
typedef struct dmw_run {
  int64_t n_walkers;
  uint64_t seed;             /* Philox4x32-10 key */
  int32_t n_sub;             /* physical sub-steps per waveform sample */
  float step_len;            /* sqrt(6 D dt_sub) */
  float gamma_dt_sub;        /* GAMMA * dt / n_sub */
  float dt_sub;
  float inv_T2;              /* 1/T2 (has_t2) */
  float inv_T1;              /* 1/T1 (has_t1) */
  int32_t has_t2;
  int32_t has_t1;
  int32_t kahan;             /* 1: compensated (Kahan) phase accumulation */
  int32_t meas_block;        /* measurements per kernel pass: 0 = automatic */
  const float* r0;           /* (n_walkers, 3) start positions, or NULL = sample in-kernel */
  const float* inj_normals;  /* (n_walkers, n_t*n_sub, 3) raw normal triples, or NULL = Philox */
  const float* inj_uniforms; /* (n_walkers, n_t*n_sub) Powles uniforms (permeable walls) */
  int64_t first_walker;      /* global index of this call's walker 0: the call walks walkers
                                [first_walker, first_walker + n_walkers) of a larger walk, with
                                their Philox streams and in-kernel start positions (a shard of a
                                walk split over calls or devices); 0 for a whole walk.
                                first_walker + n_walkers must not exceed 2^32. */
} dmw_run;

/* How a walk accumulated the phase (dmw_result.phase_form).  Automatic
 * measurement blocks (run->meas_block = 0) with at least 32 sub-steps per sample
 * and more than 16 measurements park a Kahan-compensated float32 phase in device
 * memory between samples; otherwise the phase stays in registers: compensated on
 * request (run->kahan), on the matrix cores (float16 hi/lo operands, float32
 * accumulation) for blocks of 32 or more, plain float32 below. */
enum dmw_phase_form {
  DMW_PHASE_REGISTERS = 0,
  DMW_PHASE_COMPENSATED = 1,
  DMW_PHASE_MATRIX_CORES = 2,
  DMW_PHASE_PARKED = 3
};

/* Outputs.  signal is required; every other array is optional (NULL). */
typedef struct dmw_result {
  double* signal;            /* (n_meas) or (n_echo, n_meas): Re <w e^{i phi}> */
  double* signal_im;         /* same shape: Im <w e^{i phi}> */
  float* final_pos;          /* (n_walkers, 3) unfolded, native frame */
  int32_t* comp_origin;      /* (n_walkers) compartment at t = 0 */
  int32_t* comp_final;       /* (n_walkers) compartment after the last sample */
  int32_t* comp_full;        /* (n_walkers, n_t) compartment after every sample */
  float* pos_full;           /* (n_walkers, n_t, 3) position after every sample */
  uint32_t* n_cross;         /* (n_walkers) membrane transmissions */
  float* log_w;              /* (n_walkers) accumulated log-weight */
  float* phi;                /* (n_walkers, n_meas) accumulated phase */
  double kernel_ms;          /* out: time of the walk kernel(s) on the device */
  uint64_t n_hits;           /* out: wall events (reflections + transmissions) */
  uint64_t n_cand;           /* out: candidate objects tested (packed substrates) */
  int32_t want_counters;     /* in: 1 = fill n_hits / n_cand */
  int32_t phase_form;        /* out: enum dmw_phase_form of the walk */
} dmw_result;

/* Number of visible HIP devices (0 when none). */
int dmw_device_count(void);

/* Select the device used by this thread's calls (default 0). */
int dmw_set_device(int device);

/* Architecture name of a device ("gfx950:sramecc+:xnack-", ...) into buf (len
 * bytes, NUL-terminated, truncated if short).  No counterpart in the reference:
 * keys the autotuner's per-architecture cache. */
int dmw_device_arch(int32_t device, char* buf, int32_t len);

/* Message of the last failure on the calling thread. */
const char* dmw_last_error(void);

/* One complete walk: upload, walk, reduce, download.
 * Replaces the fused engine of simulate() (core.py:572-762) and, with
 * wave->n_echo > 0, the echo sampling of simulate_cpmg() (core.py:937-969). */
int dmw_walk(const dmw_geometry* geom, const dmw_waveform* wave,
             const dmw_run* run, dmw_result* out);

/* dmw_walk without device memory allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and kernels as
 * dmw_walk; every device buffer of the call (staged waveform and geometry tables,
 * start positions, injected increments, per-wave partial sums, per-walker outputs)
 * comes from a static device array of 64 MiB per GPU, which the library holds from
 * its load until the process ends.  A walk needing more fails with DMW_EINVAL
 * before any launch.  The first call on a device creates the thread's stream and
 * looks the array up; calls on one device share the array and are serialised by
 * the library.  kernel_ms is not measured (0). */
int dmw_walk_noalloc(const dmw_geometry* geom, const dmw_waveform* wave,
                     const dmw_run* run, dmw_result* out);

/* A walk whose inputs stay resident in device memory across executions. */
typedef struct dmw_plan dmw_plan;

int dmw_plan_create(dmw_plan** plan, const dmw_geometry* geom,
                    const dmw_waveform* wave, const dmw_run* run);
/* Run the planned walk with a new seed; inputs are already in HBM.  A plan runs
 * on the device that was selected when it was created. */
int dmw_plan_execute(dmw_plan* plan, uint64_t seed, dmw_result* out);
int dmw_plan_destroy(dmw_plan* plan);

/* Resumable walk.  No counterpart in the reference, whose simulate() walks every
 * sample in one call: this is for walks checkpointed between launches (long walks
 * under a launch time limit, schedulers that interleave work) and for timing one
 * walk in slices.  Walks samples [t, t + n_samples) of the plan's waveform,
 * clipped at n_t, from the walkers the previous call left in device memory
 * (position and periodic cell, log-weight, crossings, parked phase: 48 bytes per
 * walker, twice, besides the phase); *t_done receives the samples walked so far.
 * The first call (after dmw_plan_create, a finished walk, dmw_plan_execute or
 * dmw_plan_rewind) starts a walk with `seed`; the following calls must pass the
 * same seed and ask for the same outputs (DMW_EINVAL otherwise).  Calls before
 * the last fill out->kernel_ms (that call's launch) and out->phase_form only; the
 * call that reaches n_t fills *out as dmw_plan_execute does (counters summed over
 * the calls) and leaves the plan ready for a new walk.  The result equals
 * dmw_plan_execute's bit for bit.  A walk split over several calls needs the
 * parked phase form, which a plan takes with meas_block 0, n_sub >= 32 and
 * n_meas > 16: other plans return DMW_EUNSUPPORTED unless one call walks every
 * sample.  Calls on one plan must not overlap. */
int dmw_plan_advance(dmw_plan* plan, uint64_t seed, int32_t n_samples,
                     dmw_result* out, int32_t* t_done);
/* Abandons the walk in progress, if any. */
int dmw_plan_rewind(dmw_plan* plan);

/* ---- Batched persistent launch (many substrate instances, one launch) ----
 * Has no single counterpart in the reference, where each instance is a separate
 * simulate() call (core.py:193) -- e.g. the serial loop of simulate_mixture
 * (core.py:821-858).  Instances may be any geometry kind with any wall mode the
 * geometry supports; each names one of n_sets waveform sets, which share n_meas
 * (<= 64) and n_t.  Instance i walks the Philox stream (key seed + seed_offset,
 * instance word `stream`, or i when stream < 0): with stream 0 and seed_offset k
 * it walks the trajectories dmw_walk walks with seed + k. */
typedef struct dmw_instance {
  dmw_geometry geom;         /* any kind; packed kinds bring centers/radii */
  int64_t n_walkers;
  int32_t n_sub;
  float step_len;
  float gamma_dt_sub;
  float dt_sub;
  float inv_T2;
  int32_t has_t2;
  int32_t wave_set;          /* index into the waveform sets of dmw_batch_create */
  int32_t stream;            /* Philox instance word; < 0: the instance's index */
  uint64_t seed_offset;      /* added to the execute seed for this instance */
} dmw_instance;

typedef struct dmw_batch dmw_batch;

typedef struct dmw_batch_stats {
  double kernel_ms;          /* walk kernel time */
  double total_ms;           /* whole graph launch */
  uint64_t walker_steps;     /* sum over instances of N * n_t * n_sub */
  int32_t n_units;           /* work units (64-walker tiles) */
  int32_t grid;              /* persistent workgroups */
} dmw_batch_stats;

/* Launch shape of the batched walk; zeros (or a NULL pointer) take the defaults.
 * The results do not depend on it: every tile of 64 walkers walks its own
 * Philox streams and writes its own partial row, reduced in a fixed order. */
typedef struct dmw_batch_tuning {
  int32_t waves_per_simd;    /* persistent wavefronts per SIMD, 1..8 (0 = 4, the kernel's occupancy budget) */
  int32_t walkers_per_wave;  /* walkers a wavefront takes per queue pop, walked as consecutive
                                64-walker tiles: 64, 128, 256 or 512 (0 = 64) */
  int32_t unroll;            /* sub-step loop unrolling: 1 or 2 (0 = 1) */
} dmw_batch_tuning;

/* sets: n_sets waveforms (chi of set 0 applies to all; echo fields ignored), all with
 * the same n_meas: 1..64, or up to 128 when every instance is a permeable packing
 * (those walk in a kernel of their own, with the wall screen and the phase parked
 * in device memory). */
int dmw_batch_create(dmw_batch** batch, int32_t n_inst, const dmw_instance* inst,
                     int32_t n_sets, const dmw_waveform* sets, const dmw_batch_tuning* tuning,
                     int32_t use_graph);
/* signals: (n_inst, n_meas) Re <w e^{i phi}> per instance. */
int dmw_batch_execute(dmw_batch* batch, uint64_t seed, double* signals,
                      dmw_batch_stats* stats);
int dmw_batch_destroy(dmw_batch* batch);

/* ---- Walk once, replay many waveforms ----
 * The reference's replay backend: simulate_trajectories (dmipy_sim/core.py:972-1529)
 * walks the spins once with no waveform and saves float16 positions after every
 * sample (plus, on request, the boundary log-weight at rho/D = 1 and the
 * compartment occupancy of permeable walls with a radius: Cylinder, Sphere,
 * PermeableSlab1D, PermeableShell); replay
 * (dmipy_sim/trajectories.py:360-641) applies any waveform afterwards as the
 * float64 contraction phi = GAMMA dt G_flat . traj^T and S = mean exp(log_w) cos(phi).
 * Here the trajectories stay in device memory between the two halves. */
typedef struct dmw_traj dmw_traj;

/* Producer walk.  geom/run as for dmw_walk (run->n_sub, step_len from the
 * producer's sub-step rule; gamma_dt_sub, T2, T1, kahan, meas_block ignored);
 * channels: bit 0 = boundary log-weight, bit 1 = compartment occupancy;
 * lab_from_native: 3x3 row-major rotation applied to the stored positions of a
 * rotated geometry (NULL = store native-frame positions). */
int dmw_traj_create(dmw_traj** traj, const dmw_geometry* geom, const dmw_run* run,
                    int32_t n_t, int32_t channels, const float* lab_from_native);
/* Trajectories given by the caller: traj (n_walkers, n_t, 3) and dlog
 * (n_walkers, n_t) or NULL, float16 bit patterns. */
int dmw_traj_from_host(dmw_traj** traj, int64_t n_walkers, int32_t n_t, const uint16_t* traj_h,
                       const uint16_t* dlog_h);
/* Copy channels back as (n_walkers, n_t, 3) / (n_walkers, n_t) float16 (NULL = skip). */
int dmw_traj_download(dmw_traj* traj, uint16_t* traj_h, uint16_t* dlog_h, uint16_t* comp_h);
int dmw_traj_info(dmw_traj* traj, int64_t* n_walkers, int32_t* n_t, int32_t* channels,
                  double* walk_ms);
int dmw_traj_destroy(dmw_traj* traj);

typedef struct dmw_replay {
  int32_t n_meas;
  int32_t n_t;               /* must equal the trajectory set's n_t */
  const double* G;           /* (n_meas, n_t, 3) T/m on the trajectory time grid */
  double scale;              /* GAMMA * dt_traj */
  const double* chi;         /* (chi_rows, n_t) transverse gate on the trajectory grid */
  int32_t chi_rows;          /* 1 or n_meas */
  const double* log_w0;      /* (n_meas) walker-independent log-weight (T2, T1 terms) */
  double rho_over_D;         /* surface relaxivity / D; 0 = no surface term */
  int32_t meas_tiles;        /* 16-measurement tiles per pass: 1, 2, 4, 8 or 0 = automatic */
} dmw_replay;

/* signals (n_meas); optional phi and log_w (n_meas, n_walkers) float64. */
int dmw_replay_run(dmw_traj* traj, const dmw_replay* replay, double* signals, double* phi,
                   double* log_w, double* kernel_ms);

/* ---- Learned surrogate (no counterpart in the reference) ----
 * A GELU MLP (depth hidden layers of `width` units) with BF16 weights and FP32
 * accumulation on the matrix cores, trained by its own forward/backward/Adam
 * kernels on signals the walk produces; it maps substrate parameters plus a
 * waveform embedding to the signal.  Public parameter layout (float32, flat):
 * for each hidden layer W_l (K_l x width, row-major, K_0 = n_in) then b_l
 * (width); then w_out (width) and b_out (1).  Batches are multiples of 64. */
typedef struct dmw_mlp dmw_mlp;

int dmw_mlp_create(dmw_mlp** mlp, int32_t n_in, int32_t width, int32_t depth, int32_t max_batch);
int dmw_mlp_destroy(dmw_mlp* mlp);
int64_t dmw_mlp_num_params(dmw_mlp* mlp);
/* Set master weights (resets the Adam moments); get them back. */
int dmw_mlp_set_params(dmw_mlp* mlp, const float* params, int64_t n);
int dmw_mlp_get_params(dmw_mlp* mlp, float* params, int64_t n);
/* y (batch) = network(x (batch, n_in)). */
int dmw_mlp_forward(dmw_mlp* mlp, const float* x, int32_t batch, float* y);
/* Gradients of the mean squared error against targets t, in the public layout. */
int dmw_mlp_gradients(dmw_mlp* mlp, const float* x, const float* t, int32_t batch, float* grads,
                      float* loss);
/* Keep a dataset resident in HBM (x: (n, n_in), t: (n)). */
int dmw_mlp_set_data(dmw_mlp* mlp, const float* x, const float* t, int64_t n);
/* n_steps Adam steps on batches of resident rows indices[s*batch ... +batch). */
int dmw_mlp_train(dmw_mlp* mlp, int32_t batch, int32_t n_steps, const int32_t* indices,
                  const float* lr, float beta1, float beta2, float eps, float* losses,
                  double* train_ms);
/* Predictions for the first n resident rows, in batches of max_batch. */
int dmw_mlp_predict(dmw_mlp* mlp, int64_t n, float* y, double* kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* DMIPY_WALK_H */
