/*
 * lbm355.h - C ABI of the MI355X (gfx950) D3Q27 lattice-Boltzmann step that
 * replaces the lbm3d time step behind solver/sim_tcpc of buresjan/nonnewlbm.
 *
 * The reference has no C ABI: its hot path is reached through the sim_tcpc
 * process CLI (solver/sim_tcpc/sim_tcpc.cu:70-125,846-857) and, inside the
 * process, through C++ template policies and the State virtual hooks
 * (solver/lbm3d/state.h:125-211).  Each entry point below names the reference
 * interface it stands in for.  Plain C types only: pointers, sizes, scalars.
 *
 * Conventions
 *  - Grids are X*Y*Z cells, dense host arrays are x-fastest: i = x + X*(y + Y*z),
 *    the same order as the legacy-VTK CELL/POINT data the reference reads and
 *    writes (solver/sim_tcpc/sim_tcpc.cu:176-183, solver/lbm3d/state.hpp:815-818).
 *  - Direction index q = 9*(cx+1) + 3*(cy+1) + (cz+1), opposite(q) = 26-q
 *    (solver/lbm3d/defs.h:215-247).
 *  - Every function returns LBM355_OK (0) or a negative LBM355_E* code and never
 *    throws; lbm355_last_error() gives a thread-local message for the last failure.
 *  - A case handle is not thread-safe: drive it from one host thread.
 */
#ifndef LBM355_H
#define LBM355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM355_ABI_VERSION 1

/* status codes */
enum {
    LBM355_OK = 0,
    LBM355_EINVAL = -1,     /* bad argument or descriptor                       */
    LBM355_ENOMEM = -2,     /* host or device allocation failed / over budget    */
    LBM355_EDEVICE = -3,    /* HIP runtime error                                 */
    LBM355_ETERMINATE = -4, /* case setup found zero inlet/outlet area; the
                               reference sets lbm.terminate and exits with 2
                               (solver/sim_tcpc/sim_tcpc.cu:633-640,837)         */
    LBM355_ENOTFOUND = -5,  /* unknown id (pool)                                 */
    LBM355_EBUSY = -6       /* request still running (pool poll)                 */
};

/* collision operators: LBM_TYPE policies of the reference */
enum {
    LBM355_CUM = 0, /* LBM_CUM, solver/lbm3d/lbm_col_cum.h:7-455 (sim_tcpc default)  */
    LBM355_BGK = 1, /* LBM_BGK, solver/lbm3d/lbm_col_bgk.h:9-170                     */
    LBM355_MRT = 2, /* LBM_MRT ("MRT_LES"), solver/lbm3d/lbm_col_mrt.h:7-78          */
    LBM355_KBC = 3  /* LBM_KBC_N4, solver/lbm3d/lbm_col_kbc_n.h:1024-1259            */
};

/* population storage; arithmetic is FP64 for F64 and FP32 for the others */
enum {
    LBM355_F64 = 0,         /* TraitsDP (solver/lbm3d/defs.h:141)                */
    LBM355_F32 = 1,         /* TraitsSP, the sim_tcpc default (defs.h:140);
                               stored as f_q - w_q, so that the small
                               non-equilibrium part keeps its 24 bits           */
    LBM355_F16_OFFSET = 2,  /* (f_q - w_q) 2^e_q stored as IEEE half (w_q 2^e_q =
                               16/27), stochastically rounded, FP32 arithmetic  */
    LBM355_BF16_OFFSET = 3  /* f_q - w_q stored as bfloat16, stochastically
                               rounded, FP32 arithmetic; for precision studies
                               only: it misses the 1e-3 velocity bar (about 4e-3 on
                               the 512x384x256 junction after 40,000 steps), so
                               the pool and the sim_tcpc CLI refuse it          */
};

/* Carreau-Yasuda coupling */
enum {
    LBM355_CY_OFF = 0,      /* what a single-rank sim_tcpc run computes: the strain
                               pre-kernels never run, so S=0, the extra force is 0
                               and the flow is Newtonian at nu_ref
                               (solver/lbm3d/core.h:1159-1166,663-781)           */
    LBM355_CY_NEQ = 1,      /* local nu(gamma) with gamma from the non-equilibrium
                               second moment of each cell, folded into omega     */
    LBM355_CY_FD_FORCE = 2  /* the reference multi-rank algorithm: velocity
                               pre-pass, finite-difference strain, body force
                               2(nu-nu_ref) div S (core.h:293-384,407-576,663-781) */
};

/* inflow profile on the IVC/SVC planes */
enum {
    LBM355_PROFILE_PLUG = 0,     /* u = flux/area on every open inlet cell
                                    (solver/sim_tcpc/sim_tcpc.cu:295-309)        */
    LBM355_PROFILE_PARABOLIC = 1 /* u = u_max (1 - r^2/R^2), r from the inlet axis */
};

/* Case descriptor.  Everything the reference reads from argv, the material
 * table and the geometry VTK for one run (sim_tcpc.cu:788-836), in lattice units
 * except dl.  Zero-initialise and fill; unknown future fields stay zero. */
typedef struct lbm355_case_desc {
    int32_t nx, ny, nz;          /* cell counts (VTK dims - 1)                     */
    const int32_t* geometry;     /* nx*ny*nz input labels, value 1 = wall
                                    (readGeometry, sim_tcpc.cu:134-185)            */
    double dl;                   /* physical spacing [m]                           */
    double dt;                   /* physical time step [s] (sim_tcpc.cu:800)       */
    int32_t collision;           /* LBM355_CUM ...                                 */
    int32_t storage;             /* LBM355_F64 ...                                 */
    double nu_ref;               /* lattice viscosity (--lbm-viscosity)            */
    /* Carreau-Yasuda in lattice units (sim_tcpc.cu:801-806); cy_enabled=0 for a
       Newtonian material */
    int32_t cy_enabled;
    int32_t cy_model;            /* LBM355_CY_*                                    */
    double cy_nu0, cy_lambda, cy_a, cy_n;
    /* inflow: lattice volume fluxes (sim_tcpc.cu:813-814); the IVC (label 3, x=0)
       moves +x, the SVC (label 4, x=X-1) moves -x (sim_tcpc.cu:657-663) */
    double ivc_flux, svc_flux;
    int32_t profile;             /* LBM355_PROFILE_*                               */
    double ivc_umax, svc_umax;   /* parabolic peak velocities (lattice units)      */
    double ivc_center[2], svc_center[2]; /* (y,z) axis of each inlet; <0 = centroid */
    double ivc_radius, svc_radius;       /* parabolic radius; <=0 = from the cells  */
    /* pulsatile factor u(t) = u (1 + amp sin(2 pi t / period)), t in steps */
    double pulse_amp, pulse_period;
    int32_t stats_every;         /* accumulate mean/rms every k steps (1 = every
                                    step, as sim_tcpc.cu:347-353); 0 = off        */
    /* boundary kind of the faces x-, x+, y-, y+, z-, z+.  All zero = the sim_tcpc
       layout (x-: IVC inflow label 3, x+: SVC inflow label 4, y-: outlet label 5,
       y+: outlet label 6, z: wall; sim_tcpc.cu:524-529).  Otherwise every face is
       an LBM355_FACE_* value; an outflow on x+ gets label 7, on x- label 8.
       Faces are labelled x, then y, then z (later faces win on edges), and every
       non-wall face gets the open-boundary extrusion of sim_tcpc.cu:549-586. */
    int32_t face_bc[6];
    int32_t reserved[1];
} lbm355_case_desc;

/* face boundary kinds for lbm355_case_desc.face_bc */
enum {
    LBM355_FACE_TCPC = 0,    /* use the sim_tcpc layout (all six must be 0)        */
    LBM355_FACE_WALL = 1,
    LBM355_FACE_INFLOW = 2,  /* x faces only: f_eq(rho from the next plane, u_in)  */
    LBM355_FACE_OUTFLOW = 3  /* x or y faces: f_eq(1, u from the next plane)        */
};

/* geometry / footprint report of a created case */
typedef struct lbm355_case_info {
    int64_t fluid;       /* label 0 cells                                          */
    int64_t wall;        /* wall cells (mask or z-plane)                           */
    int64_t area_ivc, area_svc, area_out5, area_out6;  /* open BC-plane cells       */
    int64_t input_label3, input_label4, input_label5, input_label6; /* VTK counts   */
    int64_t active_cells;   /* non-wall cells updated per step                     */
    int64_t bricks_active;  /* 64-cell bricks holding an active cell               */
    int64_t bricks_stored;  /* bricks with population storage                      */
    int64_t device_bytes;   /* device memory held by the case                      */
    int64_t steps;          /* time steps done                                     */
    int32_t terminate;      /* 1 if a zero-area inlet/outlet was found             */
    int32_t reserved[5];
} lbm355_case_info;

/* dense field ids for lbm355_case_get_field (double, nx*ny*nz, x-fastest) */
enum {
    LBM355_FIELD_RHO = 0, LBM355_FIELD_VX = 1, LBM355_FIELD_VY = 2, LBM355_FIELD_VZ = 3,
    /* statistics sums (MacroTcpc e_vm_*, sim_tcpc.cu:347-353): sum rho*u,
       sum (rho*u)^2 per component */
    LBM355_FIELD_SUM_RUX = 4, LBM355_FIELD_SUM_RUY = 5, LBM355_FIELD_SUM_RUZ = 6,
    LBM355_FIELD_SUM_RRUXX = 7, LBM355_FIELD_SUM_RRUYY = 8, LBM355_FIELD_SUM_RRUZZ = 9,
    /* strain tensor S11,S12,S13,S22,S32,S33 (lattice units) from the last
       lbm355_case_compute_strain */
    LBM355_FIELD_S11 = 10, LBM355_FIELD_S12 = 11, LBM355_FIELD_S13 = 12,
    LBM355_FIELD_S22 = 13, LBM355_FIELD_S32 = 14, LBM355_FIELD_S33 = 15,
    LBM355_FIELD_LABEL = 16,     /* map label after setup (VTK "wall")            */
    LBM355_FIELD_WALLMASK = 17,  /* wall mask after the open-boundary extrusion   */
    LBM355_FIELD_FX = 18, LBM355_FIELD_FY = 19, LBM355_FIELD_FZ = 20, /* body force */
    LBM355_FIELD_COUNT = 21
};

/* --- library --------------------------------------------------------------- */
const char* lbm355_version(void);
const char* lbm355_last_error(void);
/* number of visible HIP devices (0 on a host without a GPU) */
int lbm355_device_count(int* count);

/* --- one case: replaces StateTcpc + LBM + execute (sim_tcpc.cu:438-838,
       solver/lbm3d/core.h:1476-1628) ----------------------------------------- */
typedef struct lbm355_case lbm355_case;

/* Reads the descriptor, applies the sim_tcpc label setup (setupBoundaries +
 * projectWall, sim_tcpc.cu:494-645, lbm.hpp:145-155), builds the sparse brick
 * index, uploads f = f_eq(1,0) (resetLattice, sim_tcpc.cu:647-655) and the
 * initial statistics term (LBMKernelInit, core.h:839-863).  On a zero-area
 * inlet/outlet the case is still created, info.terminate = 1 and the call
 * returns LBM355_ETERMINATE (the reference runs and exits with code 2). */
int lbm355_case_create(const lbm355_case_desc* desc, lbm355_case** out);
int lbm355_case_destroy(lbm355_case* c);
int lbm355_case_info_get(const lbm355_case* c, lbm355_case_info* info);

/* per-step inflow hook (updateKernelVelocities, sim_tcpc.cu:657-663): lattice
   volume fluxes of the IVC and SVC; parabolic profiles scale u_max by the
   ratio to the fluxes given at create */
int lbm355_case_set_inflow(lbm355_case* c, double ivc_flux, double svc_flux);

/* n time steps (SimUpdate, core.h:1029-1340, n times); asynchronous on the
   case's stream; ends with the statistics and macro fields of the last step */
int lbm355_case_step(lbm355_case* c, int64_t n);
int lbm355_case_sync(lbm355_case* c);
/* launch shape of the case's fused step: 64-cell bricks (wavefronts) per
   workgroup, 1, 2, 4 or 8 (default 4).  No counterpart in the reference, whose
   launch shape is the y-block size (--block-size, sim_tcpc.cu:36). */
int lbm355_case_set_shape(lbm355_case* c, int bricks_per_group);
/* cache policy of the even step's population reads and writes: 1 streams them
   with the non-temporal hint, 0 uses the default policy, -1 (default) picks 1
   when the case holds more than 448 MiB of device memory (larger than the
   Infinity Cache keeps between steps).  Results do not depend on it.  No
   counterpart in the reference. */
int lbm355_case_set_nontemporal(lbm355_case* c, int mode);
/* at most waves_per_cu wavefronts of the fused step per CU (4..32), through an
   LDS allocation the kernel does not use; 0 (default) leaves occupancy to the
   register count.  Fewer resident wavefronts keep the odd step's neighbour
   lines in L2 longer on cases larger than the Infinity Cache.  Results do not
   depend on it.  No counterpart in the reference. */
int lbm355_case_set_occupancy(lbm355_case* c, int waves_per_cu);

/* dense double copy of one field; bytes must be >= nx*ny*nz*8 */
int lbm355_case_get_field(lbm355_case* c, int field, double* host, size_t bytes);
/* populations after the last step in the reference's A/B sense: out[q][cell] =
   post-collision f_q of each active cell (df_out after SimUpdate); wall cells
   next to the flow carry the populations in flight through full-way
   bounce-back, other wall cells are 0.  After an odd step the outward
   directions of BC-plane cells read 0: they would leave the grid, nobody reads
   them, and the step does not store them.  bytes >= 27*nx*ny*nz*8 */
int lbm355_case_get_populations(lbm355_case* c, double* host, size_t bytes);

/* finite-difference strain S of the current velocity field with the reference
   stencil (cudaLBMKernelStress, core.h:407-576) into the S fields */
int lbm355_case_compute_strain(lbm355_case* c);

/* signed lattice flux through the four BC planes {IVC, SVC, label5, label6}:
   sum of u_axis * sign over the plane cells, sign as compute_flows.py:96-104
   (inlets reported positive inflow, compute_flows.py:133-137); mean=1 uses
   the accumulated mean momentum instead of the last step */
int lbm355_case_flux(lbm355_case* c, int mean, double out[4]);

/* zero the statistics sums (statReset, sim_tcpc.cu:665-682) */
int lbm355_case_reset_stats(lbm355_case* c);

/* reference-layout 3D output: results_<id>/vtk3D/rank000_data_<n>.vtk with the
   13 arrays of StateTcpc::outputData (sim_tcpc.cu:716-784) through the legacy
   binary writer (state.hpp:782-851); phys_* describe the material */
typedef struct lbm355_output_desc {
    double phys_density;   /* kg/m^3                                              */
    double phys_nu_inf;    /* m^2/s (physViscosity)                               */
    double time;           /* physical time written to FIELD TIME                 */
    int32_t cycle;         /* VTK counter                                         */
    int32_t stat_steps;    /* iterations - prev_iter (sim_tcpc.cu:720)            */
    int32_t material_cy;   /* 1: the material is Carreau-Yasuda, so nu/mu/wss use
                              the case's cy_* parameters even when the coupling
                              is off (localNuLbm reads the material,
                              sim_tcpc.cu:697-707); 0: follow cy_enabled        */
    int32_t reserved[5];
} lbm355_output_desc;
int lbm355_case_write_vtk(lbm355_case* c, const char* path, const lbm355_output_desc* o);

/* --- checkpoint / restart ------------------------------------------------
 * The reference's current_state layout (saveAndLoadState, state.hpp:1552-1640):
 * <dir>/config_rank000 (iterations, X, Y, Z, local X, Y, Z, final time),
 * cnt_<k>_rank000 for the 11 output counters (count, period; state.h:72),
 * df_0_rank000 and df_1_rank000 (populations in the A/B sense, both the same;
 * the next step reads df_(iterations % 2)), map_rank000 and macro_rank000.
 * Arrays are raw binary in the index order of the reference's GPU build
 * (defs.h:57-60): df[x][q][z][y] and macro[x][id][z][y] in the real type of the
 * storage (float; double for LBM355_F64), map[x][z][y] as int16; macro ids are
 * the 20 fields of MacroTcpc (sim_tcpc.cu:328-335).  lbm355_rank000 adds what
 * the reference does not keep (the statistics reset step and sample count).
 * A restart needs a case created from the same descriptor; run N steps, save,
 * load, run M equals N + M steps. */
typedef struct lbm355_state_desc {
    double phys_final_time;        /* config line 8                                 */
    int32_t counter_count[11];     /* STAT_RESET, STAT2_RESET, PRINT, VTK1D, VTK2D,
                                      VTK3D, PROBE1, PROBE2, PROBE3, SAVESTATE,
                                      VTK3DCUT                                      */
    double counter_period[11];
    int64_t stat_prev_iter;        /* step of the last statistics reset              */
    int64_t iterations;            /* set by load                                    */
    int32_t reserved[4];
} lbm355_state_desc;
int lbm355_case_save_state(lbm355_case* c, const char* dir, const lbm355_state_desc* s);
int lbm355_case_load_state(lbm355_case* c, const char* dir, lbm355_state_desc* s);

/* --- batches: many cases stepped by one launch per step ------------------ */
/* No counterpart in the reference (it runs one case per process,
   scripts/slurm_run_cases.py:104-198).  A batch steps cases that share
   collision and storage in one fused launch over the concatenated brick
   index, kept with the members' constants in device memory; graph=1 replays
   one captured hipGraph of 64 steps per group.  A member keeps its own step
   count and can be stepped alone between batch calls (the batch orders
   itself after the work queued on the member's stream, and the member's
   later work after the batch's). */
typedef struct lbm355_batch lbm355_batch;
int lbm355_batch_create(lbm355_case* const* cases, int n, int graph, lbm355_batch** out);
int lbm355_batch_step(lbm355_batch* b, int64_t n);
int lbm355_batch_destroy(lbm355_batch* b);
/* membership changes between batch calls: the device brick list and member
   constants are rebuilt and re-uploaded; the captured graph stays valid
   unless the largest group outgrows its launch grid.  EINVAL for a case of
   another collision/storage type, an fd-force case or a member already in;
   ENOTFOUND for removing a non-member.  A batch may become empty. */
int lbm355_batch_add(lbm355_batch* b, lbm355_case* c);
int lbm355_batch_remove(lbm355_batch* b, lbm355_case* c);
int lbm355_batch_size(const lbm355_batch* b, int* members, int64_t* bricks);
/* launch shape: 64-cell bricks (wavefronts) per workgroup, 1, 2, 4 or 8 */
int lbm355_batch_set_shape(lbm355_batch* b, int bricks_per_group);
int lbm355_batch_set_occupancy(lbm355_batch* b, int waves_per_cu);
/* bricks per wavefront: 1 (default) launches one wavefront per brick; k > 1
   launches the persistent form, a grid of about 1/k of that whose wavefronts
   loop over the group's bricks */
int lbm355_batch_set_per_wave(lbm355_batch* b, int per_wave);
/* out = {bricks per workgroup, non-temporal (-1 auto), wavefront cap, group
   MiB (0 = one group), bricks per wavefront, groups, membership uploads,
   graph captures} */
int lbm355_batch_config(const lbm355_batch* b, int out[8]);
/* members are stepped in groups of consecutive cases holding at most `bytes`
   of device memory (a larger case is a group of its own; 0 = one group); each
   group runs a whole chunk of steps before the next, so a group that fits the
   256 MB Infinity Cache stays resident.  Default 256 MiB.  Results do not
   depend on it.  No counterpart in the reference. */
int lbm355_batch_set_group_bytes(lbm355_batch* b, int64_t bytes);
/* times every launch shape (1, 2, 4, 8 bricks per workgroup), with and
   without the non-temporal even step, and with no cap, 16 or 12 wavefronts per
   CU, on real steps: 64 warm-up steps, then per configuration 2 warm-up steps
   (plus 64 that capture the graph when the window reaches 64 steps) and a
   timed window W of 16 x stats_every steps (a whole number of statistics
   groups; 8 without statistics, at most 128); then 2 and 4 bricks per
   wavefront, then one group, 256 MiB and 128 MiB groups on windows of 2W.
   All of them count as simulation steps of every member.  Keeps the fastest;
   *best = bricks_per_group | nontemporal << 8 | waves_per_cu << 16 |
           (group bytes >> 24) << 24 (lbm355_batch_config gives the rest) */
int lbm355_batch_autotune(lbm355_batch* b, int* best);
/* the same for one case without grouping; each configuration is timed over W
   back-to-back steps (2 warm-up steps, then lbm355_case_time_kernel of 2W),
   then the 4 fastest again, in three interleaved passes timing 4W steps
   (2 + 8W each), each keeping its fastest pass (64 + 24 (2 + 2W) + 12 (2 + 8W)
   steps in all) */
int lbm355_case_autotune(lbm355_case* c, int* best);

/* --- serving pool ----------------------------------------------------------
 * No counterpart in the reference, which runs one case per process and queues
 * cases through Slurm (scripts/slurm_run_cases.py:104-198).  A pool owns one
 * worker thread that drives the current device: requests wait in FIFO order
 * until their footprint fits the device-memory budget; admitted cases stay
 * resident, and each tick steps the ones with the least remaining work (less
 * their age in seconds, so that none waits for ever) until they hold enough
 * cells to fill the GPU, those of one collision and storage type by one fused
 * launch per step over their concatenated brick lists, all types within the
 * same tick; finished cases leave between ticks and free their memory for
 * waiting requests (continuous batching).  Pool calls are thread-safe. */
typedef struct lbm355_pool lbm355_pool;
enum {
    LBM355_REQ_QUEUED = 0,
    LBM355_REQ_RUNNING = 1,
    LBM355_REQ_DONE = 2,
    LBM355_REQ_FAILED = 3
};
typedef struct lbm355_request_status {
    int32_t state;              /* LBM355_REQ_*                                      */
    int32_t error;              /* status code of a failed request (ETERMINATE for a
                                   zero-area case, ENOMEM over budget, ...)           */
    int64_t steps_done, steps;
    int64_t active_cells, device_bytes;
    double t_submit, t_start;   /* seconds on the pool clock (0 = pool creation)     */
    double t_first_report;      /* first plane-flux report (every report_every steps) */
    double t_done;
    double flux[4];             /* last reported fluxes, as lbm355_case_flux(mean=0)  */
    double mean_flux[4];        /* at completion, lbm355_case_flux(mean=1)            */
    char message[128];          /* why a request failed                               */
} lbm355_request_status;
typedef struct lbm355_pool_info {
    int64_t submitted, queued, running, done, failed;
    int64_t budget_bytes, used_bytes;
    int64_t cell_updates;       /* sum of active cells x steps over all requests      */
    int64_t ticks;
    double busy_seconds;        /* time spent stepping                               */
    double clock;               /* pool clock now                                    */
} lbm355_pool_info;
/* budget_bytes <= 0: 90 % of the free device memory; steps_per_tick <= 0: 64;
   graph = 1 replays captured hipGraphs inside a batch */
int lbm355_pool_create(int64_t budget_bytes, int steps_per_tick, int graph, lbm355_pool** out);
/* copies the descriptor and geometry; report_every > 0 reports plane fluxes
   every so many steps; keep = 1 keeps the finished case for lbm355_pool_take.
   LBM355_EINVAL for LBM355_BF16_OFFSET storage (not admitted for serving) */
int lbm355_pool_submit(lbm355_pool* p, const lbm355_case_desc* desc, int64_t steps, int64_t report_every,
                       int keep, int64_t* id);
/* LBM355_OK when the request is done or failed, LBM355_EBUSY while queued or running */
int lbm355_pool_poll(lbm355_pool* p, int64_t id, lbm355_request_status* st);
/* timeout_s < 0 waits without limit */
int lbm355_pool_wait(lbm355_pool* p, int64_t id, double timeout_s, lbm355_request_status* st);
int lbm355_pool_drain(lbm355_pool* p, double timeout_s);
/* hands a finished keep=1 case to the caller, who destroys it */
int lbm355_pool_take(lbm355_pool* p, int64_t id, lbm355_case** out);
int lbm355_pool_stats(lbm355_pool* p, lbm355_pool_info* st);
/* stops the worker after its current tick; unfinished requests are dropped */
int lbm355_pool_destroy(lbm355_pool* p);

/* error of a case against a reference case made from the same geometry (for
   example the same junction stored in FP16 against FP64), from the macro
   output of the last step of each, reduced on the device:
   out = {sum |u - u_ref|^2, sum |u_ref|^2, sum (rho - rho_ref)^2,
          sum (rho_ref - 1)^2} over the non-wall cells, so that
   sqrt(out[0] / out[1]) is the relative L2 error of velocity.  No counterpart
   in the reference (the precision pattern states each field's error against
   an FP64 run).  Allocates a partial-sum buffer of 4 ceil(active bricks / 4)
   + 4 doubles for the call.  EINVAL when grid, labels or brick index differ
   or either case has not stepped. */
int lbm355_case_compare(lbm355_case* a, lbm355_case* ref, double out[4]);
/* the same result, bit for bit, without allocating: the partial sums go to a
   static device array of the library (one per GPU, 4 x 16384 + 4 doubles =
   512 KiB, held until the process ends), so it calls no hipMalloc, hipFree
   or allocating routine.  No counterpart in the reference; for callers that
   cannot allocate (code under graph capture of its own surroundings, tight
   serving loops, harnesses that forbid device allocation).  Limit: cases of
   at most 65,536 active bricks (4,194,304 cells), EINVAL beyond with the
   counts in the message.  Calls that overlap on one GPU share the array and
   race; it needs nothing cleared between calls. */
int lbm355_case_compare_noalloc(lbm355_case* a, lbm355_case* ref, double out[4]);

/* bytes of device memory a descriptor would need (admission control) */
int lbm355_case_footprint(const lbm355_case_desc* desc, int64_t* bytes);
/* the label setup of lbm355_case_create without a device: labels (nx*ny*nz,
   x-fastest, may be NULL) receive the final map (0 fluid, 1 wall, 3/4 x-planes,
   5/6 y-planes, 7/8 x+/x- outflows) and info the counts, areas, brick counts
   and footprint.  Returns LBM355_ETERMINATE on a zero inlet/outlet area, as
   setupBoundaries does (sim_tcpc.cu:632-639). */
int lbm355_case_labels(const lbm355_case_desc* desc, int32_t* labels, lbm355_case_info* info);

/* --- raw kernel timing (measurement only) ---------------------------------- */
/* n more time steps, timed with HIP events on the case's stream: avg_step_ms =
   whole step (the fused kernel and its launch gaps) from one event pair around the
   first (n + 1) / 2 steps, launched back to back; avg_kernel_ms = the fused
   stream-collide launch alone, from an event pair around each of the next
   min(n / 2, 512) launches (the rest of the n steps run untimed).  No counterpart in
   the reference, which only logs wall-clock MLUPS (core.h:1440-1469). */
int lbm355_case_time_kernel(lbm355_case* c, int64_t n, double* avg_step_ms, double* avg_kernel_ms);

/* device-to-device copy bandwidth in GB/s (read + write bytes / time) of a
   plain vectorised copy kernel over `bytes` bytes, best of `iters` runs:
   the on-box reference for the roofline fractions */
int lbm355_copy_bandwidth(int64_t bytes, int iters, double* gbps);

#ifdef __cplusplus
}
#endif
#endif /* LBM355_H */
