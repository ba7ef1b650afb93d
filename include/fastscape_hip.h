/*
 * fastscape_hip.h -- C ABI of the MI355X implementation of fastscapelib-fortran's per-step
 * landscape update (FastScape_Execute_Step and the state it reads and writes).
 *
 * A context (fsx_ctx) holds a batch of independent landscapes ("instances") resident in HBM; a
 * single-landscape context is the drop-in for the reference's module state
 * (src/FastScape_ctx.f90:9-36).  Instance ids run from 0 to ninst-1.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative FSX_E* code; fsx_last_error() returns a
 *    thread-local message for the last failure.  (The reference STOPs the process instead.)
 *  - Node arrays are nx*ny doubles with x fastest, as the reference's h(nx*ny).
 *  - `on_device` != 0 means the pointer is HIP device memory (e.g. a torch tensor's data_ptr());
 *    0 means host memory.  Pointers are not retained after the call returns.
 *  - Node indices returned by copy-outs are 0-based (the reference's are 1-based).
 *  - `stream` may be NULL (the context's own stream).
 */
#ifndef FASTSCAPE_HIP_H
#define FASTSCAPE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fsx_ctx fsx_ctx;

enum {
  FSX_OK = 0,
  FSX_EINVAL = -1,   /* bad argument, size or call order                              */
  FSX_EHIP = -2,     /* HIP runtime error                                              */
  FSX_ESTEP = -3,    /* a step failed on the device (see fsx_last_error for the cause) */
  FSX_ENOMEM = -4,
};

/* field ids of fsx_set_field / fsx_copy_field */
enum {
  FSX_H = 0,          /* topography h                     Set_H / Copy_H            */
  FSX_B = 1,          /* basement b                       Set_Basement / Copy_Basement */
  FSX_U = 2,          /* uplift rate u                    Set_U                     */
  FSX_PRECIP = 3,     /* precipitation                    Set_Precip                */
  FSX_ETOT = 4,       /* total erosion                    Copy_Total_Erosion        */
  FSX_ERATE = 5,      /* erosion rate                     Copy_Erosion_Rate         */
  FSX_AREA = 6,       /* drainage area                    Copy_Drainage_Area        */
  FSX_LAKE_DEPTH = 7, /* lake depth                       Copy_Lake_Depth           */
  FSX_SEDFLUX = 8,    /* sediment flux of the last step (Sedflux, no reference copy-out) */
  FSX_LENGTH = 9,     /* distance to receiver (length, internal)                      */
  FSX_HWATER = 10,    /* lake surface (hwater, internal)                              */
  FSX_KF = 11,        /* kf (write through fsx_set_erosional_parameters)             */
  FSX_KD = 12,        /* kd (idem)                                                    */
  FSX_FMIX = 13,      /* silt fraction of the marine transport (Fmix)   Init_F / Copy_F */
};
/* derived field ids of fsx_copy_derived */
enum {
  FSX_SLOPE = 0,     /* slope in degrees (slope, src/TerrainDerivatives.f90:3-36)     Copy_Slope     */
  FSX_CURVATURE = 1, /* curvature (curvature, src/TerrainDerivatives.f90:40-74)        Copy_Curvature */
  FSX_CHI = 2,       /* chi (CopyChi, src/FastScape_ctx.f90:226-247)                   Copy_Chi       */
  FSX_CATCHMENT = 3, /* catchment label (CopyCatchment, :297-307): (outlet + 1) / nn of the node's tree
                        before closed-basin rerouting, where the reference takes a random catch0 of
                        that outlet; nodes share a value exactly as in the reference  Copy_Catchment */
};
/* integer field ids of fsx_copy_ifield */
enum {
  FSX_REC = 0,    /* receiver after closed-basin rerouting, 0-based (rec)               */
  FSX_NDON = 1,   /* donor count (ndon)                                                  */
  FSX_DON = 2,    /* donors, 8 slots per node ascending, unused = -1 (don(8,nn))         */
  FSX_ORDER = 3,  /* base-level-first topological order of the nodes (replaces stack)    */
  FSX_MREC = 4,   /* multiple-flow receivers, 8 slots per node in scan order, unused = -1 (mrec) */
  FSX_MNREC = 5,  /* multiple-flow receiver count (mnrec)                                */
  FSX_MSTACK = 6, /* multiple-flow order, donors before receivers (replaces mstack)      */
};

/* Thread-local text of the last error. */
const char *fsx_last_error(void);
/* Library version string. */
const char *fsx_version(void);

/* FastScape_Init + Set_NX_NY + Setup (src/FastScape_api.f90:146-170,475-487; SetUp,
 * src/FastScape_ctx.f90:57-113) for ninst landscapes of sizes nx[k] x ny[k] (each >= 3).
 * Defaults as SetUp: u=0, etot=0, precip=1, BC 1111, no process active; h = 0 (not random). */
int fsx_create(int ninst, const int *nx, const int *ny, fsx_ctx **out);
/* FastScape_Destroy (src/FastScape_api.f90:174-184) */
int fsx_destroy(fsx_ctx *ctx);

/* Landscape pools for serving.  The reference holds one landscape per process and allocates it in
 * SetUp / frees it in Destroy (src/FastScape_ctx.f90:57-113,117-154); a pool generalises that
 * lifecycle to landscapes that come and go while others step.  fsx_create_pool reserves an arena of
 * capacity_nodes nodes and nslots landscape slots in HBM; fsx_add_instance places an nx x ny landscape
 * in the first free node range that fits (SetUp defaults as fsx_create; FSX_ENOMEM when no slot or
 * range is free) and returns its id; fsx_remove_instance frees it.  Ids of other landscapes do not
 * change.  fsx_pool_info reports free nodes, the largest free range and free slots. */
int fsx_create_pool(long long capacity_nodes, int nslots, fsx_ctx **out);
int fsx_add_instance(fsx_ctx *ctx, int nx, int ny, int *inst);
int fsx_remove_instance(fsx_ctx *ctx, int inst);
int fsx_pool_info(fsx_ctx *ctx, long long *free_nodes, long long *largest_free, int *free_slots);
/* Free and total device memory in bytes (hipMemGetInfo), for sizing pools against free HBM; the
 * reference allocates host memory in SetUp without such a check (src/FastScape_ctx.f90:64-80). */
int fsx_device_memory(long long *free_bytes, long long *total_bytes);

/* FastScape_Set_XL_YL (src/FastScape_api.f90:491-503) */
int fsx_set_xl_yl(fsx_ctx *ctx, int inst, double xl, double yl);
/* FastScape_Set_DT (src/FastScape_api.f90:507-519) */
int fsx_set_dt(fsx_ctx *ctx, int inst, double dt);
/* FastScape_Set_BC (src/FastScape_api.f90:602-614; SetBC src/FastScape_ctx.f90:565-593) */
int fsx_set_bc(fsx_ctx *ctx, int inst, int ibc);
/* FastScape_Set_Erosional_Parameters (src/FastScape_api.f90:523-536; SetErosionalParam
 * src/FastScape_ctx.f90:423-450): p < -1.5 selects single flow direction, otherwise multiple flow
 * direction with weights slope**p (p < 0: exponent from the mean slope, src/FlowRouting.f90:232-243). */
int fsx_set_erosional_parameters(fsx_ctx *ctx, int inst, const double *kf, double kfsed, double m, double n,
                                 const double *kd, double kdsed, double g1, double g2, double p, int on_device);
/* FastScape_Set_Marine_Parameters (src/FastScape_api.f90:540-552; SetMarineParam
 * src/FastScape_ctx.f90:454-472): sea level, surface porosities and their e-folding depths of silt
 * and sand, silt ratio, thickness of the mixed layer, marine diffusivities of silt and sand.  Switches
 * on the marine transport after Diffusion in every step (Marine, src/Marine.f90:1-167), which needs
 * the erosional parameters (the step's flow routing); with it on, the stream-power law leaves nodes
 * below sea level alone (src/StreamPowerLaw.f90:105,144,305,342).  The silt fraction is field
 * FSX_FMIX (default 0.5). */
int fsx_set_marine_parameters(fsx_ctx *ctx, int inst, double sealevel, double poro1, double poro2, double zporo1,
                              double zporo2, double ratio, double layer, double kdsea1, double kdsea2);
/* FastScape_Strati (src/FastScape_api.f90:745-758; Activate_Strati src/FastScape_ctx.f90:781-805):
 * nreflector stratigraphic reflectors, one more becoming active every nstep / nreflector steps, output
 * every nfreq steps.  Every step then ends with run_Strati's reflector update
 * (src/FastScape_ctx.f90:809-830) on the GPU; the reference also writes its horizon files from
 * inside the step every nfreq steps (Strati, src/Strati.f90:1-82): here the caller does that between
 * steps -- after a step whose count is a multiple of nfreq, it advances the active reflector when the
 * count is also a multiple of nstep / nreflector (fsx_strati_set_active), reads the reflectors, the
 * erosion below them, their slopes and the distances to shore (fsx_strati_read), and writes the files
 * (fastscapelib_fortran.strati does so).  vex is the host writer's.  The reflectors start at zero
 * (the reference reads them before it first sets them). */
int fsx_strati(fsx_ctx *ctx, int inst, int nstep, int nreflector, int nfreq, double vex);
/* The active reflector (1-based; ireflector in run_Strati). */
int fsx_strati_set_active(fsx_ctx *ctx, int inst, int ireflector);
/* Stratigraphy read-outs of one landscape into host memory, nreflector x nn values (reflector-major):
 * FSX_STRATI_REFLECTORS, FSX_STRATI_ERODED (erosion below each reflector, Strati's field 10),
 * FSX_STRATI_SLOPES (slope of each reflector in degrees, src/TerrainDerivatives.f90:3-36), or
 * FSX_STRATI_SHORE (distance to shore of each reflector and then of the basement, (nreflector + 1) x nn,
 * distance_to_shore src/Strati.f90:86-127, along the last step's receivers). */
enum { FSX_STRATI_REFLECTORS = 0, FSX_STRATI_ERODED = 1, FSX_STRATI_SLOPES = 2, FSX_STRATI_SHORE = 3 };
int fsx_strati_read(fsx_ctx *ctx, int inst, int which, double *dst);
/* Newton tolerance of the n != 1 stream-power solve.  No reference setter: the reference
 * hard-codes 1e-3 (src/StreamPowerLaw.f90:344), which is the default here. */
int fsx_set_newton_tol(fsx_ctx *ctx, int inst, double tol);
/* FastScape_Init_H (src/FastScape_api.f90:267-279; InitH src/FastScape_ctx.f90:339-350): h, and b = h */
int fsx_init_h(fsx_ctx *ctx, int inst, const double *h, int on_device);
/* FastScape_Set_U / Set_H / Set_Basement / Set_Precip / Init_F (src/FastScape_api.f90:618-630,664-676,
 * 696-724,283-295): field one of FSX_U, FSX_H, FSX_B, FSX_PRECIP, FSX_FMIX.  Setting FSX_U switches
 * uplift on. */
int fsx_set_field(fsx_ctx *ctx, int inst, int field, const double *src, int on_device);
/* FastScape_Set_V (src/FastScape_api.f90:634-646; SetV src/FastScape_ctx.f90:616-632): the velocity
 * field (vx, vy: nn values each); switches on the advection of h, b and etot at the start of every
 * step (Advect, src/Advect.f90:1-103). */
int fsx_set_v(fsx_ctx *ctx, int inst, const double *vx, const double *vy, int on_device);
/* FastScape_Set_All_Layers (src/FastScape_api.f90:680-692): h += dh, b += dh */
int fsx_set_all_layers(fsx_ctx *ctx, int inst, const double *dh, int on_device);
/* FastScape_Reset_Cumulative_Erosion (src/FastScape_api.f90:650-660) */
int fsx_reset_cumulative_erosion(fsx_ctx *ctx, int inst);
/* Include (1) or exclude (0) an instance from fsx_execute_steps (all are active at creation). */
int fsx_set_active(fsx_ctx *ctx, int inst, int active);

/* FastScape_Execute_Step (src/FastScape_api.f90:201-263), nsteps times, on every active
 * instance: Advect, Uplift, then routing + LocalMinima + accumulation + stream power, then Diffusion,
 * then Marine, then the stratigraphic reflectors, each where switched on.
 * Asynchronous on `stream` unless `sync` != 0; device-side failures surface as FSX_ESTEP at the
 * next synchronous call. */
int fsx_execute_steps(fsx_ctx *ctx, int nsteps, void *stream, int sync);
/* Wait for queued steps and report device-side failures. */
int fsx_synchronize(fsx_ctx *ctx);
/* Without waiting: 1 when every step queued on the context's own stream has finished (its error words
   are then checked, as after a synchronous step), 0 while some still run, < 0 on an error.  Lets a
   serving loop keep several pools' steps in flight and attend to whichever finishes.  No counterpart
   in the reference (its steps are synchronous calls). */
int fsx_query(fsx_ctx *ctx);
/* Run each step as a replay of one captured HIP graph (1, the default) or as individual kernel
 * launches (0).  Results are the same; graphs are not used while phase profiling is on.  No
 * reference counterpart: an execution option of FastScape_Execute_Step's device implementation. */
int fsx_set_graphs(fsx_ctx *ctx, int on);

/* FastScape_Copy_H / Copy_Basement / Copy_Total_Erosion / Copy_Erosion_Rate /
 * Copy_Drainage_Area / Copy_Lake_Depth (src/FastScape_api.f90:299-471) and internal fields. */
int fsx_copy_field(fsx_ctx *ctx, int inst, int field, double *dst, int on_device);
/* Integer state after the last step: receivers, donors, order (see FSX_REC...). dst holds nn
 * ints (8*nn for FSX_DON and FSX_MREC).  FSX_MREC/FSX_MNREC/FSX_MSTACK need an instance that uses
 * multiple flow direction (p >= -1.5). */
int fsx_copy_ifield(fsx_ctx *ctx, int inst, int field, int *dst, int on_device);

/* FastScape_Copy_Slope / Copy_Curvature / Copy_Chi / Copy_Catchment (src/FastScape_api.f90:379-439),
 * computed on the device from the current state of one landscape into dst (nn doubles).  Chi and the
 * catchment follow the routing of the last step (FSX_EINVAL before the first step or fsx_route). */
int fsx_copy_derived(fsx_ctx *ctx, int inst, int which, double *dst, int on_device);
/* FastScape_Get_Fluxes (src/FastScape_api.f90:762-774; compute_fluxes, src/FastScape_ctx.f90:843-886):
 * out3 = tectonic, erosion and boundary flux of one landscape's current state, computed on the device
 * (FSX_EINVAL before the first step or fsx_route). */
int fsx_get_fluxes(fsx_ctx *ctx, int inst, double *out3);
/* The same for many landscapes in one launch (no reference counterpart: the reference holds one
 * landscape).  fsx_copy_derived_batch: dst has one pointer per landscape slot (nn doubles each, NULL to
 * skip); fsx_get_fluxes_batch: out has 3 doubles per slot, NaN for a slot without a routed landscape. */
int fsx_copy_derived_batch(fsx_ctx *ctx, int which, double *const *dst, int on_device);
int fsx_get_fluxes_batch(fsx_ctx *ctx, double *out);

/* FastScape_Get_Step (src/FastScape_api.f90:572-584) */
int fsx_get_step(fsx_ctx *ctx, int inst, int *step);
/* FastScape_Get_Sizes (src/FastScape_api.f90:556-568) */
int fsx_get_sizes(fsx_ctx *ctx, int inst, int *nx, int *ny);

/* Routing only, on the current h (no uplift, no erosion): find_receiver + find_donor +
 * find_stack + LocalMinima + lake fill + accumulation, as FlowRoutingSingleFlowDirection and
 * FlowAccumulationSingleFlowDirection (src/FlowRouting.f90:43-93,123-143), or their multiple-flow
 * counterparts (:3-39,97-119).  Copy the result with fsx_copy_ifield / fsx_copy_field. */
int fsx_route(fsx_ctx *ctx, void *stream, int sync);

/* loc_min_3_indexx (src/LocalMinima.f90:947-1046) on n keys: writes the 1-based permutation
 * (the reference's indx) into perm.  Host arrays; runs the device sort. */
int fsx_indexx(int n, const double *keys, int *perm);

/* fsx_indexx without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).  keys and
 * perm are DEVICE arrays of n entries (perm receives the 1-based permutation); the work is enqueued on
 * stream (0: the null stream) and, with sync != 0, waited for and checked.  The sort's working arrays
 * are a fixed area in the library's device memory, one per GPU, held until the process ends: n is
 * limited to FSX_INDEXX_NOALLOC_MAX (larger n returns FSX_EINVAL).  Calls that overlap on one GPU
 * share that area and race; serialize them (one stream, or wait between calls). */
#define FSX_INDEXX_NOALLOC_MAX 524288
int fsx_indexx_noalloc(int n, const double *keys, int *perm, void *stream, int sync);

/* Statistics of the last step of an instance: out[0]=nbasins, out[1]=nconn, out[2]=ntree,
 * out[3]=levels, out[4]=Gauss-Seidel iterations, out[5]=Newton iterations (total). */
int fsx_get_stats(fsx_ctx *ctx, int inst, long long *out6);

/* Phase timers, the counterpart of the reference's per-phase cpu_time accumulators
 * (timeUplift/timeSPL/timeDiffusion, src/FastScape_api.f90:207-257; printed by FastScape_Debug,
 * src/FastScape_ctx.f90:553-559).  When on, every step records a HIP event pair around each phase
 * on the stream it runs on.  fsx_get_timers fills up to n accumulated milliseconds and counts
 * (phase order: uplift, receivers,
 * donors, lm_labels, lm_basins, lm_walk, lm_edges, lm_sort, lm_mst, lm_orient, lm_correct, levels, area_prep, area_sweep, spl_prep, spl_fact, spl_sweep, spl_post,
 * diffusion, mfd_route, mfd_spl, advect, marine, strati) and returns the number of phases. */
int fsx_set_profiling(fsx_ctx *ctx, int on);
/* Time one phase (0..fsx_get_timers' count-1; -1: none) live while steps replay HIP graphs: the
 * phases before it and after it run as two graphs and the phase itself is launched directly between
 * an event pair on the step's stream.  Its time accumulates in fsx_get_timers. */
int fsx_time_phase(fsx_ctx *ctx, int phase);
int fsx_get_timers(fsx_ctx *ctx, double *ms, long long *counts, int n);
int fsx_reset_timers(fsx_ctx *ctx);

/* Launch-shape autotuner.  Times the routing phases (on the current landscapes, which it does not
 * change) under candidate launch shapes -- basin-walk blocks per landscape, the pointer-jumping grid
 * cap, Euler-tour blocks per landscape, edge-sort blocks per landscape, drainage-area sweep threads
 * per landscape, spanning-forest edge-scan workgroups per landscape -- and, when some landscape takes
 * the dataflow stream-power sweep (single flow direction, g = 0), that sweep's threads per workgroup and
 * nodes per workgroup (timed on solves of the routed state that leave h and the per-step counts alone), with reps timed
 * passes per candidate, and keeps the fastest of each.  No result depends on these shapes.  Routing
 * read-outs (receivers, donors, area) then hold the tuning passes' routing of the current elevation
 * until the next step.  Also the basin size above which the pass search runs node-parallel instead of
 * one thread walking the basin (walk_big), and, when some landscape routes by multiple flow directions,
 * the nodes per workgroup of its dataflow sweeps (mfd_split), the rows per 2-D tile of the receiver and
 * donor-mask stencil kernels (tile_h: 4, 8, 16 or 32; tiles are 64 columns wide), and the Newton iterations
 * a lane of the dataflow stream-power sweep runs per pass of its wave (flow_nk: 4, 8, 16 or 32).  chosen
 * (FSX_NTUNE ints, or NULL) receives walk_gx, jump_blocks, tour_gx, nr_small_gx, area_wg, bor_gx, flow_wg,
 * flow_split, walk_big, mfd_split, tile_h, flow_nk.  No reference counterpart: the reference has no launch
 * shapes. */
#define FSX_NTUNE 12
int fsx_autotune(fsx_ctx *ctx, int reps, int *chosen);
/* The launch shapes in use (FSX_NTUNE ints, same order as fsx_autotune's chosen). */
int fsx_get_tuning(fsx_ctx *ctx, int *out);
/* Set the launch shapes (FSX_NTUNE ints, same order; e.g. ones an earlier fsx_autotune chose for
   this workload).  Each must be >= 1; walk blocks <= 4096, jump grid <= 2^20, tour blocks <= 1024,
   sort blocks <= 4096; area_wg one of 256, 512, 1024; bor_gx in [1, 1024]; flow_wg one of 256, 512,
   1024; flow_split >= 0 (nodes per workgroup of the dataflow sweep, 0: one workgroup per landscape; at
   most 16 workgroups per landscape); walk_big >= 64; mfd_split >= 1024 (at most 32 workgroups per
   landscape); tile_h one of 4, 8, 16, 32; flow_nk in [1, 1024].  A rejected call changes nothing.
   No counterpart in the reference. */
int fsx_set_tuning(fsx_ctx *ctx, const int *in);

/* Device-resident memory of a context (bytes) and the bytes per node the layout costs. */
int fsx_memory(fsx_ctx *ctx, long long *bytes, long long *bytes_per_node);

#ifdef __cplusplus
}
#endif
#endif
