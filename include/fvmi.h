/*
 * fvmi.h -- C ABI of libfvmi.so, the MI355X (gfx950) flux-and-update engine.
 *
 * This library replaces, for the dimensionally split hot path of
 * FiniteVolumeGPU, the PyCUDA device runtime (GPUSimulators/CudaContext.py:45-272),
 * the pitched padded device arrays (GPUSimulators/Common.py:498-616, 741-779)
 * and the CUDA kernels launched through `prepared_async_call`
 * (GPUSimulators/KP07_dimsplit.py:111-126, HLL2.py, WAF.py, EE2D_KP07_dimsplit.py:115-266).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on failure;
 *     fvmi_last_error() then returns a message (thread local).  No C++
 *     exception crosses this boundary.
 *   - Streams and events are opaque `void*` (hipStream_t / hipEvent_t).
 *     A NULL stream is the device's null stream.
 *   - Field buffers are row-major float32 of (ny+4) rows x pitch bytes with
 *     the interior at rows 2..ny+1, columns 2..nx+1 (2 ghost cells per side),
 *     the layout of CudaArray2D (Common.py:502-531).  Pitches are in bytes, as
 *     the reference passes `strides[0]`.  Each field must be below 2 GiB,
 *     (ny+4)*pitch < 2^31 ((ny+2)*pitch for the one-ghost-cell schemes).
 *     The kernel that marches two columns per lane (Euler in fast
 *     arithmetic) moves them with 8-byte loads and stores: its field
 *     pointers and pitches must be multiples of 8 bytes (hipMalloc /
 *     fvmi_malloc_pitch give 256-byte alignment).
 *   - Boundary conditions are packed N<<24 | S<<16 | E<<8 | W with
 *     0 = Dirichlet, 2 = Periodic, 3 = Reflective (Simulator.py:44-92,
 *     common.h:109-130) and 4 = Outflow (zero-gradient ghost copy; an
 *     extension with no counterpart in the reference).  Neumann (1) is
 *     rejected, as Simulator.py:68-72 does.
 */
#ifndef FVMI_H
#define FVMI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVMI_OK 0
#define FVMI_ERR_INVALID (-1)
#define FVMI_ERR_HIP (-2)
#define FVMI_ERR_UNSUPPORTED (-3)

#define FVMI_MATH_FAST 0  /* hardware rcp/sqrt, contraction: within rel-L1 1e-5 of exact */
#define FVMI_MATH_EXACT 1 /* IEEE div/sqrt, no FMA: bitwise equal to an IEEE float32 evaluation */

#define FVMI_MEMCPY_HTOD 1
#define FVMI_MEMCPY_DTOH 2
#define FVMI_MEMCPY_DTOD 3

/* Launch shape of the substep kernels.  A NULL pointer, or a zero field, selects
 * the default.  Replaces the compile-time BLOCK_WIDTH/BLOCK_HEIGHT defines of
 * the reference kernels (KP07_dimsplit.py:80-89). */
typedef struct fvmi_launch {
    int tile_rows;       /* rows per wavefront segment; 0 (default) = balanced over the device */
    int waves_per_block; /* strips per workgroup, 1..4 (default 4) */
    int math;            /* FVMI_MATH_FAST (default) or FVMI_MATH_EXACT */
    int xcd_remap;       /* 1 (default): consecutive strips on one XCD; 2: off */
} fvmi_launch_t;

/* ---- errors, version ------------------------------------------------------ */
const char* fvmi_last_error(void);
int fvmi_version(void);

/* ---- device / context: CudaContext.__init__, synchronize (CudaContext.py:45-120, 262-263) */
int fvmi_device_count(int* count);
/* Selects `device` for the calling thread and initialises it; `flags` is reserved (0). */
int fvmi_ctx_create(int device, int flags, void** ctx);
int fvmi_ctx_destroy(void* ctx);
/* Makes ctx's device current for the calling thread (CudaContext's context push). */
int fvmi_ctx_make_current(void* ctx);
int fvmi_ctx_synchronize(void* ctx);
int fvmi_ctx_device(void* ctx, int* device);
int fvmi_device_name(int device, char* buf, int len);
/* Number of compute units (256 on MI355X); used to size launches. */
int fvmi_device_cu_count(int device, int* cus);
/* cuda.mem_get_info(), CudaContext.py:81 */
int fvmi_mem_get_info(size_t* free_bytes, size_t* total_bytes);

/* ---- streams and events: cuda.Stream (Simulator.py:164-165), cuda.Event (Autotuner.py:186-205) */
int fvmi_stream_create(void** stream);
int fvmi_stream_destroy(void* stream);
int fvmi_stream_synchronize(void* stream);
int fvmi_event_create(void** event);
int fvmi_event_destroy(void* event);
int fvmi_event_record(void* event, void* stream);
int fvmi_event_synchronize(void* event);
/* Work queued on `stream` after this call waits for `event` (cuda.Stream.wait_for_event). */
int fvmi_stream_wait_event(void* stream, void* event);
int fvmi_event_elapsed_ms(void* start, void* end, float* ms);

/* ---- memory: gpuarray.zeros / free (Common.py:510, 535-538), Memcpy2D (:543-616) */
int fvmi_malloc(void** ptr, size_t bytes);
/* Pitched 2-D allocation; pitch is a multiple of 256 bytes. */
int fvmi_malloc_pitch(void** ptr, size_t* pitch, size_t width_bytes, size_t height);
int fvmi_free(void* ptr);
/* Page-locked host memory (cuda.pagelocked_empty, Common.py:564) */
int fvmi_host_malloc(void** ptr, size_t bytes);
int fvmi_host_free(void* ptr);
int fvmi_memcpy_async(void* dst, const void* src, size_t bytes, int kind, void* stream);
int fvmi_memcpy2d_async(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                        size_t height, int kind, void* stream);
/* Fills a width x height float32 region with `value` using a kernel (safe under
 * stream capture, unlike a memset node). */
int fvmi_fill_f32(float* ptr, size_t pitch_bytes, int width, int height, float value, void* stream);

/* Device-to-device copy of `count` floats (16-byte aligned, count % 4 == 0)
 * with a streaming kernel; also the known-size reference for HBM counters. */
int fvmi_copy_f32(float* dst, const float* src, size_t count, void* stream);

/* ---- substep kernels ------------------------------------------------------
 * One dimensionally split substep from the input fields (…0) into the output
 * fields (…1): step 0 sweeps x then y, step 1 y then x.  `dt` is the substep
 * length (the wrappers pass dt*0.5).  Boundary fill happens on the device.
 *
 * `cfl`, if not NULL, points to two device floats.  A step-1 launch atomically
 * minimises slot 1 with the per-cell admissible dt min(dx/(|u|+c), dy/(|v|+c))
 * of every cell it writes; every launch sets the other slot (1 - step) to
 * +inf.  After the usual step-0/step-1 pair slot 1 therefore holds the CFL
 * minimum of the last substep, which is what the reference's
 * `gpuarray.min(cfl_data)` reads after a full step (KP07_dimsplit.py:135-137).
 * Before the first substep set slot 1 to the initial value.
 */

/* KP07DimsplitKernel, SWE2D_KP07_dimsplit.cu:123-144 (argument order kept). */
int fvmi_swe_kp07_dimsplit_substep(int nx, int ny, float dx, float dy, float dt, float g, float theta, int step,
                                   int boundary_conditions, const float* h0, int h0_pitch, const float* hu0,
                                   int hu0_pitch, const float* hv0, int hv0_pitch, float* h1, int h1_pitch,
                                   float* hu1, int hu1_pitch, float* hv1, int hv1_pitch, float* cfl,
                                   const fvmi_launch_t* launch, void* stream);

/* HLL2Kernel, SWE2D_HLL2.cu:126-147. */
int fvmi_swe_hll2_substep(int nx, int ny, float dx, float dy, float dt, float g, float theta, int step,
                          int boundary_conditions, const float* h0, int h0_pitch, const float* hu0, int hu0_pitch,
                          const float* hv0, int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch,
                          float* hv1, int hv1_pitch, float* cfl, const fvmi_launch_t* launch, void* stream);

/* WAFKernel, SWE2D_WAF.cu:103-119.  The reference kernel has no CFL output
 * (WAF.py:82,115 passes one it ignores); here `cfl` is filled as for KP07. */
int fvmi_swe_waf_substep(int nx, int ny, float dx, float dy, float dt, float g, int step, int boundary_conditions,
                         const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                         int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1,
                         int hv1_pitch, float* cfl, const fvmi_launch_t* launch, void* stream);

/* ---- the reference's other shallow-water schemes --------------------------
 * One launch of LxFKernel (SWE2D_LxF.cu:93-110), HLLKernel (SWE2D_HLL.cu:
 * 100-115) or FORCEKernel (SWE2D_FORCE.cu:88-105): one full first-order step
 * on fields with ONE ghost cell a side ((ny+2) rows of pitch >= (nx+2)*4
 * bytes).  Arguments in the reference's order; then `cfl_slot` (0 or 1): the
 * launch minimises cfl[cfl_slot] with the per-cell admissible dt of the cells
 * it writes and sets cfl[1 - cfl_slot] to +inf, so a caller alternating the
 * slot reads the last launch's minimum without a reset in between.  LxF and
 * FORCE need dt > 0 (their fluxes divide by it). */
int fvmi_swe_lxf_substep(int nx, int ny, float dx, float dy, float dt, float g, int boundary_conditions,
                         const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                         int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1, int hv1_pitch,
                         float* cfl, int cfl_slot, const fvmi_launch_t* launch, void* stream);
int fvmi_swe_hll_substep(int nx, int ny, float dx, float dy, float dt, float g, int boundary_conditions,
                         const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                         int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1, int hv1_pitch,
                         float* cfl, int cfl_slot, const fvmi_launch_t* launch, void* stream);
int fvmi_swe_force_substep(int nx, int ny, float dx, float dy, float dt, float g, int boundary_conditions,
                           const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                           int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1,
                           int hv1_pitch, float* cfl, int cfl_slot, const fvmi_launch_t* launch, void* stream);

/* KP07Kernel, SWE2D_KP07.cu:120-146: one stage of the unsplit KP07 scheme
 * with TVD RK2, fields with TWO ghost cells a side.  step_order = step << 16 |
 * order (Simulator.stepOrderToCodedInt); the second stage of order 2 writes
 * 0.5 * (h1_old + Q), reading the output buffer first.  cfl_slot as above;
 * the CFL is that of Q before the average, as the reference computes it. */
int fvmi_swe_kp07_substep(int nx, int ny, float dx, float dy, float dt, float g, float theta, int step_order,
                          int boundary_conditions, const float* h0, int h0_pitch, const float* hu0, int hu0_pitch,
                          const float* hv0, int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch,
                          float* hv1, int hv1_pitch, float* cfl, int cfl_slot, const fvmi_launch_t* launch,
                          void* stream);

/* Euler KP07DimsplitKernel, EE2D_KP07_dimsplit.cu:126-154.  x0,y0,x1,y1 select
 * an interior sub-area [x0,x1) x [y0,y1) to update (x1 == 0 -> nx, y1 == 0 -> ny,
 * as the reference's defaults); boundary conditions are applied at the true
 * domain edges whatever the sub-area. */
int fvmi_euler_kp07_dimsplit_substep(int nx, int ny, float dx, float dy, float dt, float g, float gamma, float theta,
                                     int step, int boundary_conditions, const float* rho0, int rho0_pitch,
                                     const float* rho_u0, int rho_u0_pitch, const float* rho_v0, int rho_v0_pitch,
                                     const float* E0, int E0_pitch, float* rho1, int rho1_pitch, float* rho_u1,
                                     int rho_u1_pitch, float* rho_v1, int rho_v1_pitch, float* E1, int E1_pitch,
                                     float* cfl, int x0, int y0, int x1, int y1, const fvmi_launch_t* launch,
                                     void* stream);

/* One-rank stand-in for MPISimulator.full_exchange's east/west leg
 * (MPISimulator.py:283-295, 440-480) with a periodic-x neighbour: for each of
 * the `nvars` fields copy padded columns [nx, nx+2) -> [0, 2) and
 * [2, 4) -> [nx+2, nx+4) over all ny+4 rows, on the device. */
int fvmi_halo_copy_periodic_x(int nx, int ny, int nvars, float* const* fields, const int* pitches, void* stream);

/* One rectangular block of float32 cells: `height` rows of `width` values,
 * row pitches in bytes. */
typedef struct fvmi_region_copy {
    const float* src;
    float* dst;
    int src_pitch;
    int dst_pitch;
    int width;
    int height;
} fvmi_region_copy_t;

/* SHMEMSimulator's exchange on one device (SHMEMSimulator.py:200-263, where
 * each region goes through host memory with CudaArray2D.download/upload,
 * Common.py:543-616): copies `count` regions device to device in launches of
 * at most 64 regions, in list order between launches.  Regions of one call
 * must not overlap a region another of them writes. */
int fvmi_copy_regions_f32(const fvmi_region_copy_t* regions, int count, void* stream);

/* The same exchange for a rank that is its own neighbour in x and/or y
 * (MPIGrid with one rank and periodic edges, MPISimulator.py:247-309): with
 * wrap_y, ghost rows [0,2) <- [ny,ny+2) and [ny+2,ny+4) <- [2,4) over the
 * interior columns; with wrap_x, the ghost columns over all ny+4 rows, corners
 * taken after the row exchange (the reference's N/S-then-E/W order). */
int fvmi_halo_exchange_periodic(int nx, int ny, int nvars, float* const* fields, const int* pitches, int wrap_x,
                                int wrap_y, void* stream);

/* ArakawaA2D.check (Common.py:772-779): sum of the interior (float64) and the
 * number of non-finite interior values of one field of nx x ny interior cells
 * with x_halo ghost columns and y_halo ghost rows a side (2 for the
 * dimensionally split schemes and the unsplit KP07, 1 for LxF, HLL and FORCE),
 * written to device memory (`sum_out`: 1 double, `nonfinite_out`: 1 unsigned
 * int, both accumulated, so zero them first).  pitch >= (nx + 2 x_halo) * 4. */
int fvmi_field_stats(int nx, int ny, int x_halo, int y_halo, const float* field, int pitch, double* sum_out,
                     unsigned int* nonfinite_out, void* stream);

/* ---- batched ensemble ------------------------------------------------------
 * No counterpart in the reference, whose simulators run one grid per object
 * (Simulator.py:109-266).  Many grids of one scheme are advanced together; an
 * instance's time stepping reproduces BaseSimulator.simulate (Simulator.py:176-228)
 * on the device, so each grid evolves exactly as it would alone.
 *
 * One full step of every instance = fvmi_ensemble_prepare, fvmi_ensemble_substep
 * (step 0), fvmi_ensemble_substep (step 1), fvmi_ensemble_advance, all on one
 * stream; the sequence may be captured in a graph (fvmi_stream_begin_capture).
 */
typedef struct fvmi_instance {
    float* field[2][4]; /* [buffer][variable]: padded fields; buffer 0 holds the state between steps */
    int pitch;          /* bytes, shared by all fields of the instance */
    int nx, ny;
    int bc;             /* packed boundary conditions */
    float dx, dy, g, theta;
    float gamma, sqrt_g, inv_g, inv_dx;
    float inv_dy;
    float cfl[2];       /* CFL minimum of substep 0 / 1; cfl[1] = initial CFL before the first step */
    float dt_step;      /* float32 length of the current step */
    int nt;             /* steps taken */
    int active;         /* 0 once t >= t_end */
    int fixed_dt;       /* 1: keep `dt`, no CFL refresh (simulate(t, dt=...)) */
    double t, t_end, dt;
    double cfl_scale;
} fvmi_instance_t;

/* dt policy and CFL reset for every instance (one thread each). */
int fvmi_ensemble_prepare(fvmi_instance_t* table, int n, void* stream);
/* One substep of every active instance over the work list `segs` (int4 per
 * segment: instance, strip index, first interior row, row count; strips are
 * 124 columns wide for EE2D_KP07_dimsplit in fast arithmetic, which marches
 * two columns per lane, and 60 for the others; a segment never crosses a
 * strip).  The number of valid
 * segments is read on the device from *num_segs_dev, so a captured graph stays
 * valid when the host rewrites the list and its count between replays.
 * `waves` wavefronts are launched (<= 0: enough to fill the device); each
 * takes segments waves apart, so any count is covered.  launch->xcd_remap is
 * not used: blocks keep dispatch order, which spreads the longest-first list
 * evenly over the XCDs. */
int fvmi_ensemble_substep(int scheme, fvmi_instance_t* table, const int* segs, const int* num_segs_dev, int waves,
                          int step, const fvmi_launch_t* launch, void* stream);
/* The work split of one substep launch of the four dimensionally split entry
 * points and of the unsplit ones, computed on the host with no device call:
 * the planner's choice for scheme (3 Euler in fast arithmetic: strips of 124
 * columns; 0 KP07_dimsplit, 1 HLL2, 2 WAF and 3 Euler in exact arithmetic: 60;
 * 4, 5, 6 LxF, HLL, FORCE: 62; 7 the unsplit KP07: 60), sub-area [x0,x1)x[y0,y1) of an nx x ny domain (x1 = y1 = 0: the
 * whole domain, the only one the unsplit schemes take), the launch shape and
 * a device of `cu_count` CUs.  For every wave of the grid, 3 ints in `segs`
 * (capacity `max_waves` waves): strip, first row and row count relative to
 * the sub-area, or (-1, 0, 0) for a wave with nothing to do; *n_waves receives
 * the grid's wave count.  No counterpart in the reference (its tile grid is
 * ceil(n / block) in each direction); for tests and tools that check that the
 * segments cover the sub-area once. */
int fvmi_substep_plan(int scheme, int nx, int ny, int x0, int y0, int x1, int y1, const fvmi_launch_t* launch,
                      int cu_count, int* segs, int max_waves, int* n_waves);
/* t += dt_step, nt += 1 for every instance that stepped. */
int fvmi_ensemble_advance(fvmi_instance_t* table, int n, void* stream);
/* Writes to *count_dev the number of instances that still have to step. */
int fvmi_ensemble_count_active(const fvmi_instance_t* table, int n, int* count_dev, void* stream);

#define FVMI_SCHEME_KP07_DIMSPLIT 0
#define FVMI_SCHEME_HLL2 1
#define FVMI_SCHEME_WAF 2
#define FVMI_SCHEME_EE2D_KP07_DIMSPLIT 3

/* ---- stream capture (hipGraph) -------------------------------------------- */
int fvmi_stream_begin_capture(void* stream);
/* Ends the capture and instantiates the graph. */
int fvmi_stream_end_capture(void* stream, void** graph_exec);
int fvmi_graph_launch(void* graph_exec, void* stream);
int fvmi_graph_destroy(void* graph_exec);

#ifdef __cplusplus
}
#endif

#endif /* FVMI_H */
