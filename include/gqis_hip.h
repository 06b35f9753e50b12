/*
 * gqis_hip.h -- C ABI of the MI355X (gfx950) batched parameter-sweep integrator.
 *
 * The library replaces the reference's CuPy/NVRTC device layer for the sweep path:
 *   compile  : cp.RawKernel(kernel_code, "time_evolution_kernel", options).compile()
 *              (reference gqis/solver.py:2237-2241)
 *   launch   : time_evolution_kernel(grid, block, args)            (gqis/solver.py:2340)
 *   timing   : cp.cuda.Event() pair around the launch               (gqis/solver.py:2316-2345)
 *   checks   : cp.all(cp.isfinite(results)) after the launch        (gqis/solver.py:2351-2362)
 * Kernels are HIP source compiled at run time with hiprtc for the device's architecture; all
 * state, axis and output buffers are caller-owned device memory.  Every function returns 0 on
 * success and a negative GQ_E* code otherwise; gq_last_error() gives the message (per thread).
 *
 * This header is also prepended to every generated kernel source, so the argument structs
 * below have one definition on the host and on the device.  It uses builtin types only.
 */
#ifndef GQIS_HIP_H
#define GQIS_HIP_H

#define GQ_OK 0
#define GQ_E_INVALID (-1)     /* bad argument (maps to ValueError)                   */
#define GQ_E_COMPILE (-2)     /* hiprtc compilation failed (log in gq_last_error)    */
#define GQ_E_HIP (-3)         /* HIP runtime error                                   */
#define GQ_E_NOT_FOUND (-4)   /* kernel name not in the module                       */

/* status bits written by the kernels with one atomic OR per wavefront */
#define GQ_STATUS_NONFINITE_RESULT 1u   /* a result value is NaN or Inf              */
#define GQ_STATUS_NONFINITE_TRACE 2u    /* a time-trace value is NaN or Inf          */
#define GQ_STATUS_SUBSTEP_OVERFLOW 4u   /* an adaptive point exceeded max_substeps   */

#define GQ_MAX_SOLVER_PARAMS 8

/*
 * Arguments of one sweep launch (fixed-step and adaptive kernels).  Pointers are device
 * addresses; element type "real" is float or double as the module was compiled.
 * Flat point p = ix * num_y + iy; a launch covers p in [point_begin, point_begin+point_count)
 * (param_mode 1: the point parameters are given per launch point instead)
 * and writes its outputs at the local index q = p - point_begin:
 *   complex results   result[q]                         (real2: re, im)
 *   state results     result[i * point_count + q]       (i < state size)
 *   observable trace  trace[k * point_count + q]        (real2)
 *   state trace       trace[(k * D + i) * point_count + q]
 *   initial states    rho0[q * D + i]                   (the reference's (nx, ny, D) order)
 */
typedef struct gq_sweep_args {
    double dt;                 /* uniform tlist step, FP64 on the host                       */
    double t_in;               /* absolute time of the initial state                         */
    long long num_x;           /* sweep axis lengths                                        */
    long long num_y;
    long long point_begin;
    long long point_count;
    int num_steps;             /* tlist intervals                                            */
    int warmup_steps;          /* steps excluded from the mean                               */
    int trace_stride;
    int trace_offset;
    int num_trace;             /* 0: no trace                                                */
    int trace_point_stride;    /* leading dimension of the trace arrays (>= point_count)     */
    int param_mode;            /* 0: par_x/par_y are the sweep axes (p = ix*num_y + iy);      */
                               /* 1: they hold one value per launch point (q), for batches of */
                               /*    requests merged into one launch                          */
    int reserved;
    const void* par_x;         /* real[num_x]                                                */
    const void* par_y;         /* real[num_y]                                                */
    const void* consts;        /* real[]: runtime constants                                  */
    const void* rho0;          /* real[point_count * D] or null                              */
    const void* time_table;    /* real[stage times * time atoms] or null                     */
    double solver_param[GQ_MAX_SOLVER_PARAMS];
    void* result;
    void* trace;
    int* adaptive_steps;       /* int[2 * point_count] accepted/rejected, or null           */
    unsigned int* status;      /* one word, OR-ed with GQ_STATUS_* bits                      */
    unsigned int* work_cursor; /* one word for persistent kernels (zero before the launch)  */
} gq_sweep_args;

#ifndef __HIPCC_RTC__
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gq_ctx gq_ctx;
typedef struct gq_module gq_module;

/* One context per process and device: selects the device, reads its gcnArchName and compute
 * unit count.  Replaces CuPy's implicit current device (gqis/solver.py:16, 2237). */
int gq_ctx_create(int device_ordinal, gq_ctx** out);
void gq_ctx_destroy(gq_ctx* ctx);
/* Architecture name ("gfx950") and compute-unit count of the context's device. */
int gq_ctx_info(gq_ctx* ctx, char* arch, int arch_len, int* compute_units);

/* Compile HIP source with hiprtc for the context's architecture and load it.  Modules are
 * cached in the context by a hash of compiler version, source, options and architecture; a
 * cache hit returns the loaded module (cached_out = 1); with an on-disk cache directory set, a
 * code object found there is loaded without compiling (cached_out = 2) and new ones are stored
 * there.  Replaces cp.RawKernel(...).compile() (gqis/solver.py:2237-2241). */
/* Directory of the on-disk code-object cache (NULL or "" disables it; it must exist).  No
 * counterpart in the reference, whose kernel cache lives only in the process
 * (gqis/solver.py:2054-2060): it removes the compile from a new process's first sweep. */
int gq_ctx_set_cache_dir(gq_ctx* ctx, const char* dir);
int gq_module_compile(gq_ctx* ctx, const char* source, const char* const* options, int num_options,
                      gq_module** out, int* cached_out, double* build_seconds);
/* Compile only (no device needed): returns the code-object size, or the log on failure.
 * Used by the build check on machines without a GPU. */
int gq_compile_check(const char* source, const char* arch, const char* const* options,
                     int num_options, long long* code_size);
/* Launch kernel `name` of a module over `grid` x `block` threads with one gq_sweep_args by
 * value, on a HIP stream (null: the default stream).  Replaces
 * time_evolution_kernel(grid, block, args) (gqis/solver.py:2322-2340). */
int gq_launch_sweep(gq_module* module, const char* name, unsigned int grid, unsigned int block,
                    const gq_sweep_args* args, void* stream);
/* Number of work-groups of `block` threads of kernel `name` that fit on the device at once
 * (occupancy x compute units): the grid of a persistent launch. */
int gq_resident_groups(gq_module* module, const char* name, unsigned int block, int* groups);
/* Drop the caller's reference from gq_module_compile.  A module no caller holds stays cached
 * for reuse; beyond 128 cached modules the least recently used unheld ones are unloaded. */
void gq_module_release(gq_module* module);

/* A fixed chain of launches of one module with one argument struct, recorded once as a HIP
 * graph and replayed with a single call (the inputs live in device buffers the caller updates
 * between replays).  No counterpart in the reference, which launches once per call
 * (gqis/solver.py:2322-2340); it removes per-launch host work from repeated sweeps.  The chain
 * should start with gq_clear_flags_<f32|f64> (status word and work cursor back to zero: a
 * kernel, not a memset, inside the graph). */
typedef struct gq_graph gq_graph;
typedef struct gq_launch_desc {
    const char* name;
    unsigned int grid, block;
} gq_launch_desc;
int gq_graph_create(gq_module* module, const gq_launch_desc* launches, int count,
                    const gq_sweep_args* args, gq_graph** out);
int gq_graph_launch(gq_graph* graph, void* stream);
void gq_graph_destroy(gq_graph* graph);

/* Size of the library's per-device scratch array used by gq_launch_sweep_noalloc: 256 bytes
 * for the work cursor, the rest for a time table (4 MiB in all: 524 256 FP64 table entries). */
#define GQ_SCRATCH_BYTES (4u << 20)
#define GQ_SCRATCH_TABLE_OFFSET 256u

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the launch chain of
 * gq_graph_create, run once on `stream` like gq_launch_sweep, with the sweep's scratch -- the
 * work cursor of the persistent adaptive kernels and, when table_bytes > 0, the time table of
 * the fixed-step kernels (args->time_table and args->work_cursor are replaced) -- in a
 * library-owned static device array of GQ_SCRATCH_BYTES per GPU instead of caller buffers.
 * The caller still passes the inputs and outputs; the status word args->status (an output)
 * and the cursor are cleared by a kernel before the chain, so the chain should not start with
 * gq_clear_flags.  A time table above GQ_SCRATCH_BYTES - GQ_SCRATCH_TABLE_OFFSET bytes is
 * rejected (GQ_E_INVALID).  Nothing is allocated.  All calls on one GPU share its array, which
 * keeps the last call's contents: calls that overlap on one GPU (other streams or threads)
 * race and must be serialised by the caller. */
int gq_launch_sweep_noalloc(gq_module* module, const gq_launch_desc* launches, int count,
                            const gq_sweep_args* args, unsigned long long table_bytes,
                            void* stream);

/* Event pair timing on a stream (cp.cuda.Event, gqis/solver.py:2316-2345). */
int gq_event_create(void** event);
int gq_event_record(void* event, void* stream);
int gq_event_elapsed_ms(void* start, void* stop, float* ms);
int gq_event_destroy(void* event);
int gq_stream_synchronize(void* stream);

/* Read (and optionally clear) a device status word written by the kernels; synchronises the
 * stream first.  Replaces the isfinite reduction (gqis/solver.py:2351-2362). */
int gq_status_read(unsigned int* device_status, void* stream, unsigned int* flags, int clear);

/* sizeof(gq_sweep_args) as compiled into the library (host-side layout check). */
unsigned long long gq_sweep_args_size(void);

/* Message of the last failing call on this thread. */
const char* gq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* __HIPCC_RTC__ */

#endif /* GQIS_HIP_H */
