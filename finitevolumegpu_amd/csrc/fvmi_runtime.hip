// C ABI of libfvmi.so: device runtime, memory, and the validated entry points of
// the substep kernels.  Declarations and reference citations: include/fvmi.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/fvmi.h"
#include "fv_launch.h"
#include "fv_unsplit.h"

using namespace fvmi;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(FVMI_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

#define FVMI_HIP(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

struct Ctx {
    int device;
};

// Compute-unit count of the current device, cached per device ordinal.
int current_cu_count() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] == 0) {
        hipDeviceProp_t p;
        cache[dev] = (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0)
                         ? p.multiProcessorCount
                         : 256;
    }
    return cache[dev];
}

hipMemcpyKind kind_of(int kind) {
    switch (kind) {
        case FVMI_MEMCPY_HTOD: return hipMemcpyHostToDevice;
        case FVMI_MEMCPY_DTOH: return hipMemcpyDeviceToHost;
        case FVMI_MEMCPY_DTOD: return hipMemcpyDeviceToDevice;
        default: return hipMemcpyDefault;
    }
}

bool valid_bc(int side) { return side == BC_DIRICHLET || side == BC_PERIODIC || side == BC_REFLECTIVE || side == BC_OUTFLOW; }

int check_bc(int bc) {
    const int sides[4] = {(bc >> 24) & 0xF, (bc >> 16) & 0xF, (bc >> 8) & 0xF, bc & 0xF};
    for (int s : sides) {
        if (s == BC_NEUMANN) return fail(FVMI_ERR_UNSUPPORTED, "Neumann boundary condition not supported");
        if (!valid_bc(s)) return fail(FVMI_ERR_INVALID, "invalid boundary condition code %d", s);
    }
    return FVMI_OK;
}

int fill_launch(const fvmi_launch_t* l, SubstepArgs& a, int* math) {
    a.seg_rows = (l && l->tile_rows > 0) ? l->tile_rows : 0;  // 0 = balanced
    a.waves_per_block = (l && l->waves_per_block > 0) ? l->waves_per_block : 4;
    a.xcd_remap = (l && l->xcd_remap == 2) ? 0 : 1;
    *math = l ? l->math : FVMI_MATH_FAST;
    if (a.seg_rows > (1 << 20)) return fail(FVMI_ERR_INVALID, "tile_rows %d out of range", a.seg_rows);
    if (a.waves_per_block > 4) return fail(FVMI_ERR_INVALID, "waves_per_block %d out of range [1, 4]", a.waves_per_block);
    if (*math != FVMI_MATH_FAST && *math != FVMI_MATH_EXACT) return fail(FVMI_ERR_INVALID, "unknown math mode %d", *math);
    return FVMI_OK;
}

int run_substep(int scheme, SubstepArgs& a, int nvars, const fvmi_launch_t* launch, void* stream) {
    if (a.nx < 2 || a.ny < 2) return fail(FVMI_ERR_INVALID, "grid %dx%d too small (need nx, ny >= 2)", a.nx, a.ny);
    if ((size_t)(a.ny + 4) * (size_t)(a.nx + 4) > (size_t)1 << 30)
        return fail(FVMI_ERR_INVALID, "grid %dx%d too large", a.nx, a.ny);
    if (a.step != 0 && a.step != 1) return fail(FVMI_ERR_INVALID, "step must be 0 or 1, got %d", a.step);
    if (int rc = check_bc(a.bc)) return rc;
    if (!(a.dt >= 0.0f) || !std::isfinite(a.dt)) return fail(FVMI_ERR_INVALID, "dt must be finite and >= 0");
    if (!(a.dx > 0.0f) || !(a.dy > 0.0f)) return fail(FVMI_ERR_INVALID, "dx, dy must be > 0");
    const int min_pitch = (a.nx + 4) * (int)sizeof(float);
    // schemes that march two columns per lane move them with 8-byte loads and
    // stores: their fields and pitches must be 8-byte aligned
    const int req_math = launch ? launch->math : FVMI_MATH_FAST;
    const int align = scheme_cols_per_lane(scheme, req_math) == 2 ? 8 : 4;
    for (int k = 0; k < nvars; ++k) {
        if (!a.in[k] || !a.out[k]) return fail(FVMI_ERR_INVALID, "field pointer %d is NULL", k);
        if (a.in[k] == a.out[k]) return fail(FVMI_ERR_INVALID, "field %d: input and output alias", k);
        if (a.in_pitch[k] < min_pitch || a.out_pitch[k] < min_pitch || (a.in_pitch[k] % align) ||
            (a.out_pitch[k] % align))
            return fail(FVMI_ERR_INVALID, "field %d: pitch must be >= %d bytes and a multiple of %d", k, min_pitch,
                        align);
        if (((uintptr_t)a.in[k] % align) || ((uintptr_t)a.out[k] % align))
            return fail(FVMI_ERR_INVALID, "field %d: pointers must be %d-byte aligned", k, align);
        // fields are addressed through buffer descriptors whose range must stay
        // below the 2 GiB offset that masked lanes' stores use
        if ((size_t)(a.ny + 4) * (size_t)a.in_pitch[k] >= ((size_t)1 << 31) ||
            (size_t)(a.ny + 4) * (size_t)a.out_pitch[k] >= ((size_t)1 << 31))
            return fail(FVMI_ERR_INVALID, "field %d: (ny+4)*pitch must be below 2 GiB", k);
    }
    if (a.x1 == 0) a.x1 = a.nx;
    if (a.y1 == 0) a.y1 = a.ny;
    if (a.x0 < 0 || a.y0 < 0 || a.x1 > a.nx || a.y1 > a.ny || a.x0 >= a.x1 || a.y0 >= a.y1)
        return fail(FVMI_ERR_INVALID, "sub-area [%d,%d)x[%d,%d) outside the %dx%d domain", a.x0, a.x1, a.y0, a.y1,
                    a.nx, a.ny);
    int math = 0;
    if (int rc = fill_launch(launch, a, &math)) return rc;
    // constants of the fast policy, rounded once on the host
    a.hg = 0.5f * a.g;
    a.sqrt_g = std::sqrt(a.g > 0.0f ? a.g : 0.0f);
    a.inv_g = a.g != 0.0f ? 1.0f / a.g : 0.0f;
    a.inv_dx = 1.0f / a.dx;
    a.inv_dy = 1.0f / a.dy;
    a.dtdx = a.dt / a.dx;
    a.dtdy = a.dt / a.dy;
    set_slope_constants(a);
    a.strip_cols = scheme_strip_cols(scheme, math);
    // WAF keeps equal segments: the round-weighted plan made it 3 % slower
    const int grid = plan_substep(a, current_cu_count(), scheme != SCHEME_WAF);
    hipError_t e = (math == FVMI_MATH_EXACT) ? launch_substep_exact(scheme, a, grid, (hipStream_t)stream)
                                             : launch_substep_fast(scheme, a, grid, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "substep launch");
    return FVMI_OK;
}

__global__ void fill_kernel(float* p, size_t pitch, int w, int h, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w) return;
    for (int j = blockIdx.y; j < h; j += gridDim.y)  // grid.y is capped at 65535 rows
        reinterpret_cast<float*>(reinterpret_cast<char*>(p) + (size_t)j * pitch)[i] = v;
}

__global__ void copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Batched 2-D region copies (SHMEMSimulator's exchange): up to kMaxRegions
// regions per launch, passed by value; block row y copies region y.
constexpr int kMaxRegions = 64;
struct RegionBatch {
    fvmi_region_copy_t r[kMaxRegions];
};

__global__ void copy_regions_kernel(RegionBatch b) {
    const fvmi_region_copy_t R = b.r[blockIdx.y];
    const long long n = (long long)R.width * R.height;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int row = (int)(i / R.width);
        const int col = (int)(i - (long long)row * R.width);
        const float v = reinterpret_cast<const float*>(reinterpret_cast<const char*>(R.src) + (size_t)row * R.src_pitch)[col];
        reinterpret_cast<float*>(reinterpret_cast<char*>(R.dst) + (size_t)row * R.dst_pitch)[col] = v;
    }
}

struct HaloArgs {
    float* f[4];
    int pitch[4];
    int nx, ny, nvars;
    int wrap_x, wrap_y;
};

// One thread per ghost-frame cell and field.  The source of every ghost cell is
// read from memory that this launch does not write, so the exchange is a single
// race-free pass: north/south rows over the interior columns, then east/west
// columns over the full height with the corners taken from the wrapped rows,
// the order of MPISimulator.full_exchange (MPISimulator.py:371-480).
__global__ void halo_periodic_kernel(HaloArgs a) {
    const int W = a.nx + 4;
    const int frame = 2 * W * 2 + 2 * 2 * a.ny;  // ghost rows (full width) + ghost columns (interior rows)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int var = blockIdx.y;
    if (t >= frame || var >= a.nvars) return;
    int r, c;
    if (t < 4 * W) {  // rows 0,1,ny+2,ny+3
        const int k = t / W;
        r = (k < 2) ? k : a.ny + k;
        c = t - k * W;
    } else {  // columns 0,1,nx+2,nx+3 of the interior rows
        const int u = t - 4 * W;
        const int k = u / a.ny;
        r = 2 + (u - k * a.ny);
        c = (k < 2) ? k : a.nx + k;
    }
    const bool ghost_c = c < 2 || c >= a.nx + 2;
    const bool ghost_r = r < 2 || r >= a.ny + 2;
    int sr = r, sc = c;
    if (ghost_c) {
        // east/west leg: full height, rows as they stand after the north/south leg
        if (!a.wrap_x) return;
        sc = (c < 2) ? c + a.nx : c - a.nx;
        if (ghost_r && a.wrap_y) sr = (r < 2) ? r + a.ny : r - a.ny;
    } else {
        // north/south leg: interior columns only
        if (!a.wrap_y) return;
        sr = (r < 2) ? r + a.ny : r - a.ny;
    }
    char* base = reinterpret_cast<char*>(a.f[var]);
    const float v = reinterpret_cast<const float*>(base + (size_t)sr * a.pitch[var])[sc];
    reinterpret_cast<float*>(base + (size_t)r * a.pitch[var])[c] = v;
}

// Ensemble bookkeeping (one thread per instance); see fv_ensemble.h.
__global__ void ensemble_prepare_kernel(fvmi_instance_t* table, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fvmi_instance_t& I = table[i];
    if (!I.active) return;
    if (!(I.t < I.t_end)) {
        I.active = 0;
        return;
    }
    if (!I.fixed_dt && (I.nt % 100) == 0) I.dt = (double)I.cfl[1] * 0.5 * I.cfl_scale;
    const double rem = I.t_end - I.t;
    const float cur = (float)(I.dt < rem ? I.dt : rem);
    if (!(cur > 0.0f)) {
        I.active = 0;
        return;
    }
    I.dt_step = cur;
    I.cfl[0] = __builtin_inff();
    I.cfl[1] = __builtin_inff();
}

__global__ void ensemble_advance_kernel(fvmi_instance_t* table, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fvmi_instance_t& I = table[i];
    if (I.active) {
        I.t += (double)I.dt_step;
        I.nt += 1;
    }
}

__global__ void ensemble_count_kernel(const fvmi_instance_t* table, int n, int* out) {
    __shared__ int partial[16];
    int live = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) live += (table[i].active && table[i].t < table[i].t_end) ? 1 : 0;
    for (int off = 32; off >= 1; off >>= 1) live += __shfl_xor(live, off, 64);
    if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = live;
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) total += partial[w];
        *out = total;
    }
}

// `f` points at interior cell (0, 0); rows are `pitch` bytes apart
__global__ void field_stats_kernel(int nx, int ny, const float* f, int pitch, double* sum_out, unsigned int* bad_out) {
    double s = 0.0;
    unsigned int bad = 0;
    for (int j = blockIdx.y; j < ny; j += gridDim.y) {
        const float* row = reinterpret_cast<const float*>(reinterpret_cast<const char*>(f) + (size_t)j * pitch);
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx; i += gridDim.x * blockDim.x) {
            const float v = row[i];
            if (std::isfinite(v))
                s += (double)v;
            else
                ++bad;
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        bad += __shfl_xor(bad, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(sum_out, s);
        if (bad) atomicAdd(bad_out, bad);
    }
}

// Shared checks and launch of the unsplit / first-order schemes (gc ghost cells).
int run_unsplit(int scheme, UnsplitArgs& a, int gc, const fvmi_launch_t* launch, void* stream) {
    if (a.nx < 2 || a.ny < 2) return fail(FVMI_ERR_INVALID, "grid %dx%d too small (need nx, ny >= 2)", a.nx, a.ny);
    if ((size_t)(a.ny + 2 * gc) * (size_t)(a.nx + 2 * gc) > (size_t)1 << 30)
        return fail(FVMI_ERR_INVALID, "grid %dx%d too large", a.nx, a.ny);
    if (int rc = check_bc(a.bc)) return rc;
    if (a.cfl_slot != 0 && a.cfl_slot != 1) return fail(FVMI_ERR_INVALID, "cfl_slot must be 0 or 1");
    const bool centred = scheme == USCHEME_LXF || scheme == USCHEME_FORCE;
    if (!std::isfinite(a.dt) || (centred ? !(a.dt > 0.0f) : !(a.dt >= 0.0f)))
        return fail(FVMI_ERR_INVALID, centred ? "dt must be finite and > 0" : "dt must be finite and >= 0");
    if (!(a.dx > 0.0f) || !(a.dy > 0.0f)) return fail(FVMI_ERR_INVALID, "dx, dy must be > 0");
    const int min_pitch = (a.nx + 2 * gc) * (int)sizeof(float);
    for (int k = 0; k < 3; ++k) {
        if (!a.in[k] || !a.out[k]) return fail(FVMI_ERR_INVALID, "field pointer %d is NULL", k);
        if (a.in[k] == a.out[k]) return fail(FVMI_ERR_INVALID, "field %d: input and output alias", k);
        if (a.in_pitch[k] < min_pitch || a.out_pitch[k] < min_pitch || (a.in_pitch[k] & 3) || (a.out_pitch[k] & 3))
            return fail(FVMI_ERR_INVALID, "field %d: pitch must be >= %d bytes and a multiple of 4", k, min_pitch);
        if ((size_t)(a.ny + 2 * gc) * (size_t)a.in_pitch[k] >= ((size_t)1 << 31) ||
            (size_t)(a.ny + 2 * gc) * (size_t)a.out_pitch[k] >= ((size_t)1 << 31))
            return fail(FVMI_ERR_INVALID, "field %d: (ny+%d)*pitch must be below 2 GiB", k, 2 * gc);
    }
    SubstepArgs shape{};
    int math = 0;
    if (int rc = fill_launch(launch, shape, &math)) return rc;
    a.seg_rows = shape.seg_rows;
    a.waves_per_block = shape.waves_per_block;
    a.xcd_remap = shape.xcd_remap;
    a.hg = 0.5f * a.g;
    a.sqrt_g = std::sqrt(a.g > 0.0f ? a.g : 0.0f);
    a.inv_dx = 1.0f / a.dx;
    a.inv_dy = 1.0f / a.dy;
    a.dtdx = a.dt / a.dx;
    a.dtdy = a.dt / a.dy;
    a.half_theta = 0.5f * a.theta;
    a.inv_2theta = a.theta > 0.0f ? 0.5f / a.theta : 0.0f;
    if (centred) {
        a.lxf_kx = a.dx / (4.0f * a.dt);
        a.lxf_ky = a.dy / (4.0f * a.dt);
        a.lf_kx = a.dx / (2.0f * a.dt);
        a.lf_ky = a.dy / (2.0f * a.dt);
        a.lw_kx = a.dt / (2.0f * a.dx);
        a.lw_ky = a.dt / (2.0f * a.dy);
    }
    const int grid = plan_unsplit(a, gc, current_cu_count(), unsplit_round_ok(scheme));
    hipError_t e = (math == FVMI_MATH_EXACT) ? launch_unsplit_exact(scheme, a, grid, (hipStream_t)stream)
                                             : launch_unsplit_fast(scheme, a, grid, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "unsplit launch");
    return FVMI_OK;
}

UnsplitArgs first_order_args(int nx, int ny, float dx, float dy, float dt, float g, int bc, const float* h0,
                             int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0, int hv0_pitch, float* h1,
                             int h1_pitch, float* hu1, int hu1_pitch, float* hv1, int hv1_pitch, float* cfl,
                             int cfl_slot) {
    UnsplitArgs a{};
    a.nx = nx, a.ny = ny, a.dx = dx, a.dy = dy, a.dt = dt, a.g = g, a.bc = bc;
    a.in[0] = h0, a.in[1] = hu0, a.in[2] = hv0;
    a.in_pitch[0] = h0_pitch, a.in_pitch[1] = hu0_pitch, a.in_pitch[2] = hv0_pitch;
    a.out[0] = h1, a.out[1] = hu1, a.out[2] = hv1;
    a.out_pitch[0] = h1_pitch, a.out_pitch[1] = hu1_pitch, a.out_pitch[2] = hv1_pitch;
    a.cfl = cfl;
    a.cfl_slot = cfl_slot;
    return a;
}

}  // namespace

extern "C" {

const char* fvmi_last_error(void) { return g_err; }

int fvmi_version(void) { return 1; }

int fvmi_device_count(int* count) {
    if (!count) return fail(FVMI_ERR_INVALID, "count is NULL");
    FVMI_HIP(hipGetDeviceCount(count));
    return FVMI_OK;
}

int fvmi_ctx_create(int device, int flags, void** ctx) {
    (void)flags;
    if (!ctx) return fail(FVMI_ERR_INVALID, "ctx is NULL");
    int n = 0;
    FVMI_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FVMI_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    FVMI_HIP(hipSetDevice(device));
    FVMI_HIP(hipFree(nullptr));  // force runtime initialisation on this device
    Ctx* c = new Ctx{device};
    *ctx = c;
    return FVMI_OK;
}

int fvmi_ctx_destroy(void* ctx) {
    delete static_cast<Ctx*>(ctx);
    return FVMI_OK;
}

int fvmi_ctx_make_current(void* ctx) {
    if (!ctx) return fail(FVMI_ERR_INVALID, "ctx is NULL");
    FVMI_HIP(hipSetDevice(static_cast<Ctx*>(ctx)->device));
    return FVMI_OK;
}

int fvmi_ctx_synchronize(void* ctx) {
    if (ctx) FVMI_HIP(hipSetDevice(static_cast<Ctx*>(ctx)->device));
    FVMI_HIP(hipDeviceSynchronize());
    return FVMI_OK;
}

int fvmi_ctx_device(void* ctx, int* device) {
    if (!ctx || !device) return fail(FVMI_ERR_INVALID, "NULL argument");
    *device = static_cast<Ctx*>(ctx)->device;
    return FVMI_OK;
}

int fvmi_device_name(int device, char* buf, int len) {
    if (!buf || len <= 0) return fail(FVMI_ERR_INVALID, "bad buffer");
    hipDeviceProp_t p;
    FVMI_HIP(hipGetDeviceProperties(&p, device));
    std::snprintf(buf, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
    return FVMI_OK;
}

int fvmi_device_cu_count(int device, int* cus) {
    if (!cus) return fail(FVMI_ERR_INVALID, "cus is NULL");
    hipDeviceProp_t p;
    FVMI_HIP(hipGetDeviceProperties(&p, device));
    *cus = p.multiProcessorCount;
    return FVMI_OK;
}

int fvmi_mem_get_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(FVMI_ERR_INVALID, "NULL argument");
    FVMI_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return FVMI_OK;
}

int fvmi_stream_create(void** stream) {
    if (!stream) return fail(FVMI_ERR_INVALID, "stream is NULL");
    hipStream_t s;
    FVMI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return FVMI_OK;
}

int fvmi_stream_destroy(void* stream) {
    if (stream) FVMI_HIP(hipStreamDestroy((hipStream_t)stream));
    return FVMI_OK;
}

int fvmi_stream_synchronize(void* stream) {
    FVMI_HIP(hipStreamSynchronize((hipStream_t)stream));
    return FVMI_OK;
}

int fvmi_event_create(void** event) {
    if (!event) return fail(FVMI_ERR_INVALID, "event is NULL");
    hipEvent_t e;
    FVMI_HIP(hipEventCreate(&e));
    *event = e;
    return FVMI_OK;
}

int fvmi_event_destroy(void* event) {
    if (event) FVMI_HIP(hipEventDestroy((hipEvent_t)event));
    return FVMI_OK;
}

int fvmi_event_record(void* event, void* stream) {
    FVMI_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return FVMI_OK;
}

int fvmi_stream_wait_event(void* stream, void* event) {
    if (!event) return fail(FVMI_ERR_INVALID, "event is NULL");
    FVMI_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return FVMI_OK;
}

int fvmi_event_synchronize(void* event) {
    FVMI_HIP(hipEventSynchronize((hipEvent_t)event));
    return FVMI_OK;
}

int fvmi_event_elapsed_ms(void* start, void* end, float* ms) {
    if (!ms) return fail(FVMI_ERR_INVALID, "ms is NULL");
    FVMI_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return FVMI_OK;
}

int fvmi_malloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(FVMI_ERR_INVALID, "ptr is NULL");
    FVMI_HIP(hipMalloc(ptr, bytes));
    return FVMI_OK;
}

int fvmi_malloc_pitch(void** ptr, size_t* pitch, size_t width_bytes, size_t height) {
    if (!ptr || !pitch) return fail(FVMI_ERR_INVALID, "NULL argument");
    const size_t p = (width_bytes + 255) & ~(size_t)255;
    FVMI_HIP(hipMalloc(ptr, p * height));
    *pitch = p;
    return FVMI_OK;
}

int fvmi_free(void* ptr) {
    if (ptr) FVMI_HIP(hipFree(ptr));
    return FVMI_OK;
}

int fvmi_host_malloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(FVMI_ERR_INVALID, "ptr is NULL");
    FVMI_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return FVMI_OK;
}

int fvmi_host_free(void* ptr) {
    if (ptr) FVMI_HIP(hipHostFree(ptr));
    return FVMI_OK;
}

int fvmi_memcpy_async(void* dst, const void* src, size_t bytes, int kind, void* stream) {
    FVMI_HIP(hipMemcpyAsync(dst, src, bytes, kind_of(kind), (hipStream_t)stream));
    return FVMI_OK;
}

int fvmi_memcpy2d_async(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes, size_t height,
                        int kind, void* stream) {
    if (width_bytes > dpitch || width_bytes > spitch) return fail(FVMI_ERR_INVALID, "row width exceeds pitch");
    FVMI_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, height, kind_of(kind), (hipStream_t)stream));
    return FVMI_OK;
}

int fvmi_fill_f32(float* ptr, size_t pitch_bytes, int width, int height, float value, void* stream) {
    if (!ptr || width < 0 || height < 0 || pitch_bytes < (size_t)width * sizeof(float))
        return fail(FVMI_ERR_INVALID, "bad fill region");
    if (width == 0 || height == 0) return FVMI_OK;
    hipLaunchKernelGGL(fill_kernel, dim3((width + 255) / 256, height < 65535 ? height : 65535), dim3(256), 0,
                       (hipStream_t)stream, ptr, pitch_bytes, width, height, value);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

int fvmi_copy_f32(float* dst, const float* src, size_t count, void* stream) {
    if (!dst || !src || (count & 3) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(FVMI_ERR_INVALID, "copy needs 16-byte aligned buffers and a multiple of 4 floats");
    if (count == 0) return FVMI_OK;
    const size_t n4 = count / 4;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<float4*>(dst), reinterpret_cast<const float4*>(src), n4);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

int fvmi_swe_kp07_dimsplit_substep(int nx, int ny, float dx, float dy, float dt, float g, float theta, int step,
                                   int boundary_conditions, const float* h0, int h0_pitch, const float* hu0,
                                   int hu0_pitch, const float* hv0, int hv0_pitch, float* h1, int h1_pitch,
                                   float* hu1, int hu1_pitch, float* hv1, int hv1_pitch, float* cfl,
                                   const fvmi_launch_t* launch, void* stream) {
    SubstepArgs a{};
    a.nx = nx, a.ny = ny, a.dx = dx, a.dy = dy, a.dt = dt, a.g = g, a.theta = theta, a.step = step;
    a.bc = boundary_conditions;
    a.in[0] = h0, a.in[1] = hu0, a.in[2] = hv0;
    a.in_pitch[0] = h0_pitch, a.in_pitch[1] = hu0_pitch, a.in_pitch[2] = hv0_pitch;
    a.out[0] = h1, a.out[1] = hu1, a.out[2] = hv1;
    a.out_pitch[0] = h1_pitch, a.out_pitch[1] = hu1_pitch, a.out_pitch[2] = hv1_pitch;
    a.cfl = cfl;
    return run_substep(SCHEME_KP07, a, 3, launch, stream);
}

int fvmi_swe_lxf_substep(int nx, int ny, float dx, float dy, float dt, float g, int boundary_conditions,
                         const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                         int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1, int hv1_pitch,
                         float* cfl, int cfl_slot, const fvmi_launch_t* launch, void* stream) {
    UnsplitArgs a = first_order_args(nx, ny, dx, dy, dt, g, boundary_conditions, h0, h0_pitch, hu0, hu0_pitch, hv0,
                                     hv0_pitch, h1, h1_pitch, hu1, hu1_pitch, hv1, hv1_pitch, cfl, cfl_slot);
    return run_unsplit(USCHEME_LXF, a, 1, launch, stream);
}

int fvmi_swe_hll_substep(int nx, int ny, float dx, float dy, float dt, float g, int boundary_conditions,
                         const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                         int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1, int hv1_pitch,
                         float* cfl, int cfl_slot, const fvmi_launch_t* launch, void* stream) {
    UnsplitArgs a = first_order_args(nx, ny, dx, dy, dt, g, boundary_conditions, h0, h0_pitch, hu0, hu0_pitch, hv0,
                                     hv0_pitch, h1, h1_pitch, hu1, hu1_pitch, hv1, hv1_pitch, cfl, cfl_slot);
    return run_unsplit(USCHEME_HLL, a, 1, launch, stream);
}

int fvmi_swe_force_substep(int nx, int ny, float dx, float dy, float dt, float g, int boundary_conditions,
                           const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                           int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1,
                           int hv1_pitch, float* cfl, int cfl_slot, const fvmi_launch_t* launch, void* stream) {
    UnsplitArgs a = first_order_args(nx, ny, dx, dy, dt, g, boundary_conditions, h0, h0_pitch, hu0, hu0_pitch, hv0,
                                     hv0_pitch, h1, h1_pitch, hu1, hu1_pitch, hv1, hv1_pitch, cfl, cfl_slot);
    return run_unsplit(USCHEME_FORCE, a, 1, launch, stream);
}

int fvmi_swe_kp07_substep(int nx, int ny, float dx, float dy, float dt, float g, float theta, int step_order,
                          int boundary_conditions, const float* h0, int h0_pitch, const float* hu0, int hu0_pitch,
                          const float* hv0, int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch,
                          float* hv1, int hv1_pitch, float* cfl, int cfl_slot, const fvmi_launch_t* launch,
                          void* stream) {
    const int step = step_order >> 16, order = step_order & 0xFFFF;
    if (order < 1 || order > 2 || step < 0 || step >= order)
        return fail(FVMI_ERR_INVALID, "step_order: step %d of order %d (order 1 or 2)", step, order);
    UnsplitArgs a = first_order_args(nx, ny, dx, dy, dt, g, boundary_conditions, h0, h0_pitch, hu0, hu0_pitch, hv0,
                                     hv0_pitch, h1, h1_pitch, hu1, hu1_pitch, hv1, hv1_pitch, cfl, cfl_slot);
    a.theta = theta;
    a.stage2 = (order == 2 && step == 1) ? 1 : 0;
    return run_unsplit(USCHEME_KP07, a, 2, launch, stream);
}

int fvmi_swe_hll2_substep(int nx, int ny, float dx, float dy, float dt, float g, float theta, int step,
                          int boundary_conditions, const float* h0, int h0_pitch, const float* hu0, int hu0_pitch,
                          const float* hv0, int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch,
                          float* hv1, int hv1_pitch, float* cfl, const fvmi_launch_t* launch, void* stream) {
    SubstepArgs a{};
    a.nx = nx, a.ny = ny, a.dx = dx, a.dy = dy, a.dt = dt, a.g = g, a.theta = theta, a.step = step;
    a.bc = boundary_conditions;
    a.in[0] = h0, a.in[1] = hu0, a.in[2] = hv0;
    a.in_pitch[0] = h0_pitch, a.in_pitch[1] = hu0_pitch, a.in_pitch[2] = hv0_pitch;
    a.out[0] = h1, a.out[1] = hu1, a.out[2] = hv1;
    a.out_pitch[0] = h1_pitch, a.out_pitch[1] = hu1_pitch, a.out_pitch[2] = hv1_pitch;
    a.cfl = cfl;
    return run_substep(SCHEME_HLL2, a, 3, launch, stream);
}

int fvmi_swe_waf_substep(int nx, int ny, float dx, float dy, float dt, float g, int step, int boundary_conditions,
                         const float* h0, int h0_pitch, const float* hu0, int hu0_pitch, const float* hv0,
                         int hv0_pitch, float* h1, int h1_pitch, float* hu1, int hu1_pitch, float* hv1,
                         int hv1_pitch, float* cfl, const fvmi_launch_t* launch, void* stream) {
    SubstepArgs a{};
    a.nx = nx, a.ny = ny, a.dx = dx, a.dy = dy, a.dt = dt, a.g = g, a.step = step;
    a.bc = boundary_conditions;
    a.in[0] = h0, a.in[1] = hu0, a.in[2] = hv0;
    a.in_pitch[0] = h0_pitch, a.in_pitch[1] = hu0_pitch, a.in_pitch[2] = hv0_pitch;
    a.out[0] = h1, a.out[1] = hu1, a.out[2] = hv1;
    a.out_pitch[0] = h1_pitch, a.out_pitch[1] = hu1_pitch, a.out_pitch[2] = hv1_pitch;
    a.cfl = cfl;
    return run_substep(SCHEME_WAF, a, 3, launch, stream);
}

int fvmi_euler_kp07_dimsplit_substep(int nx, int ny, float dx, float dy, float dt, float g, float gamma, float theta,
                                     int step, int boundary_conditions, const float* rho0, int rho0_pitch,
                                     const float* rho_u0, int rho_u0_pitch, const float* rho_v0, int rho_v0_pitch,
                                     const float* E0, int E0_pitch, float* rho1, int rho1_pitch, float* rho_u1,
                                     int rho_u1_pitch, float* rho_v1, int rho_v1_pitch, float* E1, int E1_pitch,
                                     float* cfl, int x0, int y0, int x1, int y1, const fvmi_launch_t* launch,
                                     void* stream) {
    SubstepArgs a{};
    a.nx = nx, a.ny = ny, a.dx = dx, a.dy = dy, a.dt = dt, a.g = g, a.gamma = gamma, a.theta = theta;
    a.step = step, a.bc = boundary_conditions;
    a.in[0] = rho0, a.in[1] = rho_u0, a.in[2] = rho_v0, a.in[3] = E0;
    a.in_pitch[0] = rho0_pitch, a.in_pitch[1] = rho_u0_pitch, a.in_pitch[2] = rho_v0_pitch, a.in_pitch[3] = E0_pitch;
    a.out[0] = rho1, a.out[1] = rho_u1, a.out[2] = rho_v1, a.out[3] = E1;
    a.out_pitch[0] = rho1_pitch, a.out_pitch[1] = rho_u1_pitch, a.out_pitch[2] = rho_v1_pitch;
    a.out_pitch[3] = E1_pitch;
    a.cfl = cfl;
    a.x0 = x0, a.y0 = y0, a.x1 = x1, a.y1 = y1;
    return run_substep(SCHEME_EULER, a, 4, launch, stream);
}

int fvmi_halo_exchange_periodic(int nx, int ny, int nvars, float* const* fields, const int* pitches, int wrap_x,
                                int wrap_y, void* stream) {
    if (nx < 2 || ny < 2 || nvars < 1 || nvars > 4 || !fields || !pitches)
        return fail(FVMI_ERR_INVALID, "bad halo exchange arguments");
    HaloArgs a{};
    a.nx = nx, a.ny = ny, a.nvars = nvars, a.wrap_x = wrap_x ? 1 : 0, a.wrap_y = wrap_y ? 1 : 0;
    for (int k = 0; k < nvars; ++k) {
        if (!fields[k] || pitches[k] < (nx + 4) * (int)sizeof(float))
            return fail(FVMI_ERR_INVALID, "bad field %d for halo exchange", k);
        a.f[k] = fields[k];
        a.pitch[k] = pitches[k];
    }
    if (!a.wrap_x && !a.wrap_y) return FVMI_OK;
    const int frame = 4 * (nx + 4) + 4 * ny;
    hipLaunchKernelGGL(halo_periodic_kernel, dim3((frame + 255) / 256, nvars), dim3(256), 0, (hipStream_t)stream, a);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

int fvmi_copy_regions_f32(const fvmi_region_copy_t* regions, int count, void* stream) {
    if (count < 0 || (count > 0 && !regions)) return fail(FVMI_ERR_INVALID, "bad region list");
    for (int i = 0; i < count; ++i) {
        const fvmi_region_copy_t& R = regions[i];
        if (R.width < 0 || R.height < 0) return fail(FVMI_ERR_INVALID, "region %d has a negative size", i);
        if (R.width == 0 || R.height == 0) continue;
        if (!R.src || !R.dst) return fail(FVMI_ERR_INVALID, "region %d has a null pointer", i);
        if (R.src_pitch < R.width * (int)sizeof(float) || R.dst_pitch < R.width * (int)sizeof(float))
            return fail(FVMI_ERR_INVALID, "region %d: pitch below the row width", i);
    }
    for (int first = 0; first < count; first += kMaxRegions) {
        const int n = count - first < kMaxRegions ? count - first : kMaxRegions;
        RegionBatch b{};
        long long most = 0;
        for (int i = 0; i < n; ++i) {
            b.r[i] = regions[first + i];
            if (b.r[i].width <= 0 || b.r[i].height <= 0) b.r[i].width = b.r[i].height = 0;
            const long long e = (long long)b.r[i].width * b.r[i].height;
            most = e > most ? e : most;
        }
        if (most == 0) continue;
        long long bx = (most + 255) / 256;
        if (bx > 64) bx = 64;
        hipLaunchKernelGGL(copy_regions_kernel, dim3((unsigned)bx, n), dim3(256), 0, (hipStream_t)stream, b);
        FVMI_HIP(hipGetLastError());
    }
    return FVMI_OK;
}

int fvmi_halo_copy_periodic_x(int nx, int ny, int nvars, float* const* fields, const int* pitches, void* stream) {
    return fvmi_halo_exchange_periodic(nx, ny, nvars, fields, pitches, 1, 0, stream);
}

int fvmi_ensemble_prepare(fvmi_instance_t* table, int n, void* stream) {
    if (!table || n < 0) return fail(FVMI_ERR_INVALID, "bad ensemble table");
    if (n == 0) return FVMI_OK;
    hipLaunchKernelGGL(ensemble_prepare_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, table, n);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

int fvmi_substep_plan(int scheme, int nx, int ny, int x0, int y0, int x1, int y1, const fvmi_launch_t* launch,
                      int cu_count, int* segs, int max_waves, int* n_waves) {
    if (scheme < SCHEME_KP07 || scheme > SCHEME_EULER + 4) return fail(FVMI_ERR_INVALID, "unknown scheme %d", scheme);
    if (nx < 2 || ny < 2 || cu_count < 1 || !n_waves || max_waves < 0 || (max_waves > 0 && !segs))
        return fail(FVMI_ERR_INVALID, "bad plan arguments");
    if (scheme > SCHEME_EULER) {
        // LxF, HLL, FORCE (one ghost cell) and the unsplit KP07 (two): the whole domain
        if (x0 != 0 || y0 != 0 || (x1 != 0 && x1 != nx) || (y1 != 0 && y1 != ny))
            return fail(FVMI_ERR_INVALID, "the unsplit schemes update the whole domain");
        UnsplitArgs u{};
        u.nx = nx, u.ny = ny;
        SubstepArgs shape{};
        int math = 0;
        if (int rc = fill_launch(launch, shape, &math)) return rc;
        u.seg_rows = shape.seg_rows;
        u.waves_per_block = shape.waves_per_block;
        u.xcd_remap = shape.xcd_remap;
        const int grid = plan_unsplit(u, scheme == SCHEME_EULER + 4 ? 2 : 1, cu_count,
                                      unsplit_round_ok(scheme - (SCHEME_EULER + 1)));
        const long long total = (long long)grid * u.waves_per_block;
        if (total > (1LL << 30)) return fail(FVMI_ERR_INVALID, "plan too large");
        *n_waves = (int)total;
        if (total > max_waves) return fail(FVMI_ERR_INVALID, "segs holds %d waves, the plan has %lld", max_waves, total);
        for (int b = 0; b < grid; ++b) {
            for (int w = 0; w < u.waves_per_block; ++w) {
                int strip = -1, row = 0, count = 0;
                if (!unsplit_segment(u, b, grid, w, strip, row, count)) strip = -1, row = 0, count = 0;
                int* o = segs + 3 * ((long long)b * u.waves_per_block + w);
                o[0] = strip, o[1] = row, o[2] = count;
            }
        }
        return FVMI_OK;
    }
    SubstepArgs a{};
    a.nx = nx, a.ny = ny, a.x0 = x0, a.y0 = y0;
    a.x1 = x1 == 0 ? nx : x1;
    a.y1 = y1 == 0 ? ny : y1;
    if (a.x0 < 0 || a.y0 < 0 || a.x1 > nx || a.y1 > ny || a.x0 >= a.x1 || a.y0 >= a.y1)
        return fail(FVMI_ERR_INVALID, "sub-area outside the domain");
    int math = 0;
    if (int rc = fill_launch(launch, a, &math)) return rc;
    a.strip_cols = scheme_strip_cols(scheme, math);
    const int grid = plan_substep(a, cu_count, scheme != SCHEME_WAF);
    const long long total = (long long)grid * a.waves_per_block;
    if (total > (1LL << 30)) return fail(FVMI_ERR_INVALID, "plan too large");
    *n_waves = (int)total;
    if (total > max_waves) return fail(FVMI_ERR_INVALID, "segs holds %d waves, the plan has %lld", max_waves, total);
    for (int b = 0; b < grid; ++b) {
        for (int w = 0; w < a.waves_per_block; ++w) {
            int strip = -1, row = 0, count = 0;
            if (!wave_segment(a, b, grid, w, strip, row, count)) strip = -1, row = 0, count = 0;
            int* o = segs + 3 * ((long long)b * a.waves_per_block + w);
            o[0] = strip, o[1] = row, o[2] = count;
        }
    }
    return FVMI_OK;
}

int fvmi_ensemble_advance(fvmi_instance_t* table, int n, void* stream) {
    if (!table || n < 0) return fail(FVMI_ERR_INVALID, "bad ensemble table");
    if (n == 0) return FVMI_OK;
    hipLaunchKernelGGL(ensemble_advance_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, table, n);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

int fvmi_ensemble_count_active(const fvmi_instance_t* table, int n, int* count_dev, void* stream) {
    if (!table || n < 0 || !count_dev) return fail(FVMI_ERR_INVALID, "bad ensemble count arguments");
    hipLaunchKernelGGL(ensemble_count_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, table, n, count_dev);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

int fvmi_ensemble_substep(int scheme, fvmi_instance_t* table, const int* segs, const int* num_segs_dev, int waves,
                          int step, const fvmi_launch_t* launch, void* stream) {
    if (!table || !segs || !num_segs_dev) return fail(FVMI_ERR_INVALID, "bad ensemble work list");
    if (scheme < SCHEME_KP07 || scheme > SCHEME_EULER) return fail(FVMI_ERR_INVALID, "unknown scheme %d", scheme);
    if (step != 0 && step != 1) return fail(FVMI_ERR_INVALID, "step must be 0 or 1");
    EnsembleArgs e{};
    e.table = table;
    e.segs = reinterpret_cast<const int4*>(segs);
    e.num_segs = num_segs_dev;
    e.step = step;
    e.waves_per_block = (launch && launch->waves_per_block > 0) ? launch->waves_per_block : 4;
    if (e.waves_per_block > 4) return fail(FVMI_ERR_INVALID, "waves_per_block out of range [1, 4]");
    // The work list is sorted longest first: dispatch order (block i on XCD
    // i % 8) spreads it evenly over the XCDs, whereas the contiguous remap of
    // the single-grid kernel would give XCD 0 the longest segments and leave
    // the last XCDs with the empty tail of the list.
    e.xcd_remap = 0;
    const int math = launch ? launch->math : FVMI_MATH_FAST;
    if (waves <= 0) waves = current_cu_count() * 4 * 4;  // four waves per SIMD
    e.num_blocks = (waves + e.waves_per_block - 1) / e.waves_per_block;
    int grid = e.num_blocks;
    if (e.xcd_remap) grid = ((grid + 7) / 8) * 8;
    hipError_t err = (math == FVMI_MATH_EXACT) ? launch_ensemble_exact(scheme, e, grid, (hipStream_t)stream)
                                               : launch_ensemble_fast(scheme, e, grid, (hipStream_t)stream);
    if (err != hipSuccess) return hip_fail(err, "ensemble substep launch");
    return FVMI_OK;
}

int fvmi_stream_begin_capture(void* stream) {
    FVMI_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return FVMI_OK;
}

int fvmi_stream_end_capture(void* stream, void** graph_exec) {
    if (!graph_exec) return fail(FVMI_ERR_INVALID, "graph_exec is NULL");
    hipGraph_t g = nullptr;
    FVMI_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph_exec = ge;
    return FVMI_OK;
}

int fvmi_graph_launch(void* graph_exec, void* stream) {
    FVMI_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return FVMI_OK;
}

int fvmi_graph_destroy(void* graph_exec) {
    if (graph_exec) FVMI_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return FVMI_OK;
}

int fvmi_field_stats(int nx, int ny, int x_halo, int y_halo, const float* field, int pitch, double* sum_out,
                     unsigned int* nonfinite_out, void* stream) {
    if (nx < 1 || ny < 1 || x_halo < 0 || y_halo < 0 || !field || !sum_out || !nonfinite_out ||
        (long long)pitch < (long long)(nx + 2 * x_halo) * (long long)sizeof(float))
        return fail(FVMI_ERR_INVALID, "bad field_stats arguments");
    const int bx = (nx + 255) / 256 < 16 ? (nx + 255) / 256 : 16;
    const int by = ny < 4096 ? ny : 4096;
    const float* interior = reinterpret_cast<const float*>(reinterpret_cast<const char*>(field) +
                                                           (size_t)y_halo * (size_t)pitch) + x_halo;
    hipLaunchKernelGGL(field_stats_kernel, dim3(bx, by), dim3(256), 0, (hipStream_t)stream, nx, ny, interior, pitch,
                       sum_out, nonfinite_out);
    FVMI_HIP(hipGetLastError());
    return FVMI_OK;
}

}  // extern "C"
