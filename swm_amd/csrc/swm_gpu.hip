// C-ABI of the MI355X SWM time-stepping path (declarations: include/swm_gpu.h).
//
// Host side: uploads the static fields (compressing mask-row-uniform ones to two
// values per row), builds per-row coefficient tables with the reference's
// expression order in FP64 on the host, keeps u/v/eta (ping-pong) and the AB
// history (3-slot ring: the reference's G shift becomes a role rotation) on the
// device in one allocation of planes, and launches one fused step kernel per
// time step (steady-state cycles replayed as HIP graphs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "swm_gpu.h"
#include "swm_kernels.h"

using swm::Dyn;
using swm::Fld;
using swm::Statics;

namespace {

// Graph capture on one stream is invalidated by allocations and synchronous
// copies that other host threads make while it is open.  Entry points that
// allocate, free or copy synchronously hold this lock; the graph captures
// hold it only while the capture is open (not while replays run).
std::recursive_mutex g_hip_mu;
using HipLock = std::lock_guard<std::recursive_mutex>;

// Step-kernel shapes, selectable with SWM_GPU_TILE for tuning.  march = 0:
// the LDS tile kernel with tx x ty outputs per 256-thread block.  march = 1:
// the marching-wavefront kernel, four wavefronts per block, each owning 62
// columns and ty rows; ty = 0 picks the smallest ty for which every
// wavefront of the grid is resident at once (one round, no tail).  march = 2:
// the persistent tile kernel (tx x ty tiles, each block a run of tiles sized
// so that all blocks are resident at once).  late = 1: the LDS tile kernel
// issuing the outputs' AB-history loads after the halo stage (faster when the
// grid is only a few rounds of resident blocks).
struct TileShape { int march, tx, ty, late; };
constexpr TileShape kTiles[] = {{0, 64, 4, 0}, {0, 32, 8, 0}, {0, 64, 8, 0}, {0, 128, 4, 0}, {0, 32, 16, 0},
                                {1, 62, 0, 0}, {1, 62, 8, 0}, {1, 62, 16, 0}, {1, 62, 4, 0},
                                {2, 64, 4, 0}, {2, 32, 8, 0}, {0, 32, 8, 1}, {0, 64, 4, 1}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
constexpr int kMaxTracers = 2;

enum FieldId {
    F_HE, F_HU, F_HV, F_IMU, F_IMV, F_IME, F_GSU, F_GSV, F_FXC, F_FYC, F_FEC, F_TAUX, F_TAUY,
    F_ORX, F_OIX, F_ORY, F_OIY, F_ORE, F_OIE, F_UBS, F_VBS, F_ZBS, F_COUNT
};
enum TabId {
    T_CU, T_CV, T_CE, T_CH, T_FH, T_ZA, T_EDU, T_ACE, T_LMU, T_LMV, T_PLLC, T_PLTC, T_TA, T_TB, T_TC,
    T_T9, T_T11, T_T12, T_COUNT
};

struct FieldDev {            // device pointers of an uploaded static field
    void* full = nullptr;
    void* rows = nullptr;
};

// Index mirror of a memoryChunk (src/memchunk_module.f90:37-273).  The chunk's
// slots hold record numbers instead of data (every record stays resident on
// the device) and the time axis is built with the same arithmetic, so each
// query returns the reference's bracketing records and abscissas.
struct ChunkIdx {
    int64_t nrec = 0, chunkSize = 0, fileRec = 1, chunkCounter = 1, nFpC = 0;
    int64_t req_chunk = 0;
    bool isInitialised = false, isPersistent = false, isConstant = false;
    double dt = 0.0, tOffset = 0.0, req_dt = 0.0, last_t = -HUGE_VAL;
    std::vector<int64_t> slot;   // 1-based record held by each slot
    std::vector<double> time;

    // getVarChunkFromDisk, 181-207 (nstart is updated as the reference's fileRec is)
    void var_chunk(size_t at, int64_t& nstart, int64_t len) {
        if (nstart > nrec) nstart = nstart - nFpC * nrec;
        if (nstart == 0) nstart = nrec;
        const int64_t nend = std::min(nstart + len - 1, nrec);
        const int64_t len2 = std::max<int64_t>(nend - nstart, 0) + 1;
        for (int64_t q = 0; q < len2; ++q) slot[at + q] = nstart + q;
        if (len2 < len) {
            int64_t nstart2 = 1;
            var_chunk(at + len2, nstart2, len - len2);
        }
    }
    // getChunkFromDisk, 149-170
    void get_chunk() {
        if (isInitialised) {
            fileRec = fileRec + chunkSize - 1;
            tOffset = tOffset - dt;
        }
        if (!isPersistent || !isInitialised) var_chunk(0, fileRec, chunkSize);
        if (!isConstant)
            for (int64_t i = 0; i < chunkSize; ++i) {
                time[i] = tOffset;
                tOffset = tOffset + dt;
            }
        chunkCounter = 1;
    }
    // initMemChunk, 77-115
    void init(int64_t n, int64_t cs, double rec_dt) {
        *this = ChunkIdx();
        nrec = n;
        req_chunk = cs;
        req_dt = rec_dt;
        chunkSize = cs;
        if (n <= 1) {
            isPersistent = isConstant = true;
            chunkSize = 1;
        } else if (cs >= n + 1) {
            isPersistent = true;
            chunkSize = n + 1;
        }
        dt = isConstant ? 0.0 : rec_dt;
        slot.assign((size_t)chunkSize, 1);
        time.assign((size_t)chunkSize, 0.0);
        get_chunk();
        nFpC = std::max<int64_t>((int64_t)std::floor((float)chunkSize / (float)nrec), 1);
        isInitialised = true;
    }
    // getChunkData, 214-243: records r0, r1 (0-based) and the interpolation
    // scalars a = t - time(cc), b = time(cc+1) - time(cc).  Model time only
    // moves forward in the reference; an earlier time (a context resumed or
    // restored at a smaller itt) starts from a fresh chunk.
    void query(double t, int64_t& r0, int64_t& r1, bool& exact, double& a, double& b) {
        if (t < last_t) init(nrec, req_chunk, req_dt);
        last_t = t;
        if (isConstant) {
            r0 = r1 = slot[0] - 1;
            exact = true;
            a = 0.0;
            b = 1.0;
            return;
        }
        if (chunkCounter >= chunkSize) get_chunk();
        while (time[(size_t)chunkCounter] < t) {
            chunkCounter = chunkCounter + 1;
            if (chunkCounter >= chunkSize) get_chunk();
        }
        const double x0 = time[(size_t)chunkCounter - 1];
        r0 = slot[(size_t)chunkCounter - 1] - 1;
        if (t == x0) {
            r1 = r0;
            exact = true;
            a = 0.0;
            b = 1.0;
            return;
        }
        r1 = slot[(size_t)chunkCounter] - 1;
        exact = false;
        a = t - x0;
        b = time[(size_t)chunkCounter] - x0;
    }
};

struct FStream {
    int kind = 0, comp = 0;      // SWM_STREAM_*, SWM_COMP_*
    void* rec = nullptr;         // nrec pitched planes of the context's type
    ChunkIdx idx;
};

template <typename T>
using SStepArr = swm::SStep<T>[2];

struct Tracer {
    double kappa = 0;
    bool cbuf_valid = false;
    FieldDev C0, gC, impl;
};

}  // namespace

struct swm_ctx {
    swm_desc d{};
    int Nx = 0, Ny = 0, P = 0;
    bool fp32 = false;
    size_t esz = 8;
    std::string err;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_total_ms = 0, last_kernel_ms = 0;
    std::vector<void*> allocs;
    bool has_osc = false, has_taux = false, has_tauy = false;
    // state planes (swm::PL_*), one allocation
    void* state = nullptr;
    long long plane = 0;         // elements per plane
    int cur = 0;                 // ping-pong index of level N0
    int ring = 0;                // slot holding G(1) (oldest); G(2) = ring+1, newest = ring+2
    int64_t itt = 0;
    int64_t itt_upload = 0;      // itt of the last state hand-over: the other level holds the
                                 // state before the last step only when itt > itt_upload
    uint8_t* mask = nullptr;
    std::vector<int8_t> hmask[4];
    FieldDev fields[F_COUNT];
    void* tab[T_COUNT] = {};
    void* statics = nullptr;     // device Statics<T>
    void* rowtab = nullptr;      // device [Ny][RT_N] packed per-row values (Statics::rowtab)
    double ab[5][5] = {};
    std::vector<Tracer> trc;
    std::vector<FStream> streams;   // time-interpolated forcing streams, in processing order
    FStream dyn[3];                 // DYNFROMFILE datasets of u, v, eta (rec null when absent)
    int tile = 0;
    // Steady steps replayed from captured six-step graphs (SWM_GPU_GRAPH=1,
    // swm_gpu_set_launch_mode, or the choice of swm_gpu_tune).
    // Off by default: on MI355X direct launches measured faster at every grid
    // size from 128^2 to 1024^2 (e.g. 1024^2 FP64: 46.34 against 46.95 us per
    // step; 128^2: 9.38 against 9.89 us), the graphs' per-launch overhead
    // exceeding what they save in host work.
    bool use_graph = false;
    int num_cu = 0;              // compute units of the device
    double* sums_buf = nullptr;  // conservation-sum partials and result
    double* norm_buf = nullptr;  // ||B|| partials and result of the device-resident SOR
    bool chi_valid[2] = {};      // sor_buf[w][0] holds a solution to warm-start from
    int sor_every = 1;           // fixed-omega mode: convergence test every sor_every sweeps
    bool sor_blocked = true;     // fixed-omega mode: several sweeps per launch when sor_every allows
    int sor_halo = -1;           // LDS halo of the blocked sweeps (2 x sweeps per launch); -1: from sor_every
    int sor_rows = 16;           // rows of the blocked sweeps' 64-column tile (16, or 32 with <= 2 sweeps per launch)
    double ms_direct = -1.0, ms_graph = -1.0;   // per-step times of the last tuning run's launch forms
    struct Avg {                 // averaging buffer of a diag task
        int field;               // 0 u, 1 v, 2 eta, 3 CH of tracer `tracer`
        int tracer;
        int square;
        double* buf;             // FP64, pitched like the fields
        double weight;           // bufferCount
    };
    std::vector<Avg> avgs;
    int chi_last = -1;           // precision slot of the last solution
    double* ce64 = nullptr;      // cos(lat_eta) rows in FP64
    std::vector<double> ce_host;
    long march_slots = 0;        // resident wavefronts of the marching kernel (0: not queried)
    long persist_slots[2] = {};  // resident blocks of the persistent tile kernels (0: not queried)
    hipGraphExec_t graphs[6] = {};   // steady-state 6-step cycles, keyed by (cur, ring)
    // diagnostics scratch
    void* diag[14] = {};
    void* tdiag[kMaxTracers][3] = {};
    void* scratch = nullptr;     // one plane (divergence output)
    double* psi_buf[4] = {};     // FP64 planes: pu, pv, column sums, psi
    // SOR
    void* sor_tab[5] = {};      // c1, c3, c4, c5 and 1/c5
    void* sor_relax = nullptr;
    int sor_nrelax = 0;
    double* sor_partial = nullptr;
    int* sor_done = nullptr;
    long long* sor_iters = nullptr;
    double* sor_norm = nullptr;
    void* sor_buf[2][2] = {};    // [fp32][chi|B]
    void* sor_alt[2] = {};       // second chi buffer of the blocked sweeps (ping-pong)
    void* sor_mid[2][3][2] = {}; // reference mode, K sweeps per launch: chi after each of the first K - 1, by launch parity
    double* sor_partial_mid = nullptr;   // and their partial sums (3 sweeps x two sets by launch parity)
    void* sor_ftab[2][5] = {};   // coefficient tables in the sweep precision
    void* dza = nullptr;
    // elliptic-solver build (SWM_ELLIPTIC_SOLVER): the reference-mode velocity
    // potential kept between steps (calc_lib::chi), solved at most once per
    // step when a diagnostic asks for it, and the non-divergent flow
    void* chi_ref = nullptr;     // chi, context precision
    void* chi_refB = nullptr;    // its right-hand side
    double* nd_buf[2] = {};      // u_nd, v_nd (FP64)
    double* gza = nullptr;       // bc / (A cos_u dLambda) per row, ocean and land
    std::vector<double> cu_host; // cos(lat_u) rows
    int64_t chi_ref_itt = -1;    // itt of the last solve (-1: none)
    int64_t chi_ref_iters = 0;
    double chi_ref_anorm = 0.0;
    swm_batch* batch = nullptr;  // batch this context is attached to
};

namespace {

int fail(swm_ctx* c, const std::string& msg, int code = 2) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), std::string(#call) + ": " + hipGetErrorString(e_), 3);          \
    } while (0)

// Host transfers.  Every copy between host and device memory passes through
// one pinned (page-locked) staging buffer of the library and is a plain 1-D
// DMA copy on the stream that orders it with the kernels (the contexts' and
// batches' streams are non-blocking, so legacy null-stream copies would not
// be ordered with them), waited for before the call returns.  Caller memory
// is only ever touched by host memcpy: the DMA engines never see pageable
// memory, and pitched planes are packed and unpacked on the host.
constexpr size_t kStageMax = (size_t)32 << 20;
char* g_stage = nullptr;
size_t g_stage_bytes = 0;

// A staging buffer of at least min(want, kStageMax) bytes (callers hold
// g_hip_mu).  grow = false never allocates: it fails when the buffer (made by
// the first upload of any context, at least 1 MiB) is missing or smaller.
hipError_t stage_get(size_t want, char** p, size_t* n, bool grow = true) {
    want = std::max<size_t>(std::min(want, kStageMax), (size_t)1 << 20);
    if (g_stage_bytes < want && !grow) return hipErrorOutOfMemory;
    if (g_stage_bytes < want) {
        if (g_stage) (void)hipHostFree(g_stage);
        g_stage = nullptr;
        g_stage_bytes = 0;
        hipError_t e = hipHostMalloc((void**)&g_stage, want, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        g_stage_bytes = want;
    }
    *p = g_stage;
    *n = g_stage_bytes;
    return hipSuccess;
}

hipError_t h2d(hipStream_t s, void* dst, const void* src, size_t bytes, bool grow = true) {
    HipLock lock(g_hip_mu);
    char* st = nullptr;
    size_t cap = 0;
    hipError_t e = stage_get(bytes, &st, &cap, grow);
    for (size_t off = 0; e == hipSuccess && off < bytes; off += cap) {
        const size_t len = std::min(cap, bytes - off);
        std::memcpy(st, (const char*)src + off, len);
        e = hipMemcpyAsync((char*)dst + off, st, len, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    return e;
}

hipError_t d2h(hipStream_t s, void* dst, const void* src, size_t bytes, bool grow = true) {
    HipLock lock(g_hip_mu);
    char* st = nullptr;
    size_t cap = 0;
    hipError_t e = stage_get(bytes, &st, &cap, grow);
    for (size_t off = 0; e == hipSuccess && off < bytes; off += cap) {
        const size_t len = std::min(cap, bytes - off);
        e = hipMemcpyAsync(st, (const char*)src + off, len, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess) std::memcpy((char*)dst + off, st, len);
    }
    return e;
}

// Rows of `width` bytes: host rows spaced spitch, device rows spaced dpitch
// (dpitch >= width; the padding bytes of each device row are written as zero).
hipError_t h2d_rows(hipStream_t s, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                    size_t rows) {
    HipLock lock(g_hip_mu);
    char* st = nullptr;
    size_t cap = 0;
    hipError_t e = stage_get(rows * dpitch, &st, &cap);
    const size_t per = std::max<size_t>(1, cap / dpitch);
    if (e == hipSuccess && per * dpitch > cap) return hipErrorInvalidValue;   // one row must fit
    for (size_t r0 = 0; e == hipSuccess && r0 < rows; r0 += per) {
        const size_t nr = std::min(per, rows - r0);
        for (size_t r = 0; r < nr; ++r) {
            std::memcpy(st + r * dpitch, (const char*)src + (r0 + r) * spitch, width);
            if (dpitch > width) std::memset(st + r * dpitch + width, 0, dpitch - width);
        }
        e = hipMemcpyAsync((char*)dst + r0 * dpitch, st, nr * dpitch, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    return e;
}

hipError_t d2h_rows(hipStream_t s, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                    size_t rows) {
    HipLock lock(g_hip_mu);
    char* st = nullptr;
    size_t cap = 0;
    hipError_t e = stage_get(rows * spitch, &st, &cap);
    const size_t per = std::max<size_t>(1, cap / spitch);
    if (e == hipSuccess && per * spitch > cap) return hipErrorInvalidValue;
    for (size_t r0 = 0; e == hipSuccess && r0 < rows; r0 += per) {
        const size_t nr = std::min(per, rows - r0);
        e = hipMemcpyAsync(st, (const char*)src + r0 * spitch, nr * spitch, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess)
            for (size_t r = 0; r < nr; ++r) std::memcpy((char*)dst + (r0 + r) * dpitch, st + r * spitch, width);
    }
    return e;
}

hipError_t copy_sync(swm_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (kind == hipMemcpyHostToDevice) return h2d(c->stream, dst, src, bytes);
    if (kind == hipMemcpyDeviceToHost) return d2h(c->stream, dst, src, bytes);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

hipError_t copy2d_sync(swm_ctx* c, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                       size_t rows, hipMemcpyKind kind) {
    if (kind == hipMemcpyHostToDevice) return h2d_rows(c->stream, dst, dpitch, src, spitch, width, rows);
    if (kind == hipMemcpyDeviceToHost) return d2h_rows(c->stream, dst, dpitch, src, spitch, width, rows);
    hipError_t e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, kind, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

hipError_t zero_sync(swm_ctx* c, void* dst, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

void* dalloc(swm_ctx* c, size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    c->allocs.push_back(p);
    return p;
}

void dfree(swm_ctx* c, void* p) {
    if (!p) return;
    auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
    if (it != c->allocs.end()) c->allocs.erase(it);
    (void)hipFree(p);
}

// Upload a host double vector as the context's element type.
void* upload_vec(swm_ctx* c, const std::vector<double>& h) {
    void* p = dalloc(c, h.size() * c->esz);
    if (!p) return nullptr;
    if (c->fp32) {
        std::vector<float> f(h.begin(), h.end());
        if (copy_sync(c, p, f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    } else {
        if (copy_sync(c, p, h.data(), h.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    }
    return p;
}

// Upload a host (Nx,Ny) double array into a pitched device plane of the context's type.
int upload_grid(swm_ctx* c, void* dst, const double* src) {
    const size_t n = (size_t)c->Nx * c->Ny;
    if (c->fp32) {
        std::vector<float> f(n);
        for (size_t k = 0; k < n; ++k) f[k] = (float)src[k];
        HIPCHK(c, copy2d_sync(c, dst, c->P * 4, f.data(), c->Nx * 4, c->Nx * 4, c->Ny, hipMemcpyHostToDevice));
    } else {
        HIPCHK(c, copy2d_sync(c, dst, c->P * 8, src, c->Nx * 8, c->Nx * 8, c->Ny, hipMemcpyHostToDevice));
    }
    return 0;
}

int download_grid(swm_ctx* c, double* dst, const void* src) {
    const size_t n = (size_t)c->Nx * c->Ny;
    if (c->fp32) {
        std::vector<float> f(n);
        HIPCHK(c, copy2d_sync(c, f.data(), c->Nx * 4, src, c->P * 4, c->Nx * 4, c->Ny, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) dst[k] = f[k];
    } else {
        HIPCHK(c, copy2d_sync(c, dst, c->Nx * 8, src, c->P * 8, c->Nx * 8, c->Ny, hipMemcpyDeviceToHost));
    }
    return 0;
}

// bfloat16 low part of the FP32 state, rounded to nearest even (the step
// kernels' lo_enc / lo_dec)
uint16_t lo_enc_host(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
double lo_dec_host(uint16_t h) {
    const uint32_t b = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

bool carries_low_part(const swm_ctx* c) { return c->fp32 && (c->d.flags & SWM_FP32_LOW_PART); }

// A state field: FP64 as is; FP32 as hi = (float)x in the field's plane and,
// with SWM_FP32_LOW_PART, lo = x - hi (bfloat16) in its low-part plane (the
// low-part planes stay zero without it).
int upload_state_grid(swm_ctx* c, void* dst, void* lo_dst, const double* src) {
    if (!carries_low_part(c)) return upload_grid(c, dst, src);
    const size_t n = (size_t)c->Nx * c->Ny;
    std::vector<float> f(n);
    std::vector<uint16_t> l(n);
    for (size_t k = 0; k < n; ++k) {
        f[k] = (float)src[k];
        l[k] = std::isfinite(f[k]) ? lo_enc_host((float)(src[k] - (double)f[k])) : (uint16_t)0;
    }
    HIPCHK(c, copy2d_sync(c, dst, c->P * 4, f.data(), c->Nx * 4, c->Nx * 4, c->Ny, hipMemcpyHostToDevice));
    HIPCHK(c, copy2d_sync(c, lo_dst, c->P * 2, l.data(), c->Nx * 2, c->Nx * 2, c->Ny, hipMemcpyHostToDevice));
    return 0;
}

int download_state_grid(swm_ctx* c, double* dst, const void* src, const void* lo_src) {
    if (!carries_low_part(c) || !lo_src) return download_grid(c, dst, src);
    const size_t n = (size_t)c->Nx * c->Ny;
    std::vector<float> f(n);
    std::vector<uint16_t> l(n);
    HIPCHK(c, copy2d_sync(c, f.data(), c->Nx * 4, src, c->P * 4, c->Nx * 4, c->Ny, hipMemcpyDeviceToHost));
    HIPCHK(c, copy2d_sync(c, l.data(), c->Nx * 2, lo_src, c->P * 2, c->Nx * 2, c->Ny, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; ++k) dst[k] = (double)f[k] + lo_dec_host(l[k]);
    return 0;
}

void* alloc_grid(swm_ctx* c) {
    void* p = dalloc(c, (size_t)c->Ny * c->P * c->esz);
    if (p && zero_sync(c, p, (size_t)c->Ny * c->P * c->esz) != hipSuccess) return nullptr;
    return p;
}

void* plane_ptr(swm_ctx* c, int k) { return (char*)c->state + (size_t)k * c->plane * c->esz; }

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

// Upload a static field.  NULL -> the constant default on every point.  If each
// row has one value on the mask's ocean points and one on its land points the
// field is stored as two values per row, else as a full pitched array.
// The implicit damping factor impl = 1 + dt*gamma as the step kernels take
// it: FP64 contexts hold impl and divide by it; FP32 contexts hold
// kap = 1 - 1/impl and form x - x*kap, which keeps the drag to float's
// relative precision (impl itself, rounded to float, would keep only the
// few bits of dt*gamma that lie above ulp(1)).  buf holds the FP32 values.
const double* damp_static(swm_ctx* c, const double* impl, std::vector<double>& buf) {
    if (!c->fp32 || !impl) return impl;
    const size_t n = (size_t)c->Nx * c->Ny;
    buf.resize(n);
    for (size_t k = 0; k < n; ++k) buf[k] = (impl[k] - 1.0) / impl[k];
    return buf.data();
}

int upload_field(swm_ctx* c, FieldDev* out, const double* h, double dflt, const int8_t* mask) {
    const int Nx = c->Nx, Ny = c->Ny;
    std::vector<double> rows(2 * (size_t)Ny, dflt);
    bool compress = true;
    if (h) {
        for (int j = 0; j < Ny && compress; ++j) {
            bool have[2] = {false, false};
            for (int i = 0; i < Nx; ++i) {
                const double x = h[(size_t)j * Nx + i];
                const int s = mask[(size_t)j * Nx + i] ? 0 : 1;
                if (!have[s]) { rows[2 * j + s] = x; have[s] = true; }
                else if (!same_bits(rows[2 * j + s], x)) { compress = false; break; }
            }
        }
    }
    if (compress) {
        out->rows = upload_vec(c, rows);
        out->full = nullptr;
        if (!out->rows) return fail(c, "field upload failed", 3);
    } else {
        void* p = alloc_grid(c);
        if (!p) return fail(c, "out of device memory", 3);
        int rc = upload_grid(c, p, h);
        if (rc) return rc;
        out->full = p;
        out->rows = nullptr;
    }
    return 0;
}

template <typename T>
Fld<T> F(const FieldDev& f) {
    return Fld<T>{(const T*)f.full, (const T*)f.rows};
}

// Per-row coefficient tables, each built with the reference's expression.
int build_tables(swm_ctx* c) {
    const swm_desc& d = c->d;
    const int Ny = c->Ny;
    const double A = d.A, dL = d.dLambda, dT = d.dTheta, Ah = d.Ah;
    const double *cu = d.grid[SWM_GRID_U].cos_lat, *cv = d.grid[SWM_GRID_V].cos_lat,
                 *ce = d.grid[SWM_GRID_ETA].cos_lat, *cH = d.grid[SWM_GRID_H].cos_lat;
    const double *tu = d.grid[SWM_GRID_U].tan_lat, *tv = d.grid[SWM_GRID_V].tan_lat,
                 *te = d.grid[SWM_GRID_ETA].tan_lat, *tH = d.grid[SWM_GRID_H].tan_lat;
    std::vector<double> t[T_COUNT];
    t[T_CU].assign(cu, cu + Ny);
    t[T_CV].assign(cv, cv + Ny);
    t[T_CE].assign(ce, ce + Ny);
    t[T_CH].assign(cH, cH + Ny);
    t[T_FH].assign(d.grid[SWM_GRID_H].f, d.grid[SWM_GRID_H].f + Ny);
    t[T_ZA].resize(2 * Ny); t[T_EDU].resize(Ny); t[T_ACE].resize(Ny);
    t[T_LMU].resize(6 * Ny); t[T_LMV].resize(6 * Ny); t[T_PLLC].resize(12 * Ny); t[T_PLTC].resize(12 * Ny);
    t[T_TA].resize(2 * Ny); t[T_TB].resize(2 * Ny); t[T_TC].resize(2 * Ny);
    t[T_T9].resize(Ny); t[T_T11].resize(Ny); t[T_T12].resize(Ny);
    for (int j = 0; j < Ny; ++j) {
        const int jp = (j + 1) % Ny;
        // pder_zonal3D factor bc/(A*cos*dLambda) on the H grid (src/calc_lib.f90:439)
        t[T_ZA][2 * j + 0] = 1.0 / (A * cH[j] * dL);
        t[T_ZA][2 * j + 1] = d.lbc / (A * cH[j] * dL);
        t[T_EDU][j] = A * cu[j] * dL;
        t[T_ACE][j] = A * ce[j];
        for (int ot = 1; ot <= 2; ++ot) {
            // SWM_LateralMixing_init (src/swm_lateralmixing_module.f90:88-94), bc = 1 on ocean points
            const double bc = 1.0;
            double* lu = &t[T_LMU][(j * 2 + ot - 1) * 3];
            lu[0] = -(bc / A / cu[j] / dL);
            lu[1] = -(bc / A / dT - 2.0 * tu[j] / A / (double)ot);
            lu[2] = -(-(bc / A / dT) - 2.0 * tu[j] / A / (double)ot);
            double* lv = &t[T_LMV][(j * 2 + ot - 1) * 3];
            lv[0] = -(bc / A / cv[j] / dL);
            lv[1] = -(-(bc / A / dT) + 2.0 * tv[j] / A / (double)ot);
            lv[2] = -(bc / A / dT + 2.0 * tv[j] / A / (double)ot);
        }
        for (int a1 = 0; a1 < 2; ++a1)
            for (int a0 = 0; a0 < 2; ++a0) {
                // _init_p_coefficients (src/swm_lateralmixing_module.f90:135-142), bc = 1 on ocean points
                const double bc = 1.0;
                int ot = std::max(1, a0 + a1);
                double* pl = &t[T_PLLC][(j * 4 + a1 * 2 + a0) * 3];   // a1 = ov(j+1), a0 = ov(j)
                pl[0] = -(Ah * bc / dL / A / ce[j]);
                pl[1] = Ah / A * (bc / dT + (double)a1 * te[j] / (double)ot);
                pl[2] = Ah / A * (-(bc / dT) + (double)a0 * te[j] / (double)ot);
                double* pt = &t[T_PLTC][(j * 4 + a1 * 2 + a0) * 3];   // a1 = ou(j), a0 = ou(j-1)
                pt[0] = -(Ah * bc / dL / A / cH[j]);
                pt[1] = -(Ah / A * (bc / dT + (double)a1 * tH[j] / (double)ot));
                pt[2] = -(Ah / A * (-(bc / dT) + (double)a0 * tH[j] / (double)ot));
            }
        // TRC_initCoeffs magnitudes (src/tracer_module.f90:318-338)
        for (int den = 1; den <= 2; ++den) {
            t[T_TA][2 * j + den - 1] = 1.0 * 1.0 * 1.0 / (double)den / A / dL / ce[j];
            t[T_TB][2 * j + den - 1] = 1.0 * 1.0 * cv[jp] * 1.0 / (double)den / A / dT / ce[j];
            t[T_TC][2 * j + den - 1] = 1.0 * 1.0 * cv[j] * 1.0 / (double)den / A / dT / ce[j];
        }
        t[T_T9][j] = 1.0 * 1.0 / (dL * dL) / (A * A) / cu[j] / ce[j];
        t[T_T11][j] = 1.0 * 1.0 * cv[jp] / (dT * dT) / (A * A) / ce[j];
        t[T_T12][j] = 1.0 * 1.0 * cv[j] / (dT * dT) / (A * A) / ce[j];
    }
    for (int k = 0; k < T_COUNT; ++k) {
        c->tab[k] = upload_vec(c, t[k]);
        if (!c->tab[k]) return fail(c, "table upload failed", 3);
    }
    // computeDivergence zonal factor on the eta grid
    std::vector<double> dza(2 * Ny);
    for (int j = 0; j < Ny; ++j) {
        dza[2 * j] = 1.0 / (A * ce[j] * dL);
        dza[2 * j + 1] = d.lbc / (A * ce[j] * dL);
    }
    c->dza = upload_vec(c, dza);
    // SOR coefficients (src/ElSolv_SOR.f90:54-62)
    std::vector<double> c1(Ny), c3(Ny), c4(Ny), c5(12 * (size_t)Ny);
    for (int j = 0; j < Ny; ++j) {
        const int jp = (j + 1) % Ny;
        double x = A * ce[j] * dL;
        x = x * x;
        const double y = A * A * ce[j] * (dT * dT);
        c1[j] = 1.0 / x;
        c3[j] = cv[jp] * 1.0 / y;
        c4[j] = cv[j] * 1.0 / y;
        for (int su = 0; su <= 2; ++su)
            for (int ovjp = 0; ovjp < 2; ++ovjp)
                for (int ov = 0; ov < 2; ++ov)
                    c5[((j * 3 + su) * 2 + ovjp) * 2 + ov] =
                        (double)(-su) / x + (-(cv[jp] * (double)ovjp) - cv[j] * (double)ov) / y;
    }
    c->sor_tab[0] = upload_vec(c, c1);
    c->sor_tab[1] = upload_vec(c, c3);
    c->sor_tab[2] = upload_vec(c, c4);
    c->sor_tab[3] = upload_vec(c, c5);
    std::vector<double> c5r(c5.size());   // correctly rounded reciprocals for dvr (1/-0 = -inf)
    for (size_t k = 0; k < c5.size(); ++k) c5r[k] = 1.0 / c5[k];
    c->sor_tab[4] = upload_vec(c, c5r);
    if (!c->dza || !c->sor_tab[0] || !c->sor_tab[1] || !c->sor_tab[2] || !c->sor_tab[3] || !c->sor_tab[4])
        return fail(c, "table upload failed", 3);
    return 0;
}

// (Re)write the device Statics<T> from the context.
template <typename T>
int write_statics_t(swm_ctx* c) {
    const swm_desc& d = c->d;
    Statics<T> a{};
    a.Nx = c->Nx; a.Ny = c->Ny; a.P = c->P;
    a.mix = (d.flags & SWM_LATERAL_MIXING) ? 1 : 0;
    a.carry = (c->fp32 && (d.flags & SWM_FP32_LOW_PART)) ? 1 : 0;
    a.quad = (d.flags & SWM_QUADRATIC_BOTTOM_FRICTION) ? 1 : 0;
    a.barotropic = (d.flags & SWM_BAROTROPIC) ? 1 : 0;
    a.has_tau_scale = (d.flags & SWM_TAU_SCALE) ? 1 : 0;
    a.has_taux = c->has_taux;
    a.has_tauy = c->has_tauy;
    a.has_oy = d.osc_re[1] ? 1 : 0;
    a.has_oe = d.osc_re[2] ? 1 : 0;
    a.has_ox = d.osc_re[0] ? 1 : 0;
    a.G = (T)d.G; a.A = (T)d.A; a.dL = (T)d.dLambda; a.dT = (T)d.dTheta; a.RHO0 = (T)d.RHO0;
    a.minD = (T)d.minD; a.lbc = (T)d.lbc; a.tau_scale = (T)d.tau_scale; a.AdT = (T)(d.A * d.dTheta);
    a.rdL = (T)(1.0 / d.dLambda); a.rdT = (T)(1.0 / d.dTheta); a.rAdT = (T)(1.0 / (d.A * d.dTheta));
    a.eps = std::is_same<T, double>::value ? (T)2.220446049250313e-16 : (T)1.1920929e-07f;
    a.mask = c->mask;
    a.He = F<T>(c->fields[F_HE]); a.Hu = F<T>(c->fields[F_HU]); a.Hv = F<T>(c->fields[F_HV]);
    a.imu = F<T>(c->fields[F_IMU]); a.imv = F<T>(c->fields[F_IMV]); a.ime = F<T>(c->fields[F_IME]);
    a.gsu = F<T>(c->fields[F_GSU]); a.gsv = F<T>(c->fields[F_GSV]);
    a.Fxc = F<T>(c->fields[F_FXC]); a.Fyc = F<T>(c->fields[F_FYC]); a.Fec = F<T>(c->fields[F_FEC]);
    a.taux = F<T>(c->fields[F_TAUX]); a.tauy = F<T>(c->fields[F_TAUY]);
    a.oRx = F<T>(c->fields[F_ORX]); a.oIx = F<T>(c->fields[F_OIX]);
    a.oRy = F<T>(c->fields[F_ORY]); a.oIy = F<T>(c->fields[F_OIY]);
    a.oRe = F<T>(c->fields[F_ORE]); a.oIe = F<T>(c->fields[F_OIE]);
    a.ubs = F<T>(c->fields[F_UBS]); a.vbs = F<T>(c->fields[F_VBS]); a.zbs = F<T>(c->fields[F_ZBS]);
    a.cu = (const T*)c->tab[T_CU]; a.cv = (const T*)c->tab[T_CV]; a.ce = (const T*)c->tab[T_CE];
    a.cH = (const T*)c->tab[T_CH]; a.fH = (const T*)c->tab[T_FH]; a.za = (const T*)c->tab[T_ZA];
    a.edu = (const T*)c->tab[T_EDU]; a.ace = (const T*)c->tab[T_ACE];
    a.lmu = (const T*)c->tab[T_LMU]; a.lmv = (const T*)c->tab[T_LMV];
    a.pllc = (const T*)c->tab[T_PLLC]; a.pltc = (const T*)c->tab[T_PLTC];
    a.tA = (const T*)c->tab[T_TA]; a.tB = (const T*)c->tab[T_TB]; a.tC = (const T*)c->tab[T_TC];
    a.t9 = (const T*)c->tab[T_T9]; a.t11 = (const T*)c->tab[T_T11]; a.t12 = (const T*)c->tab[T_T12];
    for (size_t n = 0; n < c->trc.size(); ++n) {
        a.kappa[n] = (T)c->trc[n].kappa;
        a.C0[n] = F<T>(c->trc[n].C0);
        a.gC[n] = F<T>(c->trc[n].gC);
        a.timpl[n] = F<T>(c->trc[n].impl);
        for (int k = 0; k < 3; ++k) a.tdiag[n][k] = (T*)c->tdiag[n][k];
    }
    for (int k = 0; k < 14; ++k) a.diag[k] = (T*)c->diag[k];
    a.nstream = (int)c->streams.size();
    a.sdiv = (a.barotropic || d.variant == SWM_FULLY_NONLINEAR) ? 1 : 0;
    for (size_t n = 0; n < c->streams.size(); ++n) {
        a.skind[n] = c->streams[n].kind == SWM_STREAM_WIND ? 0 : 1;
        a.scomp[n] = c->streams[n].comp;
        a.srec[n] = (const T*)c->streams[n].rec;
    }
    a.dyn_u = (const T*)c->dyn[0].rec;
    a.dyn_v = (const T*)c->dyn[1].rec;
    static_assert(swm::RT_N >= swm::RT_TRC + swm::RT_PER_TRC * kMaxTracers, "rowtab has room for every tracer");
    if (!c->statics && !(c->statics = dalloc(c, sizeof(Statics<T>)))) return fail(c, "out of device memory", 3);
    if (!c->rowtab && !(c->rowtab = dalloc(c, (size_t)c->Ny * swm::RT_N * sizeof(T)))) return fail(c, "out of device memory", 3);
    a.rowtab = (const T*)c->rowtab;
    HIPCHK(c, copy_sync(c, c->statics, &a, sizeof(a), hipMemcpyHostToDevice));
    const int nrt = c->Ny * swm::RT_N;
    swm::pack_rowtab<T><<<(nrt + 255) / 256, 256, 0, c->stream>>>((const Statics<T>*)c->statics, (T*)c->rowtab, c->Ny);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int write_statics(swm_ctx* c) { return c->fp32 ? write_statics_t<float>(c) : write_statics_t<double>(c); }

template <typename T>
int march_blocks_per_cu(swm_ctx* c) {
    const int var = c->d.variant;
    const int nt = (int)c->trc.size();
    int nb = 0;
    hipError_t e = hipSuccess;
#define SWM_OCC(V, N) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, swm::march_kernel<T, V, N>, 256, 0)
    if (var == SWM_FULLY_NONLINEAR) {
        if (nt == 0) SWM_OCC(swm::VAR_FN, 0);
        else if (nt == 1) SWM_OCC(swm::VAR_FN, 1);
        else SWM_OCC(swm::VAR_FN, 2);
    } else {
        if (nt == 0) SWM_OCC(swm::VAR_SOR, 0);
        else if (nt == 1) SWM_OCC(swm::VAR_SOR, 1);
        else SWM_OCC(swm::VAR_SOR, 2);
    }
#undef SWM_OCC
    return e == hipSuccess ? nb : 0;
}

template <typename T, int TX, int TY>
int persist_blocks_per_cu(swm_ctx* c) {
    const int var = c->d.variant;
    const int nt = (int)c->trc.size();
    int nb = 0;
    hipError_t e = hipSuccess;
#define SWM_OCC(V, N) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, swm::step_kernel_p<T, V, N, TX, TY>, 256, 0)
    if (var == SWM_FULLY_NONLINEAR) {
        if (nt == 0) SWM_OCC(swm::VAR_FN, 0);
        else if (nt == 1) SWM_OCC(swm::VAR_FN, 1);
        else SWM_OCC(swm::VAR_FN, 2);
    } else {
        if (nt == 0) SWM_OCC(swm::VAR_SOR, 0);
        else if (nt == 1) SWM_OCC(swm::VAR_SOR, 1);
        else SWM_OCC(swm::VAR_SOR, 2);
    }
#undef SWM_OCC
    return e == hipSuccess ? nb : 0;
}

// Rows per wavefront of the marching kernel: the smallest count (at least 4)
// whose grid fits the device's resident wavefront slots in one round.
int march_rows(swm_ctx* c, int ntx) {
    if (c->march_slots <= 0) return 8;
    const long segs = std::max<long>(1, c->march_slots / ntx);
    const int ty = (int)((c->Ny + segs - 1) / segs);
    return std::max(4, ty);
}

template <typename T>
Dyn<T> make_dyn(swm_ctx* c, int tile) {
    Dyn<T> d{};
    d.s = (const Statics<T>*)c->statics;
    d.state = (T*)c->state;
    d.plane = c->plane;
    const TileShape ts = kTiles[tile];
    d.ntx = (c->Nx + ts.tx - 1) / ts.tx;
    if (ts.march == 1 && !ts.ty && c->march_slots <= 0)
        c->march_slots = (long)march_blocks_per_cu<T>(c) * 4 * c->num_cu;
    if (ts.march == 2) {   // persistent tiles: nblk counts tiles, the grid is nblk / tpb blocks
        d.ty = ts.ty;
        d.nblk = d.ntx * ((c->Ny + ts.ty - 1) / ts.ty);
        const int k = ts.tx == 64 ? 0 : 1;
        if (c->persist_slots[k] <= 0) {
            const int nb = ts.tx == 64 ? persist_blocks_per_cu<T, 64, 4>(c) : persist_blocks_per_cu<T, 32, 8>(c);
            c->persist_slots[k] = (long)std::max(1, nb) * std::max(1, c->num_cu);
        }
        d.tpb = (int)std::max<long>(1, (d.nblk + c->persist_slots[k] - 1) / c->persist_slots[k]);
    } else {
        d.ty = ts.ty ? ts.ty : march_rows(c, d.ntx);
        const int nty = (c->Ny + d.ty - 1) / d.ty;
        d.nblk = d.ntx * (ts.march ? (nty + 3) / 4 : nty);
    }
    d.cur = c->cur;
    d.ring = c->ring;
    d.tstep = 1;
    d.cbuf = 0;
    for (size_t n = 0; n < c->trc.size(); ++n)
        if (c->trc[n].cbuf_valid) d.cbuf |= 1 << n;
    return d;
}

template <typename T, int TX, int TY, bool DIAG, bool LATE = false>
void launch_tile(swm_ctx* c, const Dyn<T>& d) {
    const int var = c->d.variant;
    const int nt = (int)c->trc.size();
    dim3 grid(d.nblk), block(256);
#define SWM_LAUNCH(V, N) \
    hipLaunchKernelGGL((swm::step_kernel<T, V, N, TX, TY, DIAG, LATE>), grid, block, 0, c->stream, d)
    if (var == SWM_LINEARISED_MEAN_STATE) {   // no tracers on the linear variants
        SWM_LAUNCH(swm::VAR_MS, 0);
    } else if (var == SWM_FULLY_NONLINEAR) {
        if (nt == 0) SWM_LAUNCH(swm::VAR_FN, 0);
        else if (nt == 1) SWM_LAUNCH(swm::VAR_FN, 1);
        else SWM_LAUNCH(swm::VAR_FN, 2);
    } else {
        if (nt == 0) SWM_LAUNCH(swm::VAR_SOR, 0);
        else if (nt == 1) SWM_LAUNCH(swm::VAR_SOR, 1);
        else SWM_LAUNCH(swm::VAR_SOR, 2);
    }
#undef SWM_LAUNCH
}

template <typename T>
void launch_march(swm_ctx* c, const Dyn<T>& d) {
    const int var = c->d.variant;
    const int nt = (int)c->trc.size();
    dim3 grid(d.nblk), block(256);
#define SWM_LAUNCH(V, N) hipLaunchKernelGGL((swm::march_kernel<T, V, N>), grid, block, 0, c->stream, d)
    if (var == SWM_FULLY_NONLINEAR) {
        if (nt == 0) SWM_LAUNCH(swm::VAR_FN, 0);
        else if (nt == 1) SWM_LAUNCH(swm::VAR_FN, 1);
        else SWM_LAUNCH(swm::VAR_FN, 2);
    } else {
        if (nt == 0) SWM_LAUNCH(swm::VAR_SOR, 0);
        else if (nt == 1) SWM_LAUNCH(swm::VAR_SOR, 1);
        else SWM_LAUNCH(swm::VAR_SOR, 2);
    }
#undef SWM_LAUNCH
}

template <typename T, int TX, int TY>
void launch_persist(swm_ctx* c, const Dyn<T>& d) {
    const int var = c->d.variant;
    const int nt = (int)c->trc.size();
    dim3 grid((d.nblk + d.tpb - 1) / d.tpb), block(256);
#define SWM_LAUNCH(V, N) hipLaunchKernelGGL((swm::step_kernel_p<T, V, N, TX, TY>), grid, block, 0, c->stream, d)
    if (var == SWM_FULLY_NONLINEAR) {
        if (nt == 0) SWM_LAUNCH(swm::VAR_FN, 0);
        else if (nt == 1) SWM_LAUNCH(swm::VAR_FN, 1);
        else SWM_LAUNCH(swm::VAR_FN, 2);
    } else {
        if (nt == 0) SWM_LAUNCH(swm::VAR_SOR, 0);
        else if (nt == 1) SWM_LAUNCH(swm::VAR_SOR, 1);
        else SWM_LAUNCH(swm::VAR_SOR, 2);
    }
#undef SWM_LAUNCH
}

template <typename T, bool DIAG>
int launch_step_t(swm_ctx* c, const Dyn<T>& d, int tile) {
    if (DIAG) {
        launch_tile<T, 64, 8, true>(c, d);
    } else {
        switch (tile) {
        case 1: launch_tile<T, 32, 8, false>(c, d); break;
        case 2: launch_tile<T, 64, 8, false>(c, d); break;
        case 3: launch_tile<T, 128, 4, false>(c, d); break;
        case 4: launch_tile<T, 32, 16, false>(c, d); break;
        case 5: case 6: case 7: case 8: launch_march<T>(c, d); break;
        case 9: launch_persist<T, 64, 4>(c, d); break;
        case 10: launch_persist<T, 32, 8>(c, d); break;
        case 11: launch_tile<T, 32, 8, false, true>(c, d); break;
        case 12: launch_tile<T, 64, 4, false, true>(c, d); break;
        default: launch_tile<T, 64, 4, false>(c, d); break;
        }
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}

template <typename T>
void set_osc(swm_ctx* c, Dyn<T>& d, double tf) {
    for (int comp = 0; comp < 3; ++comp) {
        if (!c->d.osc_re[comp]) continue;
        const double arg = c->d.osc_omega[comp] * tf * c->d.dt;
        d.oc[comp] = (T)std::cos(arg);
        d.os[comp] = (T)std::sin(arg);
    }
}

// Place of every time-interpolated stream at model time tf*dt (tf = itt of
// the SWM_forcing_update that set the forcing; the reference passes itt*dt).
template <typename T>
void set_streams(swm_ctx* c, Dyn<T>& d, double tf) {
    for (size_t n = 0; n < c->streams.size(); ++n) {
        int64_t r0 = 0, r1 = 0;
        bool exact = true;
        double a = 0.0, b = 1.0;
        c->streams[n].idx.query(tf * c->d.dt, r0, r1, exact, a, b);
        d.ss[n].o0 = r0 * c->plane;
        d.ss[n].o1 = r1 * c->plane;
        d.ss[n].exact = exact ? 1 : 0;
        d.ss[n].a = (T)a;
        d.ss[n].b = (T)b;
    }
}

// Place of the DYNFROMFILE datasets at model time tf*dt (DFF_advance is
// called with itt*dt at every model_advance; the step that starts from itt
// reads the master fields of that advance).
template <typename T>
void set_dyn(swm_ctx* c, SStepArr<T>& ds, double tf) {
    for (int f = 0; f < 2; ++f) {
        if (!c->dyn[f].rec) continue;
        int64_t r0 = 0, r1 = 0;
        bool exact = true;
        double a = 0.0, b = 1.0;
        c->dyn[f].idx.query(tf * c->d.dt, r0, r1, exact, a, b);
        ds[f].o0 = r0 * c->plane;
        ds[f].o1 = r1 * c->plane;
        ds[f].exact = exact ? 1 : 0;
        ds[f].a = (T)a;
        ds[f].b = (T)b;
    }
}

template <typename T>
int step_one(swm_ctx* c) {
    c->itt += 1;
    Dyn<T> d = make_dyn<T>(c, c->tile);
    const int tstep = (int)std::min<int64_t>(3, c->itt);
    d.tstep = tstep;
    d.dtc1 = (T)(c->d.dt * c->ab[tstep - 1][0]);
    d.dtc2 = (T)(c->d.dt * c->ab[tstep - 1][1]);
    d.dtc3 = (T)(c->d.dt * c->ab[tstep - 1][2]);
    // the forcing used by step itt was set by SWM_forcing_update at itt-1
    set_osc(c, d, (double)(c->itt - 1));
    set_streams(c, d, (double)(c->itt - 1));
    set_dyn<T>(c, d.ds, (double)(c->itt - 1));
    int rc = launch_step_t<T, false>(c, d, c->tile);
    if (rc) return rc;
    // model_advance: N0 <- N0p1, G(1) <- G(2), G(2) <- G(3)
    c->cur = 1 - c->cur;
    c->ring = (c->ring + 1) % 3;
    for (auto& tr : c->trc) tr.cbuf_valid = false;
    // averaging diag tasks: add dt * field**p of the new N0 (Diag after model_advance)
    for (auto& av : c->avgs) {
        const int pl = av.field == 0 ? swm::PL_U + c->cur : av.field == 1 ? swm::PL_V + c->cur
                     : av.field == 2 ? swm::PL_E + c->cur : swm::PL_TRC + swm::PL_PER_TRC * av.tracer + c->cur;
        const long n = (long)c->Nx * c->Ny;
        hipLaunchKernelGGL(swm::avg_accumulate<T>, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->Nx, c->Ny, c->P,
                           (const T*)plane_ptr(c, pl), av.buf, c->d.dt, av.square);
        av.weight += c->d.dt;
    }
    return 0;
}

void drop_graphs(swm_ctx* c) {
    for (auto& g : c->graphs)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

// Capture the six steady-state steps that start from ping-pong level `cur`
// and ring slot `ring` into c->graphs[cur * 3 + ring].  Capturing records the
// launches without running them, so the host-side step state is put back.
template <typename T>
int capture_graph(swm_ctx* c, int cur, int ring) {
    const int key = cur * 3 + ring;
    if (c->graphs[key]) return 0;
    HipLock lock(g_hip_mu);
    const int cur0 = c->cur, ring0 = c->ring;
    const int64_t itt0 = c->itt;
    std::vector<double> w0;
    for (auto& av : c->avgs) w0.push_back(av.weight);
    std::vector<bool> cb0;
    for (auto& tr : c->trc) {
        cb0.push_back(tr.cbuf_valid);
        tr.cbuf_valid = false;    // steady steps read no host C buffer
    }
    c->cur = cur;
    c->ring = ring;
    if (c->itt < 2) c->itt = 2;   // steady steps use the full AB3 coefficients
    hipGraph_t g = nullptr;
    int rc = 0;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int k = 0; k < 6 && !rc; ++k) rc = step_one<T>(c);
        e = hipStreamEndCapture(c->stream, &g);
    }
    c->cur = cur0;
    c->ring = ring0;
    c->itt = itt0;
    for (size_t q = 0; q < w0.size(); ++q) c->avgs[q].weight = w0[q];
    for (size_t q = 0; q < cb0.size(); ++q) c->trc[q].cbuf_valid = cb0[q];
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(c, std::string("graph capture: ") + hipGetErrorString(e), 3);
    e = hipGraphInstantiate(&c->graphs[key], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->graphs[key] = nullptr;
        return fail(c, std::string("graph instantiate: ") + hipGetErrorString(e), 3);
    }
    // move the graph's one-time device set-up out of the first replay
    HIPCHK(c, hipGraphUpload(c->graphs[key], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Steady state (itt >= 3, no first-step tracer buffer, no oscillating or
// time-interpolated stream, whose per-step arguments change every step):
// the launch arguments repeat with the period 6 of the ping-pong x ring roles,
// so six steps are captured once per phase and replayed as one graph.
bool steady_ok(const swm_ctx* c) {
    return c->use_graph && !c->has_osc && c->streams.empty() && !c->dyn[0].rec && !c->dyn[1].rec && !c->dyn[2].rec;
}

// Build, before any step is enqueued, every graph that step_many(nsteps) will
// replay, so that the host-side capture and instantiation stay outside the
// device-time window of swm_gpu_step.
template <typename T>
int prepare_graphs(swm_ctx* c, int64_t nsteps) {
    if (!steady_ok(c)) return 0;
    int cur = c->cur, ring = c->ring;
    int64_t itt = c->itt, done = 0;
    bool cbuf = false;
    for (auto& tr : c->trc) cbuf = cbuf || tr.cbuf_valid;
    while (done < nsteps) {
        if (!cbuf && itt >= 2 && nsteps - done >= 6) {
            if (int rc = capture_graph<T>(c, cur, ring)) return rc;
            return 0;   // a six-step cycle leaves (cur, ring) unchanged: one graph serves the rest
        }
        itt += 1;
        done += 1;
        cur = 1 - cur;
        ring = (ring + 1) % 3;
        cbuf = false;
    }
    return 0;
}

template <typename T>
int step_many(swm_ctx* c, int64_t nsteps) {
    int64_t done = 0;
    while (done < nsteps) {
        bool cbuf = false;
        for (auto& tr : c->trc) cbuf = cbuf || tr.cbuf_valid;
        const bool steady = steady_ok(c) && !cbuf && c->itt >= 2;
        if (steady && nsteps - done >= 6) {
            const int key = c->cur * 3 + c->ring;
            if (int rc = capture_graph<T>(c, c->cur, c->ring)) return rc;
            HIPCHK(c, hipGraphLaunch(c->graphs[key], c->stream));
            c->itt += 6;
            done += 6;
            for (auto& av : c->avgs)   // the graph holds their accumulations; bufferCount adds dt per step
                for (int q = 0; q < 6; ++q) av.weight += c->d.dt;
        } else {
            int rc = step_one<T>(c);
            if (rc) return rc;
            done += 1;
        }
    }
    return 0;
}

// prev: evaluate on the state before the last step (the other ping-pong level),
// which is what the lateral-mixing register holds after a step.
template <typename T>
int materialise(swm_ctx* c, bool prev = false) {
    bool fresh = false;
    for (int k = 0; k < 14; ++k)
        if (!c->diag[k]) {
            if (!(c->diag[k] = alloc_grid(c))) return fail(c, "out of device memory", 3);
            fresh = true;
        }
    for (size_t n = 0; n < c->trc.size(); ++n)
        for (int k = 0; k < 3; ++k)
            if (!c->tdiag[n][k]) {
                if (!(c->tdiag[n][k] = alloc_grid(c))) return fail(c, "out of device memory", 3);
                fresh = true;
            }
    if (fresh) {
        int rc = write_statics(c);
        if (rc) return rc;
    }
    Dyn<T> d = make_dyn<T>(c, 2);   // the materialisation kernel uses the 64 x 8 tile (kTiles[2])
    if (prev) d.cur = 1 - c->cur;
    // forcing as left by the last SWM_forcing_update
    set_osc(c, d, (double)c->itt);
    set_streams(c, d, (double)c->itt);
    return launch_step_t<T, true>(c, d, 0);
}

int alloc_state(swm_ctx* c, int ntrc) {
    const int planes = swm::PL_TRC + swm::PL_PER_TRC * ntrc;
    void* p = dalloc(c, (size_t)planes * c->plane * c->esz);
    if (!p) return fail(c, "out of device memory", 3);
    HIPCHK(c, zero_sync(c, p, (size_t)planes * c->plane * c->esz));
    if (c->state) {
        const int old = swm::PL_TRC + swm::PL_PER_TRC * (ntrc - 1);
        HIPCHK(c, copy_sync(c, p, c->state, (size_t)old * c->plane * c->esz, hipMemcpyDeviceToDevice));
        dfree(c, c->state);
    }
    c->state = p;
    return 0;
}

int member_refresh(swm_ctx* c);

}  // namespace

extern "C" {

static thread_local std::string g_create_err;

static int create_impl(swm_ctx* c, const swm_desc* d);

int swm_gpu_create(const swm_desc* d, swm_ctx** out) {
    HipLock lock(g_hip_mu);
    if (!out) return 1;
    *out = nullptr;
    g_create_err.clear();
    if (!d) { g_create_err = "null descriptor"; return 1; }
    if (d->abi_version != SWM_GPU_ABI_VERSION) { g_create_err = "ABI version mismatch"; return 1; }
    if (d->Nx < 3 || d->Ny < 3) { g_create_err = "Nx and Ny must be >= 3"; return 1; }
    if (d->variant != SWM_FULLY_NONLINEAR && d->variant != SWM_LINEARISED_STATE_OF_REST &&
        d->variant != SWM_LINEARISED_MEAN_STATE) {
        g_create_err = "unknown variant";
        return 4;
    }
    if (d->variant == SWM_LINEARISED_MEAN_STATE) {
        if (!d->u_bs || !d->v_bs || !d->zeta_bs) {
            g_create_err = "LINEARISED_MEAN_STATE needs the basic state u_bs, v_bs and zeta_bs";
            return 1;
        }
        // the G_eta basic-state term reads column ip1(j) of row j
        // (src/swm_timestep_module.f90:360); ip1 has Nx entries
        if (d->Ny > d->Nx) {
            g_create_err = "LINEARISED_MEAN_STATE needs Ny <= Nx (the reference indexes ip1(j))";
            return 4;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { g_create_err = "no HIP device"; return 5; }
    swm_ctx* c = new swm_ctx();
    const int rc = create_impl(c, d);
    if (rc) {
        g_create_err = c->err.empty() ? "create failed" : c->err;
        swm_gpu_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

static int create_impl(swm_ctx* c, const swm_desc* d) {
    c->d = *d;
    c->Nx = d->Nx;
    c->Ny = d->Ny;
    c->fp32 = d->precision == SWM_FP32;
    c->esz = c->fp32 ? 4 : 8;
    // Default shape from measured step times on MI355X (1024^2 .. 4096^2):
    // 32x8 tiles carry less halo per output cell than 64x4 and win at every
    // size; loading the AB history after the halo stage wins in FP64 at every
    // size and in FP32 up to 2^22 cells.
    const long long cells = (long long)d->Nx * d->Ny;
    c->tile = (c->fp32 && cells > (1LL << 22)) ? 1 : 11;
    if (const char* e = getenv("SWM_GPU_TILE")) c->tile = std::max(0, std::min(kNumTiles - 1, atoi(e)));
    if (d->variant == SWM_LINEARISED_MEAN_STATE && kTiles[c->tile].march) c->tile = 1;   // LDS tile kernels only
    if (const char* e = getenv("SWM_GPU_GRAPH")) c->use_graph = atoi(e) != 0;
    if (const char* e = getenv("SWM_GPU_SOR_BLOCK")) c->sor_blocked = atoi(e) != 0;
    const int align = c->fp32 ? 32 : 16;   // 128-byte aligned rows
    c->P = (c->Nx + align - 1) / align * align;
    c->plane = (long long)c->P * c->Ny;
    c->has_osc = d->osc_re[0] || d->osc_re[1] || d->osc_re[2];
    c->has_taux = d->tau_x != nullptr;
    c->has_tauy = d->tau_y != nullptr;
    for (int g = 0; g < 4; ++g) {
        if (!d->grid[g].ocean || !d->grid[g].cos_lat || !d->grid[g].tan_lat || !d->grid[g].f)
            return fail(c, "every grid needs ocean, cos_lat, tan_lat and f", 1);
        c->hmask[g].assign(d->grid[g].ocean, d->grid[g].ocean + (size_t)c->Nx * c->Ny);
    }
    if (!d->grid[SWM_GRID_ETA].H) return fail(c, "grid[ETA].H is required", 1);
    if ((d->osc_re[0] && !d->osc_im[0]) || (d->osc_re[1] && !d->osc_im[1]) || (d->osc_re[2] && !d->osc_im[2]))
        return fail(c, "oscillating forcing needs both real and imaginary parts", 1);
    if ((d->tau_x || d->tau_y) && d->variant != SWM_FULLY_NONLINEAR)
        return fail(c, "per-step windstress streams are a FULLY_NONLINEAR feature; fold constant "
                       "windstress into F_*_const for linear variants", 1);
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        int dev = 0;
        HIPCHK(c, hipGetDevice(&dev));
        HIPCHK(c, hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    // time_integration_init (src/time_integration_module.f90:44-51)
    const double chi = d->ab_chi;
    c->ab[0][0] = 1.0;
    c->ab[1][0] = -0.5 - chi; c->ab[1][1] = 1.5 + chi;
    c->ab[2][0] = 5.0 / 12.0; c->ab[2][1] = -16.0 / 12.0; c->ab[2][2] = 23.0 / 12.0;
    // packed masks: bit0 H, bit1 u, bit2 v, bit3 eta
    {
        std::vector<uint8_t> m((size_t)c->Ny * c->P, 0);
        for (int j = 0; j < c->Ny; ++j)
            for (int i = 0; i < c->Nx; ++i) {
                const size_t k = (size_t)j * c->Nx + i;
                m[(size_t)j * c->P + i] = (uint8_t)((c->hmask[0][k] ? 1 : 0) | (c->hmask[1][k] ? 2 : 0) |
                                                    (c->hmask[2][k] ? 4 : 0) | (c->hmask[3][k] ? 8 : 0));
            }
        c->mask = (uint8_t*)dalloc(c, m.size());
        if (!c->mask) return fail(c, "out of device memory", 3);
        HIPCHK(c, copy_sync(c, c->mask, m.data(), m.size(), hipMemcpyHostToDevice));
    }
    const int8_t *mu = c->hmask[1].data(), *mv = c->hmask[2].data(), *me = c->hmask[3].data();
    int rc = 0;
    std::vector<double> kap[3];
    const double* imp[3] = {damp_static(c, d->impl_u, kap[0]), damp_static(c, d->impl_v, kap[1]),
                            damp_static(c, d->impl_eta, kap[2])};
    const double idflt = c->fp32 ? 0.0 : 1.0;
    struct { int id; const double* h; double dflt; const int8_t* m; } ups[] = {
        {F_HE, d->grid[SWM_GRID_ETA].H, 0.0, me}, {F_HU, d->grid[SWM_GRID_U].H, 0.0, mu},
        {F_HV, d->grid[SWM_GRID_V].H, 0.0, mv},   {F_IMU, imp[0], idflt, mu},
        {F_IMV, imp[1], idflt, mv},               {F_IME, imp[2], idflt, me},
        {F_GSU, d->gamma_sq_u, 0.0, mu},          {F_GSV, d->gamma_sq_v, 0.0, mv},
        {F_FXC, d->F_x_const, 0.0, mu},           {F_FYC, d->F_y_const, 0.0, mv},
        {F_FEC, d->F_eta_const, 0.0, me},         {F_TAUX, d->tau_x, 0.0, mu},
        {F_TAUY, d->tau_y, 0.0, mv},              {F_ORX, d->osc_re[0], 0.0, mu},
        {F_OIX, d->osc_im[0], 0.0, mu},           {F_ORY, d->osc_re[1], 0.0, mv},
        {F_OIY, d->osc_im[1], 0.0, mv},           {F_ORE, d->osc_re[2], 0.0, me},
        {F_OIE, d->osc_im[2], 0.0, me},           {F_UBS, d->u_bs, 0.0, mu},
        {F_VBS, d->v_bs, 0.0, mv},                {F_ZBS, d->zeta_bs, 0.0, c->hmask[0].data()}};
    for (auto& u : ups)
        if ((rc = upload_field(c, &c->fields[u.id], u.h, u.dflt, u.m))) return rc;
    if ((rc = build_tables(c))) return rc;
    if ((rc = alloc_state(c, 0))) return rc;
    if (!(c->scratch = alloc_grid(c))) return fail(c, "out of device memory", 3);
    c->ce_host.assign(d->grid[SWM_GRID_ETA].cos_lat, d->grid[SWM_GRID_ETA].cos_lat + c->Ny);
    c->cu_host.assign(d->grid[SWM_GRID_U].cos_lat, d->grid[SWM_GRID_U].cos_lat + c->Ny);
    // host pointers in the descriptor are only valid during this call
    for (int g = 0; g < 4; ++g) c->d.grid[g] = swm_grid_desc{};
    c->d.impl_u = c->d.impl_v = c->d.impl_eta = c->d.gamma_sq_u = c->d.gamma_sq_v = nullptr;
    c->d.F_x_const = c->d.F_y_const = c->d.F_eta_const = c->d.tau_x = c->d.tau_y = nullptr;
    c->d.u_bs = c->d.v_bs = c->d.zeta_bs = nullptr;
    for (int k = 0; k < 3; ++k) {
        c->d.osc_re[k] = d->osc_re[k] ? (const double*)1 : nullptr;   // presence flag only
        c->d.osc_im[k] = nullptr;
    }
    return write_statics(c);
}

int swm_gpu_upload_state(swm_ctx* c, const double* u, const double* v, const double* eta, const double* G_u,
                         const double* G_v, const double* G_eta, int64_t itt) {
    HipLock lock(g_hip_mu);
    if (!c) return 1;
    if (c->batch) return fail(c, "context is attached to a batch; detach it first", 1);
    if (!u || !v || !eta) return fail(c, "u, v and eta are required", 1);
    if (itt < 0) return fail(c, "itt must be >= 0", 1);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    const size_t n = (size_t)c->Nx * c->Ny;
    c->cur = 0;
    c->ring = 0;
    int rc;
    if ((rc = upload_state_grid(c, plane_ptr(c, swm::PL_U), plane_ptr(c, swm::PL_LO), u)) ||
        (rc = upload_state_grid(c, plane_ptr(c, swm::PL_V), plane_ptr(c, swm::PL_LO + 1), v)) ||
        (rc = upload_state_grid(c, plane_ptr(c, swm::PL_E), plane_ptr(c, swm::PL_LO + 2), eta)))
        return rc;
    const size_t pbytes = (size_t)c->plane * c->esz;
    for (int l = 0; l < 3; ++l) {
        const double* src[3] = {G_u ? G_u + l * n : nullptr, G_v ? G_v + l * n : nullptr,
                                G_eta ? G_eta + l * n : nullptr};
        const int pl[3] = {swm::PL_GU + l, swm::PL_GV + l, swm::PL_GE + l};
        for (int f = 0; f < 3; ++f) {
            if (src[f]) {
                if ((rc = upload_grid(c, plane_ptr(c, pl[f]), src[f]))) return rc;
            } else {   // no history given: zero tendencies (0.0 is all-zero bits)
                HIPCHK(c, hipMemsetAsync(plane_ptr(c, pl[f]), 0, pbytes, c->stream));
            }
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->itt = itt;
    c->itt_upload = itt;
    return 0;
}

int swm_gpu_add_tracer(swm_ctx* c, const swm_tracer_desc* t, int32_t* tracer_id) {
    HipLock lock(g_hip_mu);
    if (!c || !t) return 1;
    if (c->batch) return fail(c, "context is attached to a batch; detach it first", 1);
    if ((int)c->trc.size() >= kMaxTracers) return fail(c, "at most 2 tracers are fused into the step kernel", 4);
    if (!t->CH || !t->C) return fail(c, "tracer CH and C are required", 1);
    if (c->d.variant != SWM_FULLY_NONLINEAR)
        return fail(c, "tracers need FULLY_NONLINEAR (h = D); the linear D alias makes C = CH/0 on land", 4);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    const size_t n = (size_t)c->Nx * c->Ny;
    const int nt = (int)c->trc.size();
    int rc;
    if ((rc = alloc_state(c, nt + 1))) return rc;
    const int base = swm::PL_TRC + swm::PL_PER_TRC * nt;
    // the tracer's planes follow the context's current ping-pong / ring roles
    if ((rc = upload_state_grid(c, plane_ptr(c, base + c->cur), plane_ptr(c, base + swm::PL_TLO), t->CH))) return rc;
    if ((rc = upload_grid(c, plane_ptr(c, base + 5), t->C))) return rc;
    for (int l = 0; l < 3; ++l) {
        void* dst = plane_ptr(c, base + 2 + (c->ring + l) % 3);
        if (t->G_CH) {
            if ((rc = upload_grid(c, dst, t->G_CH + l * n))) return rc;
        } else {
            HIPCHK(c, hipMemsetAsync(dst, 0, (size_t)c->plane * c->esz, c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Tracer tr;
    tr.kappa = t->kappa_h;
    tr.cbuf_valid = true;
    const int8_t* me = c->hmask[3].data();
    std::vector<double> kap;
    if ((rc = upload_field(c, &tr.C0, t->C0, 0.0, me)) || (rc = upload_field(c, &tr.gC, t->gamma_C, 0.0, me)) ||
        (rc = upload_field(c, &tr.impl, damp_static(c, t->impl, kap), c->fp32 ? 0.0 : 1.0, me)))
        return rc;
    c->trc.push_back(tr);
    c->march_slots = 0;   // the kernel instance (and its occupancy) depends on the tracer count
    c->persist_slots[0] = c->persist_slots[1] = 0;
    if ((rc = write_statics(c))) return rc;
    if (tracer_id) *tracer_id = nt;
    return 0;
}

int swm_gpu_step(swm_ctx* c, int64_t nsteps) {
    if (!c) return 1;
    if (c->batch) return fail(c, "context is attached to a batch; it advances through swm_gpu_batch_step", 1);
    if (nsteps < 0) return fail(c, "nsteps must be >= 0", 1);
    if (int rc = c->fp32 ? prepare_graphs<float>(c, nsteps) : prepare_graphs<double>(c, nsteps)) return rc;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int rc = c->fp32 ? step_many<float>(c, nsteps) : step_many<double>(c, nsteps);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_total_ms = ms;
    c->last_kernel_ms = nsteps > 0 ? ms / (double)nsteps : 0.0;
    return 0;
}

int64_t swm_gpu_itt(const swm_ctx* c) {
    if (!c) return -1;
    HipLock lock(g_hip_mu);
    if (c->batch && member_refresh(const_cast<swm_ctx*>(c))) return -1;
    return c->itt;
}

int swm_gpu_last_timing(const swm_ctx* c, double* total_ms, double* step_kernel_ms) {
    if (!c) return 1;
    if (total_ms) *total_ms = c->last_total_ms;
    if (step_kernel_ms) *step_kernel_ms = c->last_kernel_ms;
    return 0;
}

}  // extern "C"

namespace {

int elliptic_chi(swm_ctx* c);
int elliptic_nd(swm_ctx* c);

// computeStreamfunction of the master u, v at N0 into psi_buf[3]; und / vnd
// (FP64 planes, optional) replace u, v by the non-divergent flow.
template <typename T>
int psi_t(swm_ctx* c, const double* und = nullptr, const double* vnd = nullptr) {
    const size_t bytes = (size_t)c->plane * sizeof(double);
    for (auto& b : c->psi_buf)
        if (!b && !(b = (double*)dalloc(c, bytes))) return fail(c, "out of device memory", 3);
    const Statics<T>* sp = (const Statics<T>*)c->statics;
    const T* u = (const T*)plane_ptr(c, swm::PL_U + c->cur);
    const T* v = (const T*)plane_ptr(c, swm::PL_V + c->cur);
    const long n = (long)c->Nx * c->Ny;
    hipLaunchKernelGGL(swm::psi_prep<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, sp, u, v, und,
                       vnd, c->psi_buf[0], c->psi_buf[1]);
    hipLaunchKernelGGL(swm::psi_col<T>, dim3((c->Nx + 255) / 256), dim3(256), 0, c->stream, sp,
                       (const double*)c->psi_buf[0], c->psi_buf[2], c->psi_buf[3]);
    const unsigned segs = (unsigned)((c->Nx - 1 + 255) / 256);
    hipLaunchKernelGGL(swm::psi_row<T>, dim3(segs, 1), dim3(256), 0, c->stream, sp, 0,
                       (const double*)c->psi_buf[1], (const double*)c->psi_buf[2], c->psi_buf[3]);
    if (c->Ny > 1)
        hipLaunchKernelGGL(swm::psi_row<T>, dim3(segs, (unsigned)(c->Ny - 1)), dim3(256), 0, c->stream, sp, 1,
                           (const double*)c->psi_buf[1], (const double*)c->psi_buf[2], c->psi_buf[3]);
    HIPCHK(c, hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int swm_gpu_sync(swm_ctx* c, const char* name, int32_t tracer_id, double* host_out) {
    HipLock lock(g_hip_mu);
    if (!c || !name || !host_out) return 1;
    if (int rc = member_refresh(c)) return rc;
    std::string s(name);
    for (auto& ch : s) ch = (char)toupper((unsigned char)ch);
    const bool ell = (c->d.flags & SWM_ELLIPTIC_SOLVER) != 0;
    if (s == "PSI") {   // computeStreamfunction of u, v at N0 (diagVar DVARNAME_PSI)
        if (tracer_id >= 0) return fail(c, "PSI is not a tracer field", 1);
        int rc = 0;
        if (ell) {   // of the non-divergent flow u - grad chi
            if ((rc = elliptic_nd(c))) return rc;
            rc = c->fp32 ? psi_t<float>(c, c->nd_buf[0], c->nd_buf[1]) : psi_t<double>(c, c->nd_buf[0], c->nd_buf[1]);
        } else {
            rc = c->fp32 ? psi_t<float>(c) : psi_t<double>(c);
        }
        if (rc) return rc;
        HIPCHK(c, copy2d_sync(c, host_out, c->Nx * 8, c->psi_buf[3], c->P * 8, c->Nx * 8, c->Ny,
                              hipMemcpyDeviceToHost));
        return 0;
    }
    if ((s == "U_ND" || s == "V_ND") && ell) {   // computeNonDivergentFlowField (diagVar U_ND, V_ND)
        if (tracer_id >= 0) return fail(c, s + " is not a tracer field", 1);
        if (int rc = elliptic_nd(c)) return rc;
        HIPCHK(c, copy2d_sync(c, host_out, c->Nx * 8, c->nd_buf[s == "U_ND" ? 0 : 1], c->P * 8, c->Nx * 8, c->Ny,
                              hipMemcpyDeviceToHost));
        return 0;
    }
    if (s == "CHI" && ell) {   // computeVelocityPotential (diagVar CHI): the reference-mode solve
        if (int rc = elliptic_chi(c)) return rc;
        const size_t n = (size_t)c->Nx * c->Ny;
        if (c->fp32) {
            std::vector<float> f(n);
            HIPCHK(c, copy2d_sync(c, f.data(), c->Nx * 4, c->chi_ref, c->P * 4, c->Nx * 4, c->Ny, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) host_out[k] = f[k];
        } else {
            HIPCHK(c, copy2d_sync(c, host_out, c->Nx * 8, c->chi_ref, c->P * 8, c->Nx * 8, c->Ny, hipMemcpyDeviceToHost));
        }
        return 0;
    }
    if (s == "CHI") {   // the last device-resident SOR solution
        if (c->chi_last < 0) return fail(c, "no velocity potential has been computed", 1);
        const int w = c->chi_last;
        const size_t n = (size_t)c->Nx * c->Ny;
        if (w == 1) {
            std::vector<float> f(n);
            HIPCHK(c, copy2d_sync(c, f.data(), c->Nx * 4, c->sor_buf[1][0], c->P * 4, c->Nx * 4, c->Ny, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k) host_out[k] = f[k];
        } else {
            HIPCHK(c, copy2d_sync(c, host_out, c->Nx * 8, c->sor_buf[0][0], c->P * 8, c->Nx * 8, c->Ny,
                                   hipMemcpyDeviceToHost));
        }
        return 0;
    }
    const int n0 = c->cur;
    const int r1 = (c->ring + 1) % 3, r0 = c->ring;
    const void* src = nullptr;
    const void* lo_src = nullptr;   // low part of an FP32 state field
    if (tracer_id >= 0) {
        if (tracer_id >= (int)c->trc.size()) return fail(c, "no such tracer", 1);
        const int base = swm::PL_TRC + swm::PL_PER_TRC * tracer_id;
        if (s == "CH") src = plane_ptr(c, base + n0), lo_src = plane_ptr(c, base + swm::PL_TLO);
        else if (s == "G_CH") src = plane_ptr(c, base + 2 + r1);   // newest tendency (G(3) == G(2) after the shift)
        else if (s == "G_CH_M1") src = plane_ptr(c, base + 2 + r0);
        else if (s == "C" || s == "UHC" || s == "VHC") {
            int rc = c->fp32 ? materialise<float>(c) : materialise<double>(c);
            if (rc) return rc;
            src = c->tdiag[tracer_id][s == "C" ? 0 : (s == "UHC" ? 1 : 2)];
        } else {
            return fail(c, "unknown tracer field " + s, 1);
        }
    } else if ((s == "U" || s == "V" || s == "ETA") && c->dyn[s == "U" ? 0 : (s == "V" ? 1 : 2)].rec) {
        // master field = ocean * DYNFROMFILE data at itt*dt + SWM field (model_advance)
        const int f = s == "U" ? 0 : (s == "V" ? 1 : 2);
        const int pl = f == 0 ? swm::PL_U : (f == 1 ? swm::PL_V : swm::PL_E);
        const uint8_t bit = f == 0 ? swm::OC_U : (f == 1 ? swm::OC_V : swm::OC_E);
        int64_t r0 = 0, r1 = 0;
        bool exact = true;
        double qa = 0.0, qb = 1.0;
        c->dyn[f].idx.query((double)c->itt * c->d.dt, r0, r1, exact, qa, qb);
        const long n = (long)c->Nx * c->Ny;
        if (c->fp32) {
            swm::SStep<float> st{r0 * c->plane, r1 * c->plane, exact ? 1 : 0, (float)qa, (float)qb};
            hipLaunchKernelGGL(swm::master_kernel<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                               c->Nx, c->Ny, c->P, c->mask, bit, (const float*)c->dyn[f].rec, st,
                               (const float*)plane_ptr(c, pl + n0), (float*)c->scratch);
        } else {
            swm::SStep<double> st{r0 * c->plane, r1 * c->plane, exact ? 1 : 0, qa, qb};
            hipLaunchKernelGGL(swm::master_kernel<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                               c->Nx, c->Ny, c->P, c->mask, bit, (const double*)c->dyn[f].rec, st,
                               (const double*)plane_ptr(c, pl + n0), (double*)c->scratch);
        }
        HIPCHK(c, hipGetLastError());
        src = c->scratch;
    } else if (s == "SWM_U" || s == "U" || s == "U_ND") {   // u_nd = u without a solver
        src = plane_ptr(c, swm::PL_U + n0), lo_src = plane_ptr(c, swm::PL_LO);
    } else if (s == "SWM_V" || s == "V" || s == "V_ND") {
        src = plane_ptr(c, swm::PL_V + n0), lo_src = plane_ptr(c, swm::PL_LO + 1);
    } else if (s == "SWM_ETA" || s == "ETA") {
        src = plane_ptr(c, swm::PL_E + n0), lo_src = plane_ptr(c, swm::PL_LO + 2);
    }
    else if (s == "G_U") src = plane_ptr(c, swm::PL_GU + r1);
    else if (s == "G_V") src = plane_ptr(c, swm::PL_GV + r1);
    else if (s == "G_ETA") src = plane_ptr(c, swm::PL_GE + r1);
    else if (s == "G_U_M1") src = plane_ptr(c, swm::PL_GU + r0);
    else if (s == "G_V_M1") src = plane_ptr(c, swm::PL_GV + r0);
    else if (s == "G_ETA_M1") src = plane_ptr(c, swm::PL_GE + r0);
    else {
        static const char* names[14] = {"SWM_D", "SWM_DH", "SWM_DU", "SWM_DV", "SWM_RELVORT", "SWM_POT", "SWM_EDENS",
                                        "SWM_MU", "SWM_MV", "F_X", "F_Y", "F_ETA", "SWM_LATMIX_U", "SWM_LATMIX_V"};
        int k = -1;
        for (int q = 0; q < 14; ++q)
            if (s == names[q]) k = q;
        if (k < 0) return fail(c, "unknown field " + s, 1);
        if (k >= 12) {
            // swm_latmix_u/v are computed at the start of SWM_timestep_step from the
            // state before the step (src/swm_timestep_module.f90:171-173) and hold
            // initVar's 0 before the first step.
            if (c->itt == c->itt_upload && c->itt > 0)
                return fail(c, "the lateral-mixing trend of the last step is not on the device after a "
                               "mid-run hand-over; step once first", 1);
            if (c->itt == 0 || !(c->d.flags & SWM_LATERAL_MIXING)) {
                for (size_t q = 0; q < (size_t)c->Nx * c->Ny; ++q) host_out[q] = 0.0;
                return 0;
            }
        }
        int rc = c->fp32 ? materialise<float>(c, k >= 12) : materialise<double>(c, k >= 12);
        if (rc) return rc;
        src = c->diag[k];
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return download_state_grid(c, host_out, src, lo_src);
}

}  // extern "C"

namespace {

// Device buffers of the SOR driver, allocated on first use and kept: chi/B in
// the sweep precision, coefficient tables in that precision, the omega schedule.
int sor_prepare(swm_ctx* c, bool f32, int nblk) {
    if (!c->sor_partial) {
        // partial sums of the colour launches (2 x nblk) or of the blocked
        // sweeps (one per 64 x 16 tile, two sets by launch parity)
        const size_t nb = std::max<size_t>(2 * (size_t)nblk, 2 * (size_t)((c->Nx + 63) / 64) * ((c->Ny + 15) / 16));
        c->sor_partial = (double*)dalloc(c, sizeof(double) * nb);
        c->sor_done = (int*)dalloc(c, sizeof(int));
        c->sor_iters = (long long*)dalloc(c, sizeof(long long));
        c->sor_norm = (double*)dalloc(c, sizeof(double));
        if (!c->sor_partial || !c->sor_done || !c->sor_iters || !c->sor_norm)
            return fail(c, "out of device memory", 3);
    }
    const int w = f32 ? 1 : 0;
    const size_t bytes = (size_t)c->P * c->Ny * (f32 ? 4 : 8);
    if (!c->sor_buf[w][0]) {
        c->sor_buf[w][0] = dalloc(c, bytes);
        c->sor_buf[w][1] = dalloc(c, bytes);
        c->sor_alt[w] = dalloc(c, bytes);
        if (!c->sor_buf[w][0] || !c->sor_buf[w][1] || !c->sor_alt[w]) return fail(c, "out of device memory", 3);
    }
    if (!c->sor_ftab[w][0]) {
        const int Ny = c->Ny;
        const size_t sz[5] = {(size_t)Ny, (size_t)Ny, (size_t)Ny, 12 * (size_t)Ny, 12 * (size_t)Ny};
        for (int q = 0; q < 5; ++q) {
            if (f32 == c->fp32) { c->sor_ftab[w][q] = c->sor_tab[q]; continue; }
            // tables were uploaded in the context precision; convert on the host
            std::vector<double> h(sz[q]);
            std::vector<float> f(sz[q]);
            HIPCHK(c, copy_sync(c, h.data(), c->sor_tab[q], sz[q] * 8, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < sz[q]; ++k) f[k] = (float)h[k];
            if (!(c->sor_ftab[w][q] = dalloc(c, sz[q] * 4))) return fail(c, "out of device memory", 3);
            HIPCHK(c, copy_sync(c, c->sor_ftab[w][q], f.data(), sz[q] * 4, hipMemcpyHostToDevice));
        }
    }
    if (!c->sor_relax) {
        // reference omega schedule (src/ElSolv_SOR.f90:175-210), kept until stationary
        const double rJ = (double)9.9999178e-1f;
        std::vector<double> relax;
        double wv = 1.0;
        relax.push_back(wv);
        for (int h = 1; h < (1 << 20); ++h) {
            const double nw = (h == 1) ? 1.0 / (1.0 - .5 * (rJ * rJ)) : 1.0 / (1.0 - .25 * (rJ * rJ) * wv);
            relax.push_back(nw);
            if (h > 2 && nw == wv) break;
            wv = nw;
        }
        if (!(c->sor_relax = dalloc(c, relax.size() * 8))) return fail(c, "out of device memory", 3);
        HIPCHK(c, copy_sync(c, c->sor_relax, relax.data(), relax.size() * 8, hipMemcpyHostToDevice));
        c->sor_nrelax = (int)relax.size();
    }
    return 0;
}

// Fixed-omega sweeps l .. l+K-1 in one blocked launch (sor_block) from chi_in
// into chi_out, then the convergence test if sweep l+K-1 is a test sweep.
// With mid (reference mode) the launch also keeps chi after each of its first
// K - 1 sweeps in sor_mid[.][sweep][parity] and their partial sums in
// sor_partial_mid, and block 0 of the next launch tests the sweeps in order.
template <typename T, int K, int TY>
void sor_block_launch(swm_ctx* c, const void* chi_in, void* chi_out, void* dB, void* const* tabs, double omega,
                      double tol, double bnorm, int64_t l, bool check, bool ref_mode, int parity, int64_t prev_sweep,
                      bool mid = false) {
    constexpr int TX = 64;
    const int ntx = (c->Nx + TX - 1) / TX;
    const int nb = ntx * ((c->Ny + TY - 1) / TY);
    swm::SorArgs<T> a{};
    a.Nx = c->Nx; a.Ny = c->Ny; a.P = c->P; a.color = 0; a.sweep = (int)l;
    a.chi = nullptr; a.B = (const T*)dB; a.mask = c->mask;
    a.relax = ref_mode ? (const double*)c->sor_relax : nullptr;
    a.nrelax = c->sor_nrelax; a.omega = omega;
    a.c1t = (const T*)tabs[0]; a.c3t = (const T*)tabs[1]; a.c4t = (const T*)tabs[2]; a.c5t = (const T*)tabs[3]; a.c5r = (const T*)tabs[4];
    // partial sums alternate between two halves of the buffer by launch parity:
    // block 0 of the next launch reads this launch's while its blocks write theirs
    a.partial = c->sor_partial + (size_t)parity * nb;
    a.done = c->sor_done;
    a.residual_mode = ref_mode ? 0 : 1;
    a.accumulate = check ? 1 : 0;
    a.prev_partial = c->sor_partial + (size_t)(parity ^ 1) * nb;
    a.prev_check = prev_sweep > 0 ? 1 : 0;
    a.prev_sweep = (int)prev_sweep;
    a.eps = tol;
    a.bnorm = bnorm;
    a.iters = c->sor_iters;
    a.norm_out = c->sor_norm;
    if (mid) {
        const int w = std::is_same<T, float>::value ? 1 : 0;
        a.mid = 1;
        a.prev_mid = a.prev_check;
        for (int q = 0; q < K - 1; ++q) {
            a.chi_mid[q] = (T*)c->sor_mid[w][q][parity];
            a.partial_mid[q] = c->sor_partial_mid + (size_t)(2 * q + parity) * nb;
            a.prev_partial_mid[q] = c->sor_partial_mid + (size_t)(2 * q + (parity ^ 1)) * nb;
        }
    }
    hipLaunchKernelGGL((swm::sor_block<T, TX, TY, K>), dim3(nb), dim3(256), 0, c->stream, a, (const T*)chi_in,
                       (T*)chi_out, ntx);
}

template <typename T>
int sor_sweeps(swm_ctx* c, void* dchi, void* dB, void* const* tabs, bool ref_mode, double omega, double tol,
                      double bnorm, int64_t max_sweeps, int nblk, int* host_done) {
    const int every = ref_mode ? 1 : c->sor_every;   // the reference tests after every sweep
    // sweeps enqueued between reads of the convergence flag: 32 at first (a
    // warm-started solve stops early without queueing many idle launches),
    // doubling up to 256 for long solves (fewer host round trips)
    int64_t batch = std::max(32, 4 * every);
    const int64_t batch_max = std::max(256, 4 * every);
    *host_done = 0;
    int64_t l = 1;
    // Blocked sweeps, chi alternating between dchi and the second buffer: K = 4
    // or 2 sweeps per launch in fixed-omega mode when the test cadence is a
    // multiple of it, else (and in the reference mode, which tests every sweep
    // with its Chebyshev omega) one sweep per launch.
    // sweeps per launch K: a set halo (2K cells), else the largest of 4, 2, 1
    // that divides the test cadence; the reference mode tests every sweep (K <= 1)
    int K = c->sor_halo >= 0 ? c->sor_halo / 2 : (every % 4 == 0 ? 4 : (every % 2 == 0 ? 2 : 1));
    // the reference mode tests every sweep: a launch of K > 1 sweeps keeps the
    // chi and residual of its first K - 1 sweeps too (mid); K from the halo,
    // 4 by default
    if (ref_mode) K = c->sor_halo < 0 ? 4 : c->sor_halo / 2;
    if (!c->sor_blocked) K = 0;
    const bool mid = ref_mode && K > 1;
    if (!ref_mode && K > 1 && every % K) return fail(c, "the SOR halo's sweeps per launch must divide the test cadence", 1);
    const int rows = (K == 4) ? 16 : c->sor_rows;   // 64x32 tiles with a 4-sweep halo exceed 64 KiB of LDS
    if (mid) {
        const int w = std::is_same<T, float>::value ? 1 : 0;
        const size_t bytes = (size_t)c->P * c->Ny * sizeof(T);
        if (!c->sor_partial_mid) {
            const size_t nbm = 6 * (size_t)((c->Nx + 63) / 64) * ((c->Ny + 15) / 16);
            c->sor_partial_mid = (double*)dalloc(c, sizeof(double) * nbm);
            if (!c->sor_partial_mid) return fail(c, "out of device memory", 3);
        }
        for (int q = 0; q < K - 1; ++q)
            for (int par = 0; par < 2; ++par)
                if (!c->sor_mid[w][q][par] && !(c->sor_mid[w][q][par] = dalloc(c, bytes)))
                    return fail(c, "out of device memory", 3);
    }
    if (K) {
        // Launch n does sweeps n*K+1 .. (n+1)*K from one buffer into the other
        // (dchi -> alt for even n); its convergence test runs in block 0 of
        // launch n+1, or in a one-block check at the end of a batch.  After a
        // passed test later launches return at once, so the converged chi is
        // in the buffer of the launch whose sweep passed.
        const int w = std::is_same<T, float>::value ? 1 : 0;
        void* cur = dchi;
        void* alt = c->sor_alt[w];
        const int ntx = (c->Nx + 63) / 64, nb = ntx * ((c->Ny + rows - 1) / rows);
        const int64_t nfull = max_sweeps / K * K;
        int64_t n = 0;
        while (l + K - 1 <= nfull && !*host_done) {
            const int64_t lend = std::min<int64_t>(nfull, l + batch - 1);
            batch = std::min(2 * batch, batch_max);
            int64_t pending = 0;   // sweep of the last launch's test not yet run (0: none)
            for (; l + K - 1 <= lend; l += K, ++n) {
                const bool check = ((l + K - 1) % every) == 0 || l + K - 1 == max_sweeps;
                const int par = (int)(n & 1);
                if (K == 4) sor_block_launch<T, 4, 16>(c, cur, alt, dB, tabs, omega, tol, bnorm, l, check, ref_mode, par, pending, mid);
                else if (K == 2 && rows == 32) sor_block_launch<T, 2, 32>(c, cur, alt, dB, tabs, omega, tol, bnorm, l, check, ref_mode, par, pending, mid);
                else if (K == 2) sor_block_launch<T, 2, 16>(c, cur, alt, dB, tabs, omega, tol, bnorm, l, check, ref_mode, par, pending, mid);
                else if (rows == 32) sor_block_launch<T, 1, 32>(c, cur, alt, dB, tabs, omega, tol, bnorm, l, check, ref_mode, par, pending);
                else sor_block_launch<T, 1, 16>(c, cur, alt, dB, tabs, omega, tol, bnorm, l, check, ref_mode, par, pending);
                pending = check ? l + K - 1 : 0;
                std::swap(cur, alt);
            }
            if (pending) {   // the last launch's test(s): its earlier sweeps in order (mid), then its last
                if (mid)
                    for (int q = 0; q < K - 1; ++q)
                        hipLaunchKernelGGL(swm::sor_check, dim3(1), dim3(256), 0, c->stream,
                                           (const double*)c->sor_partial_mid + (size_t)(2 * q + ((n - 1) & 1)) * nb,
                                           nb, (int)(pending - (K - 1) + q), tol, bnorm, 0, c->sor_done,
                                           c->sor_iters, c->sor_norm);
                hipLaunchKernelGGL(swm::sor_check, dim3(1), dim3(256), 0, c->stream,
                                   (const double*)c->sor_partial + (size_t)((n - 1) & 1) * nb, nb, (int)pending, tol,
                                   bnorm, ref_mode ? 0 : 1, c->sor_done, c->sor_iters, c->sor_norm);
            }
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, copy_sync(c, host_done, c->sor_done, sizeof(int), hipMemcpyDeviceToHost));
        }
        if (*host_done) {   // the buffer of the launch whose test passed
            long long it = 0;
            HIPCHK(c, copy_sync(c, &it, c->sor_iters, sizeof(it), hipMemcpyDeviceToHost));
            const int64_t nconv = (it - 1) / K;   // launch n did sweeps n*K+1 .. (n+1)*K
            const int64_t q = (it - 1) % K;       // the passed sweep's place in it
            if (mid && q < K - 1) cur = c->sor_mid[w][q][nconv & 1];
            else cur = (nconv & 1) ? dchi : c->sor_alt[w];
        }
        if (cur != dchi)   // the result in the caller's buffer
            HIPCHK(c, hipMemcpyAsync(dchi, cur, (size_t)c->P * c->Ny * sizeof(T), hipMemcpyDeviceToDevice,
                                     c->stream));
        if (*host_done) return 0;
    }
    while (l <= max_sweeps && !*host_done) {
        const int64_t lend = std::min<int64_t>(max_sweeps, l + batch - 1);
        batch = std::min(2 * batch, batch_max);
        for (; l <= lend; ++l) {
            const bool check = (l % every) == 0 || l == max_sweeps;
            for (int color = 0; color < 2; ++color) {
                swm::SorArgs<T> a{};
                a.Nx = c->Nx; a.Ny = c->Ny; a.P = c->P; a.color = color; a.sweep = (int)l;
                a.chi = (T*)dchi; a.B = (const T*)dB; a.mask = c->mask;
                a.relax = ref_mode ? (const double*)c->sor_relax : nullptr;
                a.nrelax = c->sor_nrelax; a.omega = omega;
                a.c1t = (const T*)tabs[0]; a.c3t = (const T*)tabs[1]; a.c4t = (const T*)tabs[2]; a.c5t = (const T*)tabs[3]; a.c5r = (const T*)tabs[4];
                a.partial = c->sor_partial + color * nblk;
                a.done = c->sor_done;
                a.residual_mode = ref_mode ? 0 : 1;
                a.accumulate = check ? 1 : 0;
                hipLaunchKernelGGL(swm::sor_sweep<T>, dim3(nblk), dim3(256), 0, c->stream, a);
            }
            if (check)
                hipLaunchKernelGGL(swm::sor_check, dim3(1), dim3(256), 0, c->stream, (const double*)c->sor_partial,
                                   2 * nblk, (int)l, tol, bnorm, ref_mode ? 0 : 1, c->sor_done, c->sor_iters,
                                   c->sor_norm);
        }
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, copy_sync(c, host_done, c->sor_done, sizeof(int), hipMemcpyDeviceToHost));
    }
    return 0;
}

int sor_common(swm_ctx* c, const double* B, double* chi, int ref_mode, double omega, double tol,
                      int64_t max_sweeps, int fp32_sweeps, int64_t* iters, double* norm) {
    const int Nx = c->Nx, Ny = c->Ny;
    const int half = (Nx + 1) / 2;
    const int nblk = (int)(((long)Ny * half + 255) / 256);
    const bool f32 = fp32_sweeps != 0 || c->fp32;
    int rc = sor_prepare(c, f32, nblk);
    if (rc) return rc;
    const int w = f32 ? 1 : 0;
    void* dchi = c->sor_buf[w][0];
    void* dB = c->sor_buf[w][1];
    const size_t n = (size_t)Nx * Ny;
    const size_t pitch_b = (size_t)c->P * (f32 ? 4 : 8);
    if (f32) {
        std::vector<float> fc(n), fb(n);
        for (size_t k = 0; k < n; ++k) { fc[k] = (float)chi[k]; fb[k] = (float)B[k]; }
        HIPCHK(c, copy2d_sync(c, dchi, pitch_b, fc.data(), Nx * 4, Nx * 4, Ny, hipMemcpyHostToDevice));
        HIPCHK(c, copy2d_sync(c, dB, pitch_b, fb.data(), Nx * 4, Nx * 4, Ny, hipMemcpyHostToDevice));
    } else {
        HIPCHK(c, copy2d_sync(c, dchi, pitch_b, chi, Nx * 8, Nx * 8, Ny, hipMemcpyHostToDevice));
        HIPCHK(c, copy2d_sync(c, dB, pitch_b, B, Nx * 8, Nx * 8, Ny, hipMemcpyHostToDevice));
    }
    // ||B|| over ocean eta points, in FP64 whatever the sweep precision
    double bnorm = 0.0;
    if (!ref_mode) {
        const int8_t* me = c->hmask[3].data();
        for (size_t k = 0; k < n; ++k) {
            const double b = f32 ? (double)(float)B[k] : B[k];
            if (me[k]) bnorm += b * b;
        }
        bnorm = std::sqrt(bnorm);
    }
    HIPCHK(c, hipMemsetAsync(c->sor_done, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->sor_iters, 0, sizeof(long long), c->stream));
    int host_done = 0;
    rc = f32 ? sor_sweeps<float>(c, dchi, dB, c->sor_ftab[w], ref_mode, omega, tol, bnorm, max_sweeps, nblk, &host_done)
             : sor_sweeps<double>(c, dchi, dB, c->sor_ftab[w], ref_mode, omega, tol, bnorm, max_sweeps, nblk, &host_done);
    if (rc) return rc;
    long long it = 0;
    double nrm = 0.0;
    HIPCHK(c, copy_sync(c, &it, c->sor_iters, sizeof(it), hipMemcpyDeviceToHost));
    HIPCHK(c, copy_sync(c, &nrm, c->sor_norm, sizeof(nrm), hipMemcpyDeviceToHost));
    if (f32) {
        std::vector<float> fc(n);
        HIPCHK(c, copy2d_sync(c, fc.data(), Nx * 4, dchi, pitch_b, Nx * 4, Ny, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) chi[k] = fc[k];
    } else {
        HIPCHK(c, copy2d_sync(c, chi, Nx * 8, dchi, pitch_b, Nx * 8, Ny, hipMemcpyDeviceToHost));
    }
    *iters = it;
    *norm = nrm;
    c->chi_valid[w] = true;
    c->chi_last = w;
    return host_done ? 0 : 1;
}

template <typename TF, typename TS>
int velpot_t(swm_ctx* c, double omega, double rel_tol, int64_t max_sweeps, bool reset, int64_t* iters,
             double* rel_res) {
    const int Nx = c->Nx, Ny = c->Ny;
    const int half = (Nx + 1) / 2;
    const int nblk = (int)(((long)Ny * half + 255) / 256);
    const bool f32 = std::is_same<TS, float>::value;
    int rc = sor_prepare(c, f32, nblk);
    if (rc) return rc;
    const int w = f32 ? 1 : 0;
    void* dchi = c->sor_buf[w][0];
    TS* dB = (TS*)c->sor_buf[w][1];
    const long n = (long)Nx * Ny;
    if (!c->norm_buf && !(c->norm_buf = (double*)dalloc(c, sizeof(double) * 1025)))
        return fail(c, "out of device memory", 3);
    // right-hand side: computeDivergence of the current velocity (calc_lib.f90:500-525)
    hipLaunchKernelGGL((swm::divergence_kernel<TF, TS>), dim3((n + 255) / 256), dim3(256), 0, c->stream, Nx, Ny, c->P,
                       (const TF*)plane_ptr(c, swm::PL_U + c->cur), (const TF*)plane_ptr(c, swm::PL_V + c->cur),
                       c->mask, (const TF*)c->dza, (const TF*)c->tab[T_CV], (const TF*)c->tab[T_CE],
                       (TF)(c->d.A * c->d.dTheta), (TF)c->d.lbc, dB);
    HIPCHK(c, hipGetLastError());
    if (reset || !c->chi_valid[w])
        HIPCHK(c, hipMemsetAsync(dchi, 0, (size_t)c->P * Ny * sizeof(TS), c->stream));
    // ||B|| over ocean eta points in FP64 (fixed order)
    const int nb = (int)std::min<long>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(swm::sq_norm_kernel<TS>, dim3(nb), dim3(256), 0, c->stream, Nx, Ny, c->P, (const TS*)dB, c->mask,
                       c->norm_buf);
    hipLaunchKernelGGL(swm::sq_norm_final, dim3(1), dim3(256), 0, c->stream, (const double*)c->norm_buf, nb,
                       c->norm_buf + 1024);
    HIPCHK(c, hipGetLastError());
    double bnorm = 0.0;
    HIPCHK(c, hipMemsetAsync(c->sor_done, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->sor_iters, 0, sizeof(long long), c->stream));
    HIPCHK(c, copy_sync(c, &bnorm, c->norm_buf + 1024, sizeof(double), hipMemcpyDeviceToHost));
    int host_done = 0;
    if ((rc = sor_sweeps<TS>(c, dchi, dB, c->sor_ftab[w], false, omega, rel_tol, bnorm, max_sweeps, nblk, &host_done)))
        return rc;
    long long it = 0;
    double nrm = 0.0;
    HIPCHK(c, copy_sync(c, &it, c->sor_iters, sizeof(it), hipMemcpyDeviceToHost));
    HIPCHK(c, copy_sync(c, &nrm, c->sor_norm, sizeof(nrm), hipMemcpyDeviceToHost));
    *iters = it;
    *rel_res = nrm;
    c->chi_valid[w] = true;
    c->chi_last = w;
    return host_done ? 0 : 1;
}

// computeVelocityPotential of the elliptic-solver build (calc_lib.f90:500-525)
// on the device: B = computeDivergence of u, v at N0 (607-622), then
// main_ElSolv_SOR (ElSolv_SOR.f90:150-222: Chebyshev omega, stop when
// sqrt(sum res^2) < epsilon) on chi kept between calls, as calc_lib::chi is,
// with epsilon = EPS (calc_lib.h:12, the default-kind literal 1e-4, i.e. its
// single-precision value) and first_guess = (itt == 1).  At most once per
// step, as chi_computed is reset by advanceCalcLib (calc_lib.f90:198-201).
// Sweeps in the context's precision.
template <typename T>
int elliptic_chi_t(swm_ctx* c) {
    if (c->chi_ref_itt == c->itt) return 0;
    const int Nx = c->Nx, Ny = c->Ny;
    const int nblk = (int)(((long)Ny * ((Nx + 1) / 2) + 255) / 256);
    const bool f32 = std::is_same<T, float>::value;
    if (int rc = sor_prepare(c, f32, nblk)) return rc;
    const size_t bytes = (size_t)c->P * Ny * sizeof(T);
    if (!c->chi_ref) {
        if (!(c->chi_ref = dalloc(c, bytes)) || !(c->chi_refB = dalloc(c, bytes))) return fail(c, "out of device memory", 3);
        HIPCHK(c, zero_sync(c, c->chi_ref, bytes));   // calc_lib::chi starts at 0
    }
    const long n = (long)Nx * Ny;
    hipLaunchKernelGGL((swm::divergence_kernel<T, T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, Nx, Ny,
                       c->P, (const T*)plane_ptr(c, swm::PL_U + c->cur), (const T*)plane_ptr(c, swm::PL_V + c->cur),
                       c->mask, (const T*)c->dza, (const T*)c->tab[T_CV], (const T*)c->tab[T_CE],
                       (T)(c->d.A * c->d.dTheta), (T)c->d.lbc, (T*)c->chi_refB);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemsetAsync(c->sor_done, 0, sizeof(int), c->stream));
    HIPCHK(c, hipMemsetAsync(c->sor_iters, 0, sizeof(long long), c->stream));
    const float mx = (float)std::max(Nx, Ny);
    const int64_t max_count = c->itt == 1 ? (int64_t)lrintf(1000.f * mx) : (int64_t)lrintf(10.f * mx);
    const double eps = (double)1e-4f;
    int host_done = 0;
    const int w = f32 ? 1 : 0;
    if (int rc = sor_sweeps<T>(c, c->chi_ref, c->chi_refB, c->sor_ftab[w], true, 1.0, eps, 0.0, max_count, nblk,
                               &host_done))
        return rc;
    long long it = 0;
    double nrm = 0.0;
    HIPCHK(c, copy_sync(c, &it, c->sor_iters, sizeof(it), hipMemcpyDeviceToHost));
    HIPCHK(c, copy_sync(c, &nrm, c->sor_norm, sizeof(nrm), hipMemcpyDeviceToHost));
    c->chi_ref_iters = host_done ? it : max_count;
    c->chi_ref_anorm = nrm;
    c->chi_ref_itt = c->itt;
    return 0;
}

int elliptic_chi(swm_ctx* c) { return c->fp32 ? elliptic_chi_t<float>(c) : elliptic_chi_t<double>(c); }

// computeNonDivergentFlowField (calc_lib.f90:548-590) into nd_buf: u - du,
// v - dv with (du, dv) = computeGradient(chi, eta_grid).
int elliptic_nd(swm_ctx* c) {
    if (int rc = elliptic_chi(c)) return rc;
    const size_t bytes = (size_t)c->P * c->Ny * sizeof(double);
    for (auto& b : c->nd_buf)
        if (!b && !(b = (double*)dalloc(c, bytes))) return fail(c, "out of device memory", 3);
    if (!c->gza) {
        std::vector<double> g(2 * (size_t)c->Ny);
        const double A = c->d.A, dL = c->d.dLambda;
        for (int j = 0; j < c->Ny; ++j) {
            g[2 * j] = 1.0 / (A * c->cu_host[j] * dL);
            g[2 * j + 1] = c->d.lbc / (A * c->cu_host[j] * dL);
        }
        if (!(c->gza = (double*)dalloc(c, g.size() * 8))) return fail(c, "out of device memory", 3);
        HIPCHK(c, copy_sync(c, c->gza, g.data(), g.size() * 8, hipMemcpyHostToDevice));
    }
    const long n = (long)c->Nx * c->Ny;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (c->fp32)
        hipLaunchKernelGGL((swm::nondivergent_kernel<float, float>), grid, block, 0, c->stream, c->Nx, c->Ny, c->P,
                           c->mask, (const float*)plane_ptr(c, swm::PL_U + c->cur),
                           (const float*)plane_ptr(c, swm::PL_V + c->cur), (const float*)c->chi_ref,
                           (const double*)c->gza, c->d.lbc, c->d.A * c->d.dTheta, c->nd_buf[0], c->nd_buf[1]);
    else
        hipLaunchKernelGGL((swm::nondivergent_kernel<double, double>), grid, block, 0, c->stream, c->Nx, c->Ny, c->P,
                           c->mask, (const double*)plane_ptr(c, swm::PL_U + c->cur),
                           (const double*)plane_ptr(c, swm::PL_V + c->cur), (const double*)c->chi_ref,
                           (const double*)c->gza, c->d.lbc, c->d.A * c->d.dTheta, c->nd_buf[0], c->nd_buf[1]);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// FP64 conservation sums of the current state into dout (3 values, copied to
// the host out): sums_kernel's per-block partials in `partial` (6 per block,
// at most 1024 blocks), the eta rows' cos(lat) in `ce`.
int sums_launch(swm_ctx* c, int32_t tracer_id, double* partial, double* dout, const double* ce, double* out,
                bool grow = true) {
    const long n = (long)c->Nx * c->Ny;
    const int nb = (int)std::min<long>(1024, (n + 255) / 256);
    const int fn = c->d.variant == SWM_FULLY_NONLINEAR;
    const int ch_plane = tracer_id >= 0 ? swm::PL_TRC + swm::PL_PER_TRC * tracer_id + c->cur : -1;
    if (c->fp32)
        hipLaunchKernelGGL(swm::sums_kernel<float>, dim3(nb), dim3(256), 0, c->stream,
                           (const Statics<float>*)c->statics, fn, (const float*)plane_ptr(c, swm::PL_E + c->cur),
                           (const float*)plane_ptr(c, swm::PL_U + c->cur), (const float*)plane_ptr(c, swm::PL_V + c->cur),
                           ch_plane >= 0 ? (const float*)plane_ptr(c, ch_plane) : nullptr, ce, partial);
    else
        hipLaunchKernelGGL(swm::sums_kernel<double>, dim3(nb), dim3(256), 0, c->stream,
                           (const Statics<double>*)c->statics, fn, (const double*)plane_ptr(c, swm::PL_E + c->cur),
                           (const double*)plane_ptr(c, swm::PL_U + c->cur),
                           (const double*)plane_ptr(c, swm::PL_V + c->cur),
                           ch_plane >= 0 ? (const double*)plane_ptr(c, ch_plane) : nullptr, ce, partial);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(swm::sums_final, dim3(1), dim3(256), 0, c->stream, (const double*)partial, nb, dout);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, d2h(c->stream, out, dout, 3 * sizeof(double), grow));
    if (tracer_id < 0) out[2] = 0.0;
    return 0;
}

// Scratch of swm_gpu_sums_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Layout (doubles): [0, 6144) per-block
// partials, [6144, 6147) the result, [6152, 6152 + kNoAllocMaxRows) the eta
// rows' cos(lat).  Every call on a GPU uses that GPU's copy, so calls that
// overlap on one GPU (two host threads, two streams) race.
constexpr int kNoAllocMaxRows = 16384;
constexpr size_t kScratchResult = 6 * 1024;
constexpr size_t kScratchRows = 6 * 1024 + 8;
constexpr size_t kScratchDoubles = kScratchRows + kNoAllocMaxRows;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is only a stand-in for the symbol).
int scratch_ptr(swm_ctx* c, double** out) {
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(c, "device index beyond the scratch table", 4);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(c, "hipGetSymbolAddress(scratch) returned no address", 3);
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

}  // namespace

extern "C" {

int swm_gpu_sor(swm_ctx* c, const double* B, double* chi, double epsilon, int32_t first_guess, int64_t* iters,
                double* anorm) {
    HipLock lock(g_hip_mu);
    if (!c || !B || !chi || !iters || !anorm) return 1;
    const int mx = std::max(c->Nx, c->Ny);
    const int64_t max_count = first_guess ? (int64_t)lrintf(1000.f * (float)mx) : (int64_t)lrintf(10.f * (float)mx);
    return sor_common(c, B, chi, 1, 0.0, epsilon, max_count, 0, iters, anorm);
}

int swm_gpu_sor_fixed(swm_ctx* c, const double* B, double* chi, double omega, double rel_tol, int64_t max_sweeps,
                      int32_t fp32_sweeps, int64_t* iters, double* rel_res) {
    HipLock lock(g_hip_mu);
    if (!c || !B || !chi || !iters || !rel_res) return 1;
    return sor_common(c, B, chi, 0, omega, rel_tol, max_sweeps, fp32_sweeps, iters, rel_res);
}

int swm_gpu_set_kernel_shape(swm_ctx* c, int32_t shape) {
    if (!c) return 1;
    if (shape < 0 || shape >= kNumTiles) return fail(c, "no such kernel shape", 1);
    if (c->d.variant == SWM_LINEARISED_MEAN_STATE && kTiles[shape].march)
        return fail(c, "LINEARISED_MEAN_STATE runs on the LDS tile kernels (shapes 0-4, 11, 12)", 4);
    if ((c->dyn[0].rec || c->dyn[1].rec) && kTiles[shape].march == 1)
        return fail(c, "DYNFROMFILE velocities are read by the tile kernels (shapes 0-4, 9-12)", 4);
    HipLock lock(g_hip_mu);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    c->tile = shape;
    return 0;
}

int swm_gpu_kernel_shape_count(void) { return kNumTiles; }

int swm_gpu_add_forcing_stream(swm_ctx* c, int32_t kind, int32_t component, int32_t nrec, double rec_dt,
                               int32_t chunk_size, const double* records, int32_t* stream_id) {
    if (!c || !records || !stream_id) return 1;
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    if (c->batch) return fail(c, "context is attached to a batch; add streams before attaching", 1);
    if (kind != SWM_STREAM_WIND && kind != SWM_STREAM_CUSTOM && kind != SWM_STREAM_HEATING)
        return fail(c, "unknown forcing stream kind", 1);
    if (component < SWM_COMP_X || component > SWM_COMP_ETA) return fail(c, "unknown component", 1);
    if (kind == SWM_STREAM_WIND && component == SWM_COMP_ETA)
        return fail(c, "windstress acts on the momentum equations (SWM_COMP_X, SWM_COMP_Y)", 1);
    if (kind == SWM_STREAM_HEATING && component != SWM_COMP_ETA)
        return fail(c, "heating acts on the continuity equation (SWM_COMP_ETA)", 1);
    if (nrec < 1) return fail(c, "a stream needs at least one record", 1);
    if (nrec > 1 && chunk_size < 2) return fail(c, "a time-dependent dataset needs chunk_size >= 2", 1);
    if (nrec > 1 && !(rec_dt > 0.0)) return fail(c, "record spacing must be positive", 1);
    if (c->streams.size() >= 4) return fail(c, "at most 4 time-interpolated forcing streams", 1);
    FStream st;
    st.kind = kind;
    st.comp = component;
    st.rec = dalloc(c, (size_t)nrec * c->plane * c->esz);
    if (!st.rec) return fail(c, "out of device memory for the stream records", 3);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, zero_sync(c, st.rec, (size_t)nrec * c->plane * c->esz));
    for (int32_t r = 0; r < nrec; ++r) {
        int rc = upload_grid(c, (char*)st.rec + (size_t)r * c->plane * c->esz, records + (size_t)r * c->Nx * c->Ny);
        if (rc) return rc;
    }
    st.idx.init(nrec, chunk_size, rec_dt);
    c->streams.push_back(std::move(st));
    drop_graphs(c);
    *stream_id = (int32_t)c->streams.size() - 1;
    return write_statics(c);
}

int swm_gpu_add_dyn_stream(swm_ctx* c, int32_t field, int32_t nrec, double rec_dt, int32_t chunk_size,
                           const double* records) {
    if (!c || !records) return 1;
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    if (c->batch) return fail(c, "context is attached to a batch; add datasets before attaching", 1);
    if (field < SWM_COMP_X || field > SWM_COMP_ETA) return fail(c, "unknown field (SWM_COMP_X, _Y, _ETA)", 1);
    if (c->dyn[field].rec) return fail(c, "one DYNFROMFILE dataset per field", 1);
    if (nrec < 1) return fail(c, "a dataset needs at least one record", 1);
    if (nrec > 1 && chunk_size < 2) return fail(c, "a time-dependent dataset needs chunk_size >= 2", 1);
    if (nrec > 1 && !(rec_dt > 0.0)) return fail(c, "record spacing must be positive", 1);
    if (kTiles[c->tile].march == 1 && field != SWM_COMP_ETA) c->tile = 11;   // the tile kernels read the velocities
    FStream st;
    st.kind = -1;
    st.comp = field;
    st.rec = dalloc(c, (size_t)nrec * c->plane * c->esz);
    if (!st.rec) return fail(c, "out of device memory for the dataset records", 3);
    HIPCHK(c, zero_sync(c, st.rec, (size_t)nrec * c->plane * c->esz));
    for (int32_t r = 0; r < nrec; ++r) {
        int rc = upload_grid(c, (char*)st.rec + (size_t)r * c->plane * c->esz, records + (size_t)r * c->Nx * c->Ny);
        if (rc) return rc;
    }
    st.idx.init(nrec, chunk_size, rec_dt);
    // DFF_initDynFromFile's DFF_advance at itt = 0 (model.f90:108-113) reads the first time
    int64_t r0, r1;
    bool ex;
    double qa, qb;
    st.idx.query(0.0, r0, r1, ex, qa, qb);
    c->dyn[field] = std::move(st);
    drop_graphs(c);
    return write_statics(c);
}

int swm_gpu_tune(swm_ctx* c, int32_t steps, int32_t* best_shape, double* times_ms) {
    if (!c || !best_shape) return 1;
    if (c->batch) return fail(c, "context is attached to a batch; tune the batch instead", 1);
    if (steps < 1) return fail(c, "steps must be >= 1", 1);
    HipLock lock(g_hip_mu);
    const int ntrc = (int)c->trc.size();
    const size_t bytes = (size_t)(swm::PL_TRC + swm::PL_PER_TRC * ntrc) * c->plane * c->esz;
    void* snap = nullptr;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (hipMalloc(&snap, bytes) != hipSuccess) return fail(c, "out of device memory for the tuning snapshot", 3);
    const int64_t itt0 = c->itt;
    const int cur0 = c->cur, ring0 = c->ring;
    std::vector<bool> cb0;
    for (auto& t : c->trc) cb0.push_back(t.cbuf_valid);
    const int tile0 = c->tile;
    const bool graph0 = c->use_graph;
    auto restore = [&]() -> hipError_t {
        c->itt = itt0;
        c->cur = cur0;
        c->ring = ring0;
        for (size_t n = 0; n < cb0.size(); ++n) c->trc[n].cbuf_valid = cb0[n];
        return hipMemcpyAsync(c->state, snap, bytes, hipMemcpyDeviceToDevice, c->stream);
    };
    int rc = 0;
    double best = 1e300;
    int best_k = tile0;
    if (hipMemcpyAsync(snap, c->state, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) rc = 3;
    c->use_graph = false;   // time the kernels, not graph capture
    for (int k = 0; k < kNumTiles && !rc; ++k) {
        if ((c->d.variant == SWM_LINEARISED_MEAN_STATE && kTiles[k].march) ||
            ((c->dyn[0].rec || c->dyn[1].rec) && kTiles[k].march == 1)) {
            if (times_ms) times_ms[k] = -1.0;   // not available for this variant / these datasets
            continue;
        }
        c->tile = k;
        float ms = 0.f;
        // untimed run first (first-launch costs), then the timed run from the same state
        rc = c->fp32 ? step_many<float>(c, steps) : step_many<double>(c, steps);
        if (rc || restore() != hipSuccess) { rc = rc ? rc : 3; break; }
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) { rc = 3; break; }
        rc = c->fp32 ? step_many<float>(c, steps) : step_many<double>(c, steps);
        if (rc) break;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess || restore() != hipSuccess) {
            rc = 3;
            break;
        }
        if (times_ms) times_ms[k] = ms / steps;
        if (ms < best) { best = ms; best_k = k; }
    }
    // launch form of the best shape: replay of the captured six-step graphs
    // against direct launches (the untimed pass captures, the timed one replays)
    bool graphs = graph0;
    c->ms_direct = best / steps;
    c->ms_graph = -1.0;
    if (!rc && steps >= 6) {
        c->tile = best_k;
        c->use_graph = true;
        drop_graphs(c);
        if (steady_ok(c)) {
            float ms = 0.f;
            rc = c->fp32 ? step_many<float>(c, steps) : step_many<double>(c, steps);
            if (!rc && restore() == hipSuccess && hipEventRecord(c->ev0, c->stream) == hipSuccess) {
                rc = c->fp32 ? step_many<float>(c, steps) : step_many<double>(c, steps);
                if (!rc && (hipEventRecord(c->ev1, c->stream) != hipSuccess ||
                            hipEventSynchronize(c->ev1) != hipSuccess ||
                            hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess || restore() != hipSuccess))
                    rc = 3;
            } else if (!rc) {
                rc = 3;
            }
            if (!rc) {
                c->ms_graph = ms / steps;
                graphs = ms < best;
            }
        }
    }
    if (rc) (void)restore();
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(snap);
    c->use_graph = rc ? graph0 : graphs;
    c->tile = rc ? tile0 : best_k;
    drop_graphs(c);
    if (rc) return rc == 3 ? fail(c, "tuning run failed: " + c->err, 3) : rc;
    *best_shape = best_k;
    return 0;
}

int swm_gpu_avg_add(swm_ctx* c, const char* name, int32_t tracer_id, int32_t square, int32_t* avg_id) {
    if (!c || !name || !avg_id) return 1;
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    if (c->batch) return fail(c, "context is attached to a batch; register averages before attaching", 1);
    std::string s(name);
    for (auto& ch : s) ch = (char)toupper((unsigned char)ch);
    int field = -1;
    if (s == "SWM_U" || s == "U") field = 0;
    else if (s == "SWM_V" || s == "V") field = 1;
    else if (s == "SWM_ETA" || s == "ETA") field = 2;
    else if (s == "CH") field = 3;
    if (field < 0) return fail(c, "averages are kept for SWM_U, SWM_V, SWM_ETA and CH", 1);
    if (field == 3 && (tracer_id < 0 || tracer_id >= (int32_t)c->trc.size())) return fail(c, "no such tracer", 1);
    swm_ctx::Avg av{field, field == 3 ? tracer_id : -1, square ? 1 : 0, nullptr, 0.0};
    const size_t bytes = (size_t)c->P * c->Ny * sizeof(double);
    if (!(av.buf = (double*)dalloc(c, bytes))) return fail(c, "out of device memory", 3);
    HIPCHK(c, hipMemsetAsync(av.buf, 0, bytes, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);   // replayed steps now also accumulate this buffer
    c->avgs.push_back(av);
    *avg_id = (int32_t)c->avgs.size() - 1;
    return 0;
}

int swm_gpu_avg_read(swm_ctx* c, int32_t avg_id, double* host_out, double* weight, int32_t reset) {
    if (!c || !host_out) return 1;
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    if (avg_id < 0 || avg_id >= (int32_t)c->avgs.size()) return fail(c, "no such average", 1);
    swm_ctx::Avg& av = c->avgs[avg_id];
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->Nx * c->Ny;
    HIPCHK(c, copy2d_sync(c, host_out, c->Nx * 8, av.buf, c->P * 8, c->Nx * 8, c->Ny, hipMemcpyDeviceToHost));
    // task%buffer = task%buffer / task%bufferCount
    for (size_t k = 0; k < n; ++k) host_out[k] = host_out[k] / av.weight;
    if (weight) *weight = av.weight;
    if (reset) {
        HIPCHK(c, hipMemsetAsync(av.buf, 0, (size_t)c->P * c->Ny * sizeof(double), c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        av.weight = 0.0;
    }
    return 0;
}

int swm_gpu_set_sor_check_every(swm_ctx* c, int32_t every) {
    if (!c) return 1;
    if (every < 1 || every > 64) return fail(c, "check cadence must be 1..64 sweeps", 1);
    if (c->sor_halo > 2 && every % (c->sor_halo / 2))
        return fail(c, "the cadence must be a multiple of the SOR halo's sweeps per launch", 1);
    c->sor_every = every;
    return 0;
}

int swm_gpu_set_sor_tile_rows(swm_ctx* c, int32_t rows) {
    if (!c) return 1;
    if (rows != 16 && rows != 32) return fail(c, "SOR tile rows must be 16 or 32", 1);
    if (rows == 32 && c->sor_halo == 8) return fail(c, "a 4-sweep halo (8 cells) takes 16-row tiles", 1);
    c->sor_rows = rows;
    return 0;
}

int swm_gpu_set_sor_halo(swm_ctx* c, int32_t halo) {
    if (!c) return 1;
    if (halo != -1 && halo != 0 && halo != 2 && halo != 4 && halo != 8)
        return fail(c, "SOR halo must be -1 (from the cadence), 0, 2, 4 or 8 cells", 1);
    if (halo == 8 && c->sor_rows == 32) return fail(c, "a 4-sweep halo (8 cells) takes 16-row tiles", 1);
    if (halo > 2 && c->sor_every % (halo / 2))
        return fail(c, "the SOR halo's sweeps per launch (halo / 2) must divide the test cadence", 1);
    c->sor_halo = halo;
    return 0;
}

int swm_gpu_set_launch_mode(swm_ctx* c, int32_t graphs) {
    if (!c) return 1;
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->use_graph = graphs != 0;
    drop_graphs(c);
    return 0;
}

int swm_gpu_launch_mode(const swm_ctx* c, int32_t* graphs, double* ms_direct, double* ms_graph) {
    if (!c || !graphs) return 1;
    *graphs = c->use_graph ? 1 : 0;
    if (ms_direct) *ms_direct = c->ms_direct;
    if (ms_graph) *ms_graph = c->ms_graph;
    return 0;
}

int swm_gpu_velocity_potential(swm_ctx* c, double omega, double rel_tol, int64_t max_sweeps, int32_t fp32_sweeps,
                               int32_t reset, int64_t* iters, double* rel_res) {
    if (!c || !iters || !rel_res) return 1;
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    if (!(omega > 0.0 && omega < 2.0)) return fail(c, "omega must lie in (0, 2)", 1);
    if (max_sweeps < 1) return fail(c, "max_sweeps must be >= 1", 1);
    if (c->fp32) return velpot_t<float, float>(c, omega, rel_tol, max_sweeps, reset != 0, iters, rel_res);
    if (fp32_sweeps) return velpot_t<double, float>(c, omega, rel_tol, max_sweeps, reset != 0, iters, rel_res);
    return velpot_t<double, double>(c, omega, rel_tol, max_sweeps, reset != 0, iters, rel_res);
}

int swm_gpu_elliptic_stats(swm_ctx* c, int64_t* iters, double* anorm) {
    if (!c || !iters || !anorm) return 1;
    HipLock lock(g_hip_mu);
    if (c->chi_ref_itt < 0) return fail(c, "no velocity potential of the elliptic-solver build has been computed", 1);
    *iters = c->chi_ref_iters;
    *anorm = c->chi_ref_anorm;
    return 0;
}

int swm_gpu_divergence(swm_ctx* c, double* host_div) {
    HipLock lock(g_hip_mu);
    if (!c || !host_div) return 1;
    if (int rc = member_refresh(c)) return rc;
    const int n = c->Nx * c->Ny;
    const int nb = (n + 255) / 256;
    if (c->fp32)
        hipLaunchKernelGGL(swm::divergence_kernel<float>, dim3(nb), dim3(256), 0, c->stream, c->Nx, c->Ny, c->P,
                           (const float*)plane_ptr(c, swm::PL_U + c->cur), (const float*)plane_ptr(c, swm::PL_V + c->cur),
                           c->mask, (const float*)c->dza,
                           (const float*)c->tab[T_CV], (const float*)c->tab[T_CE], (float)(c->d.A * c->d.dTheta),
                           (float)c->d.lbc, (float*)c->scratch);
    else
        hipLaunchKernelGGL(swm::divergence_kernel<double>, dim3(nb), dim3(256), 0, c->stream, c->Nx, c->Ny, c->P,
                           (const double*)plane_ptr(c, swm::PL_U + c->cur),
                           (const double*)plane_ptr(c, swm::PL_V + c->cur), c->mask, (const double*)c->dza,
                           (const double*)c->tab[T_CV], (const double*)c->tab[T_CE], c->d.A * c->d.dTheta, c->d.lbc,
                           (double*)c->scratch);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return download_grid(c, host_div, c->scratch);
}

int swm_gpu_sums(swm_ctx* c, int32_t tracer_id, double* out) {
    HipLock lock(g_hip_mu);
    if (!c || !out) return 1;
    if (int rc = member_refresh(c)) return rc;
    if (tracer_id >= (int32_t)c->trc.size()) return fail(c, "tracer id out of range", 1);
    if (!c->sums_buf) {
        c->sums_buf = (double*)dalloc(c, sizeof(double) * (6 * 1024 + 3));
        c->ce64 = (double*)dalloc(c, sizeof(double) * c->Ny);
        if (!c->sums_buf || !c->ce64) return fail(c, "out of device memory", 3);
        HIPCHK(c, copy_sync(c, c->ce64, c->ce_host.data(), sizeof(double) * c->Ny, hipMemcpyHostToDevice));
    }
    return sums_launch(c, tracer_id, c->sums_buf, c->sums_buf + 6 * 1024, c->ce64, out);
}

int swm_gpu_sums_noalloc(swm_ctx* c, int32_t tracer_id, double* out) {
    if (!c || !out) return 1;
    if (c->Ny > kNoAllocMaxRows)
        return fail(c, "swm_gpu_sums_noalloc: Ny = " + std::to_string(c->Ny) + " exceeds the " +
                       std::to_string(kNoAllocMaxRows) + " rows of its static scratch", 4);
    HipLock lock(g_hip_mu);
    if (int rc = member_refresh(c)) return rc;
    if (tracer_id >= (int32_t)c->trc.size()) return fail(c, "tracer id out of range", 1);
    double* ws = nullptr;
    if (int rc = scratch_ptr(c, &ws)) return rc;
    double* partial = ws;
    double* dout = ws + kScratchResult;
    double* ce = ws + kScratchRows;
    // the cos rows are this context's; the partials and the result are
    // rewritten in full by the kernels, so nothing needs clearing.  The host
    // transfers use the staging buffer that the context's creation made (at
    // least 1 MiB, more than these 16384 * 8 + 24 bytes) and never grow it.
    HIPCHK(c, h2d(c->stream, ce, c->ce_host.data(), sizeof(double) * c->Ny, false));
    return sums_launch(c, tracer_id, partial, dout, ce, out, false);
}

const char* swm_gpu_last_error(const swm_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

void swm_gpu_destroy(swm_ctx* c) {
    HipLock lock(g_hip_mu);
    if (!c) return;
    if (c->batch) (void)swm_gpu_batch_detach(c->batch, c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched ensembles: contexts attached to one batch advance together, one
// launch (plus a tiny state-advance launch) per step, replayed from a graph.
// ---------------------------------------------------------------------------
struct swm_batch {
    int fp32 = -1;                       // precision of the members (-1: none attached yet)
    int variant = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<swm_ctx*> mem;
    std::vector<swm::MemberStep> hstep;  // host copy of the members' step state
    bool on_device = false;              // the device copy is newer than hstep
    bool dirty = true;                   // member list changed since the last build
    void* dmem = nullptr;                // Member<T>[cap]
    swm::MemberStep* dstep = nullptr;    // MemberStep[cap]
    size_t cap = 0;
    int* downer = nullptr;               // member of each work item
    size_t owner_cap = 0;
    int* dqueue = nullptr;               // per-XCD tile-queue counters of the persistent kernel, then the
                                         // finished-block counter of either kernel (kBatchQueues + 1)
    int nwork = 0, ntrc_max = 0;
    hipGraphExec_t graph = nullptr;
    bool use_graph = true;               // SWM_GPU_GRAPH=0: launch each step directly (profiling)
    int shape = 1;                       // kBatchShapes[shape]: tile and launch form (32x8 one block
                                         // per tile until a tuning run picks another)
    double last_ms = 0.0;
    std::string err;
};

namespace {

// Tile and launch form of a batched step: one block per tile, or the
// persistent tile queue (a resident grid that drains per-XCD queues).
struct BatchShape { int tx, ty, queue; };
constexpr BatchShape kBatchShapes[] = {{64, 4, 0}, {32, 8, 0}, {64, 4, 1}, {32, 8, 1}};
constexpr int kNumBatchShapes = 4;

int bfail(swm_batch* b, const std::string& msg, int code = 2) {
    if (b) b->err = msg;
    return code;
}

#define BHIPCHK(b, call)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bfail((b), std::string(#call) + ": " + hipGetErrorString(e_), 3);           \
    } while (0)

void batch_drop_graph(swm_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
}

// Copy the device step state back and hand it to the member contexts.
int batch_pull(swm_batch* b) {
    if (!b->on_device) return 0;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    if (!b->hstep.empty())
        BHIPCHK(b, d2h(b->stream, b->hstep.data(), b->dstep, b->hstep.size() * sizeof(swm::MemberStep)));
    b->on_device = false;
    for (size_t k = 0; k < b->mem.size(); ++k) {
        swm_ctx* c = b->mem[k];
        const swm::MemberStep& h = b->hstep[k];
        c->itt = h.itt;
        c->cur = h.cur;
        c->ring = h.ring;
        for (size_t n = 0; n < c->trc.size(); ++n) c->trc[n].cbuf_valid = ((h.cbuf >> n) & 1) != 0;
    }
    return 0;
}

template <typename T>
int batch_build_t(swm_batch* b) {
    const size_t n = b->mem.size();
    if (n > b->cap) {
        if (b->dmem) (void)hipFree(b->dmem);
        if (b->dstep) (void)hipFree(b->dstep);
        b->dmem = nullptr;
        b->dstep = nullptr;
        b->cap = 0;
        const size_t cap = std::max<size_t>(n, 2 * b->cap + 16);
        BHIPCHK(b, hipMalloc(&b->dmem, cap * sizeof(swm::Member<T>)));
        BHIPCHK(b, hipMalloc((void**)&b->dstep, cap * sizeof(swm::MemberStep)));
        b->cap = cap;
    }
    std::vector<swm::Member<T>> hm(n);
    long work = 0;
    int ntrc_max = 0;
    for (size_t k = 0; k < n; ++k) {
        swm_ctx* c = b->mem[k];
        swm::Member<T>& m = hm[k];
        m.s = (const Statics<T>*)c->statics;
        m.state = (T*)c->state;
        m.plane = c->plane;
        const BatchShape bs = kBatchShapes[b->shape];
        m.ntx = (c->Nx + bs.tx - 1) / bs.tx;
        m.ntile = m.ntx * ((c->Ny + bs.ty - 1) / bs.ty);
        m.wbase = (int)work;
        m.ntr = (int)c->trc.size();
        for (int t = 0; t < 3; ++t)
            for (int q = 0; q < 3; ++q) m.dtc[t][q] = (T)(c->d.dt * c->ab[t][q]);
        work += m.ntile;
        ntrc_max = std::max(ntrc_max, m.ntr);
    }
    if (work > INT32_MAX) return bfail(b, "batch too large for one launch", 4);
    b->nwork = (int)work;
    b->ntrc_max = ntrc_max;
    if ((size_t)work > b->owner_cap) {
        if (b->downer) (void)hipFree(b->downer);
        b->downer = nullptr;
        b->owner_cap = 0;
        BHIPCHK(b, hipMalloc((void**)&b->downer, (size_t)work * sizeof(int)));
        b->owner_cap = (size_t)work;
    }
    std::vector<int> owner((size_t)work);
    for (size_t k = 0; k < n; ++k)
        std::fill(owner.begin() + hm[k].wbase, owner.begin() + hm[k].wbase + hm[k].ntile, (int)k);
    if (work) BHIPCHK(b, h2d(b->stream, b->downer, owner.data(), (size_t)work * sizeof(int)));
    if (n) {
        BHIPCHK(b, h2d(b->stream, b->dmem, hm.data(), n * sizeof(swm::Member<T>)));
        BHIPCHK(b, h2d(b->stream, b->dstep, b->hstep.data(), n * sizeof(swm::MemberStep)));
    }
    b->on_device = n > 0;
    b->dirty = false;
    batch_drop_graph(b);
    return 0;
}

// Resident blocks of the persistent batch kernel: its occupancy per CU times
// the CU count, found once per instantiation.
template <typename T, int V, int NT, int TX, int TY>
int queue_grid() {
    static const int g = [] {
        int dev = 0, ncu = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, swm::batch_queue_kernel<T, V, NT, TX, TY>, 256, 0);
        return std::max(1, ncu) * std::max(1, per);
    }();
    return g;
}

template <typename T, int V, int NT, int TX, int TY>
void batch_launch_k(swm_batch* b, bool queue) {
    const swm::Member<T>* m = (const swm::Member<T>*)b->dmem;
    const int nmem = (int)b->mem.size();
    if (queue) {
        const int g = std::min(b->nwork, queue_grid<T, V, NT, TX, TY>());
        hipLaunchKernelGGL((swm::batch_queue_kernel<T, V, NT, TX, TY>), dim3(g), dim3(256), 0, b->stream, m,
                           b->dstep, b->downer, b->nwork, nmem, b->dqueue);
    } else {
        hipLaunchKernelGGL((swm::batch_kernel<T, V, NT, TX, TY>), dim3(b->nwork), dim3(256), 0, b->stream, m,
                           b->dstep, b->downer, b->nwork);
        hipLaunchKernelGGL(swm::batch_advance, dim3((nmem + 255) / 256), dim3(256), 0, b->stream, b->dstep, nmem);
    }
}

template <typename T, int V, int TX, int TY>
void batch_launch_s(swm_batch* b, bool queue) {
    switch (b->ntrc_max) {
    case 0: batch_launch_k<T, V, 0, TX, TY>(b, queue); break;
    case 1: batch_launch_k<T, V, 1, TX, TY>(b, queue); break;
    default: batch_launch_k<T, V, 2, TX, TY>(b, queue); break;
    }
}

// One batched step: the queue kernel's last block advances the members; the
// one-block-per-tile kernel is followed by batch_advance.
template <typename T, int V>
void batch_launch_v(swm_batch* b) {
    const BatchShape bs = kBatchShapes[b->shape];
    if (bs.tx == 32) batch_launch_s<T, V, 32, 8>(b, bs.queue != 0);
    else batch_launch_s<T, V, 64, 4>(b, bs.queue != 0);
}

void batch_launch(swm_batch* b) {
    if (b->fp32) {
        if (b->variant == SWM_FULLY_NONLINEAR) batch_launch_v<float, swm::VAR_FN>(b);
        else batch_launch_v<float, swm::VAR_SOR>(b);
    } else {
        if (b->variant == SWM_FULLY_NONLINEAR) batch_launch_v<double, swm::VAR_FN>(b);
        else batch_launch_v<double, swm::VAR_SOR>(b);
    }
}

int member_refresh(swm_ctx* c) {
    if (!c->batch) return 0;
    if (batch_pull(c->batch)) return fail(c, "batch: " + c->batch->err, 3);
    return 0;
}

}  // namespace

extern "C" {

int swm_gpu_batch_create(swm_batch** out) {
    HipLock lock(g_hip_mu);
    if (!out) return 1;
    *out = nullptr;
    swm_batch* b = new swm_batch();
    if (const char* e = getenv("SWM_GPU_GRAPH")) b->use_graph = atoi(e) != 0;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipMalloc((void**)&b->dqueue, (swm::kBatchQueues + 1) * sizeof(int)) != hipSuccess ||
        hipMemsetAsync(b->dqueue, 0, (swm::kBatchQueues + 1) * sizeof(int), b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        g_create_err = "batch: could not create a HIP stream/events or the tile-queue counters";
        swm_gpu_batch_destroy(b);
        return 3;
    }
    *out = b;
    return 0;
}

int swm_gpu_batch_attach(swm_batch* b, swm_ctx* c, int64_t nsteps) {
    HipLock lock(g_hip_mu);
    if (!b || !c) return 1;
    if (c->batch) return bfail(b, "context is already attached to a batch", 1);
    if (nsteps < 0) return bfail(b, "nsteps must be >= 0", 1);
    if (c->has_osc) return bfail(b, "oscillating forcing needs per-step host phases; step such a context alone", 4);
    if (c->d.variant == SWM_LINEARISED_MEAN_STATE)
        return bfail(b, "LINEARISED_MEAN_STATE contexts are stepped alone", 4);
    if (!c->streams.empty() || c->dyn[0].rec || c->dyn[1].rec || c->dyn[2].rec)
        return bfail(b, "time-interpolated forcing or DYNFROMFILE datasets need per-step host arguments; step such a "
                        "context alone", 4);
    if (!c->avgs.empty()) return bfail(b, "averaging buffers are kept by contexts stepped alone", 4);
    if (b->fp32 >= 0 && (b->fp32 != (int)c->fp32 || b->variant != c->d.variant))
        return bfail(b, "every member of a batch has the same precision and variant", 4);
    if (int rc = batch_pull(b)) return rc;
    BHIPCHK(b, hipStreamSynchronize(c->stream));
    drop_graphs(c);
    b->fp32 = c->fp32;
    b->variant = c->d.variant;
    swm::MemberStep h{};
    h.itt = c->itt;
    h.remaining = nsteps;
    h.cur = c->cur;
    h.ring = c->ring;
    for (size_t n = 0; n < c->trc.size(); ++n)
        if (c->trc[n].cbuf_valid) h.cbuf |= 1 << n;
    b->mem.push_back(c);
    b->hstep.push_back(h);
    c->batch = b;
    b->dirty = true;
    return 0;
}

int swm_gpu_batch_detach(swm_batch* b, swm_ctx* c) {
    HipLock lock(g_hip_mu);
    if (!b || !c) return 1;
    if (c->batch != b) return bfail(b, "context is not attached to this batch", 1);
    if (int rc = batch_pull(b)) return rc;
    const size_t k = (size_t)(std::find(b->mem.begin(), b->mem.end(), c) - b->mem.begin());
    b->mem.erase(b->mem.begin() + k);
    b->hstep.erase(b->hstep.begin() + k);
    c->batch = nullptr;
    b->dirty = true;
    if (b->mem.empty()) b->fp32 = -1;
    return 0;
}

int swm_gpu_batch_set_steps(swm_batch* b, swm_ctx* c, int64_t nsteps) {
    HipLock lock(g_hip_mu);
    if (!b || !c) return 1;
    if (c->batch != b) return bfail(b, "context is not attached to this batch", 1);
    if (nsteps < 0) return bfail(b, "nsteps must be >= 0", 1);
    if (int rc = batch_pull(b)) return rc;
    const size_t k = (size_t)(std::find(b->mem.begin(), b->mem.end(), c) - b->mem.begin());
    b->hstep[k].remaining = nsteps;
    b->dirty = true;
    return 0;
}

int swm_gpu_batch_remaining(swm_batch* b, const swm_ctx* c, int64_t* remaining) {
    HipLock lock(g_hip_mu);
    if (!b || !c || !remaining) return 1;
    if (c->batch != b) return bfail(b, "context is not attached to this batch", 1);
    if (int rc = batch_pull(b)) return rc;
    const size_t k = (size_t)(std::find(b->mem.begin(), b->mem.end(), c) - b->mem.begin());
    *remaining = b->hstep[k].remaining;
    return 0;
}

int swm_gpu_batch_step(swm_batch* b, int64_t nsteps, int32_t* active) {
    if (!b) return 1;
    if (nsteps < 0) return bfail(b, "nsteps must be >= 0", 1);
    if (b->dirty) {
        HipLock lock(g_hip_mu);
        if (int rc = batch_pull(b)) return rc;
        int rc = b->fp32 == 1 ? batch_build_t<float>(b) : batch_build_t<double>(b);
        if (rc) return rc;
    }
    b->last_ms = 0.0;
    if (!b->mem.empty() && nsteps > 0 && !b->use_graph) {
        BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
        for (int64_t k = 0; k < nsteps; ++k) batch_launch(b);
        BHIPCHK(b, hipGetLastError());
        BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
        BHIPCHK(b, hipEventSynchronize(b->ev1));
        float ms = 0.f;
        BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->last_ms = ms;
        b->on_device = true;
    } else if (!b->mem.empty() && nsteps > 0) {
        if (!b->graph) {
            HipLock lock(g_hip_mu);
            hipGraph_t g = nullptr;
            BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            batch_launch(b);
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (le != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return bfail(b, std::string("batch graph capture: ") + hipGetErrorString(le != hipSuccess ? le : ce), 3);
            }
            const hipError_t ie = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                b->graph = nullptr;
                return bfail(b, std::string("batch graph instantiate: ") + hipGetErrorString(ie), 3);
            }
        }
        BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
        for (int64_t k = 0; k < nsteps; ++k) BHIPCHK(b, hipGraphLaunch(b->graph, b->stream));
        BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
        BHIPCHK(b, hipEventSynchronize(b->ev1));
        float ms = 0.f;
        BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->last_ms = ms;
        b->on_device = true;
    }
    if (active) {
        HipLock lock(g_hip_mu);
        if (int rc = batch_pull(b)) return rc;
        int32_t na = 0;
        for (const auto& h : b->hstep) na += h.remaining > 0;
        *active = na;
    }
    return 0;
}

int swm_gpu_batch_set_shape(swm_batch* b, int32_t shape) {
    if (!b) return 1;
    if (shape < 0 || shape >= kNumBatchShapes) return bfail(b, "no such batch tile shape", 1);
    HipLock lock(g_hip_mu);
    if (int rc = batch_pull(b)) return rc;
    b->shape = shape;
    b->dirty = true;
    return 0;
}

int swm_gpu_batch_tune(swm_batch* b, int32_t steps, int32_t* best_shape, double* times_ms) {
    if (!b || !best_shape) return 1;
    if (steps < 1) return bfail(b, "steps must be >= 1", 1);
    if (b->mem.empty()) return bfail(b, "no members to tune on", 1);
    std::vector<void*> snaps;
    std::vector<size_t> sizes;
    std::vector<swm::MemberStep> h0;
    {
        HipLock lock(g_hip_mu);
        if (int rc = batch_pull(b)) return rc;
        h0 = b->hstep;
        for (swm_ctx* c : b->mem) {
            const size_t bytes = (size_t)(swm::PL_TRC + swm::PL_PER_TRC * c->trc.size()) * c->plane * c->esz;
            void* p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) {
                for (void* q : snaps) (void)hipFree(q);
                return bfail(b, "out of device memory for the tuning snapshot", 3);
            }
            BHIPCHK(b, hipMemcpyAsync(p, c->state, bytes, hipMemcpyDeviceToDevice, b->stream));
            snaps.push_back(p);
            sizes.push_back(bytes);
        }
    }
    const int shape0 = b->shape;
    const bool graph0 = b->use_graph;
    double best = 1e300;
    int best_k = shape0, rc = 0;
    b->use_graph = false;
    auto reset = [&](int k) {
        HipLock lock(g_hip_mu);
        b->shape = k;
        b->hstep = h0;
        for (auto& h : b->hstep) h.remaining = h.remaining > 0 ? steps : 0;
        b->on_device = false;
        b->dirty = true;
    };
    auto restore = [&]() -> int {
        HipLock lock(g_hip_mu);
        for (size_t q = 0; q < snaps.size(); ++q)
            if (hipMemcpyAsync(b->mem[q]->state, snaps[q], sizes[q], hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
                return 3;
        return 0;
    };
    for (int k = 0; k < kNumBatchShapes && !rc; ++k) {
        // untimed run first (first-launch costs), then the timed run from the same state
        for (int pass = 0; pass < 2 && !rc; ++pass) {
            reset(k);
            if ((rc = swm_gpu_batch_step(b, steps, nullptr))) break;
            rc = restore();
        }
        if (rc) break;
        if (times_ms) times_ms[k] = b->last_ms / steps;
        if (b->last_ms < best) { best = b->last_ms; best_k = k; }
    }
    HipLock lock(g_hip_mu);
    (void)hipStreamSynchronize(b->stream);
    for (void* q : snaps) (void)hipFree(q);
    b->hstep = h0;
    b->on_device = false;
    for (size_t q = 0; q < b->mem.size(); ++q) {
        swm_ctx* c = b->mem[q];
        c->itt = h0[q].itt;
        c->cur = h0[q].cur;
        c->ring = h0[q].ring;
        for (size_t n = 0; n < c->trc.size(); ++n) c->trc[n].cbuf_valid = ((h0[q].cbuf >> n) & 1) != 0;
    }
    b->use_graph = graph0;
    b->shape = rc ? shape0 : best_k;
    b->dirty = true;
    if (rc) return rc;
    *best_shape = best_k;
    return 0;
}

int32_t swm_gpu_batch_shape_count(void) { return kNumBatchShapes; }

int swm_gpu_batch_last_timing(const swm_batch* b, double* total_ms) {
    if (!b || !total_ms) return 1;
    *total_ms = b->last_ms;
    return 0;
}

const char* swm_gpu_batch_last_error(const swm_batch* b) { return b ? b->err.c_str() : g_create_err.c_str(); }

void swm_gpu_batch_destroy(swm_batch* b) {
    HipLock lock(g_hip_mu);
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    (void)batch_pull(b);
    for (swm_ctx* c : b->mem) c->batch = nullptr;
    batch_drop_graph(b);
    if (b->dmem) (void)hipFree(b->dmem);
    if (b->dstep) (void)hipFree(b->dstep);
    if (b->downer) (void)hipFree(b->downer);
    if (b->dqueue) (void)hipFree(b->dqueue);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
